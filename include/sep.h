/*
 * sep.h -- C ABI of the MI355X (gfx950) eigensolver library libsep_amd.so.
 *
 * Drop-in for the hot path of SymmetricEigenProblem.jl: the QR algorithm for real
 * symmetric tridiagonal matrices with Wilkinson shifts, and the bulk application of
 * Givens rotations in wave order.  Every entry point takes plain pointers and 64-bit
 * sizes (Julia's Int), matrices are column-major with an explicit leading dimension,
 * and everything is mutated in place, as in the reference.
 *
 * Return codes: SEP_OK (0); SEP_NOT_CONVERGED (1) where the reference returns
 * converged = false; negative values for invalid arguments (the reference's
 * @assert) and HIP errors.  sep_last_error() describes the last failure of the
 * calling thread.
 *
 * Pointer conventions: functions without a suffix take HOST pointers and copy
 * through the handle's device buffers; *_dev functions take DEVICE pointers and run
 * on the handle's stream (they return once the work is queued unless stated).
 */
#ifndef SEP_H
#define SEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEP_OK 0
#define SEP_NOT_CONVERGED 1
#define SEP_ERR_ARG -1        /* invalid size / pointer / leading dimension */
#define SEP_ERR_ODD_N -2      /* odd order: the reference's @assert iseven(n) */
#define SEP_ERR_BAD_LAYERS -3 /* max_layers < 1 (or beyond the library limit) */
#define SEP_ERR_HIP -4        /* a HIP runtime call failed */
#define SEP_ERR_NOMEM -5      /* device allocation failed */
#define SEP_ERR_UNSUPPORTED -6

typedef struct sep_handle sep_handle;

/* Library handle: owns the device buffers, two HIP streams (chase / apply) and the
 * batched solver's autotuned choice.  One handle per host thread. */
int sep_create(sep_handle **h, int device);
int sep_destroy(sep_handle *h);
/* Message of the last failing call on this thread ("" if none). */
const char *sep_last_error(void);
/* Library / kernel identification string, e.g. "libsep_amd gfx950 ...". */
const char *sep_version(void);
/* Run subsequent *_dev work on `stream` (a hipStream_t; NULL = the handle's own).  The
 * handle's own streams are blocking streams, ordered with the legacy default stream; a
 * caller that produces its device arrays on another stream passes that stream here. */
int sep_set_stream(sep_handle *h, void *stream);
int sep_synchronize(sep_handle *h);

/* Kernel timing (instrumentation, no counterpart in the reference).  When enabled, every
 * launch of the main bulk-apply kernel is bracketed by HIP events on the stream it runs
 * on; sep_get_timing returns the summed kernel time (ms), the summed time of the
 * coefficient-preparation kernels that precede it (ms) and the number of apply launches
 * since the last reset.  Enabling or disabling resets the counters. */
int sep_set_timing(sep_handle *h, int enable);
int sep_get_timing(sep_handle *h, double *apply_ms, double *prep_ms, int64_t *launches);

/* ---- Seam 2: bulk Givens application in wave order --------------------------
 * Replaces bulk_wave_order_2x2_rmul!(A, givens) and its _parallel! variant
 * (src/wave.jl:62 and :36; called from src/bulge_chasing.jl:282 and
 * benchmark/benchmark.jl:13,27).
 * A is m x n column-major (leading dimension lda >= m).  cs is the (n-1) x k grid of
 * (c,s) pairs, column-major and interleaved, i.e. the memory of a Julia
 * Matrix{Tuple{T,T}}: c of rotation (i, L) at cs[2*(i + (n-1)*L)], s right after
 * (0-based i, L).  The result is A * prod_{L} prod_{i} G(i,L)^T in layer-major order,
 * where G(i,L)^T maps columns (i, i+1): a_i' = c a_i + s a_{i+1},
 * a_{i+1}' = -s a_i + c a_{i+1} (Rotation2 rmul!, src/simple_givens.jl:53-64).
 * Unlike the reference there is no parity precondition on n or k. */
int sep_bulk_rmul_f64(sep_handle *h, int64_t m, int64_t n, double *A, int64_t lda,
                      const double *cs, int64_t k);
int sep_bulk_rmul_f32(sep_handle *h, int64_t m, int64_t n, float *A, int64_t lda,
                      const float *cs, int64_t k);
int sep_bulk_rmul_f64_dev(sep_handle *h, int64_t m, int64_t n, double *A, int64_t lda,
                          const double *cs, int64_t k);
int sep_bulk_rmul_f32_dev(sep_handle *h, int64_t m, int64_t n, float *A, int64_t lda,
                          const float *cs, int64_t k);
/* Allocation-free variant of sep_bulk_rmul_f64_dev.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels and results
 * (bit for bit) as sep_bulk_rmul_f64_dev; its coefficient table and auxiliary words live
 * in a fixed 48 MiB device array per GPU (held from the first sep_create on that GPU to
 * the end of the process), which bounds n: n <= 16384 always fits, a larger n that does
 * not fit returns SEP_ERR_ARG without launching anything.  It calls no hipMalloc/hipFree
 * and records no memset.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (two streams, two host threads) race and must be serialised by the caller. */
int sep_bulk_rmul_f64_dev_noalloc(sep_handle *h, int64_t m, int64_t n, double *A, int64_t lda,
                                  const double *cs, int64_t k);

/* ---- Seam 1: the whole solve ------------------------------------------------
 * Replaces qr_algorithm!(H::SymTridiagonal{T}, Q, max_layers=128, tol=eps(T),
 * maxiter=100n) -> (H, converged, iter)  (src/bulge_chasing.jl:146,181,287).
 * dv (n) and ev (n-1) are H.dv / H.ev; on return dv holds the eigenvalues in the
 * reference's deflation order and ev the (negligible) off-diagonal.  Q is n x n
 * column-major (ldq >= n) and is multiplied from the right by all rotations, so
 * A Q0 = Q Lambda when Q0 = I.  Pass Q = NULL for eigenvalues only.  n must be even
 * (the reference's @assert).  *converged, *iter as the reference's tuple;
 * *n_rotations (may be NULL) counts the rotations the chase created, which gives the
 * algorithmic flop count 6 n R of the eigenvector update.
 * Returns SEP_NOT_CONVERGED if iter exceeded maxiter (then, as in the reference, the
 * rotations of the unfinished block are not applied to Q). */
int sep_qr_algorithm_f64(sep_handle *h, int64_t n, double *dv, double *ev, double *Q,
                         int64_t ldq, int64_t max_layers, double tol, int64_t maxiter,
                         int32_t *converged, int64_t *iter, int64_t *n_rotations);
int sep_qr_algorithm_f32(sep_handle *h, int64_t n, float *dv, float *ev, float *Q, int64_t ldq,
                         int64_t max_layers, float tol, int64_t maxiter, int32_t *converged,
                         int64_t *iter, int64_t *n_rotations);
/* Same with device pointers; synchronises the handle before returning (the status
 * words come back to the host). */
int sep_qr_algorithm_f64_dev(sep_handle *h, int64_t n, double *dv, double *ev, double *Q,
                             int64_t ldq, int64_t max_layers, double tol, int64_t maxiter,
                             int32_t *converged, int64_t *iter, int64_t *n_rotations);
int sep_qr_algorithm_f32_dev(sep_handle *h, int64_t n, float *dv, float *ev, float *Q,
                             int64_t ldq, int64_t max_layers, float tol, int64_t maxiter,
                             int32_t *converged, int64_t *iter, int64_t *n_rotations);

/* ---- Batched solve (no counterpart in the reference) --------------------------
 * B independent problems, problem p of even order n[p] (2 <= n[p] <= 8192) stored at
 * dv + off_d[p], ev + off_d[p] (n[p]-1 used) and Q + off_q[p] (n[p] x n[p], column-major,
 * ld n[p]).  Device pointers for dv/ev/Q/converged/iters (converged/iters may be NULL);
 * n/off_d/off_q are HOST arrays.  Every problem runs qr_algorithm! semantics (tol = eps,
 * maxiter = 100 n[p]): rounds of one block of sweeps per unfinished problem (one
 * wavefront each) followed by the block's application to Q, until every problem is done.
 * The rounds are captured into a hipGraph once per buffer set, layout and configuration
 * and replayed (the environment variable SEP_BATCHED_GRAPH=0 at sep_create launches them
 * one by one instead).  Q = NULL computes eigenvalues only (no application).
 * Eigenvalues, off-diagonal and iteration counts equal per-problem sep_qr_algorithm_f64's
 * bit for bit (the same chase); Q agrees to rounding (the same three-shear apply form).
 * max_layers = 0 takes the autotuned (or default, 16) block length.  converged[p]
 * (int32) and iters[p] (int64) are written per problem.  Synchronous: returns after the
 * solve.  Returns SEP_NOT_CONVERGED if any problem did not converge. */
int sep_qr_batched_f64_dev(sep_handle *h, int64_t B, const int64_t *n, const int64_t *off_d,
                           const int64_t *off_q, double *dv, double *ev, double *Q,
                           int64_t max_layers, int32_t *converged, int64_t *iters);
/* Autotune the batched solver for this batch layout: block length (max_layers), fused
 * rotation depth and rows of Q per lane, each configuration timed on scratch copies (the
 * inputs are not modified).  Writes the chosen configuration index to *choice (may be
 * NULL); later solves of the same layout on this handle use it. */
int sep_tune_batched_f64(sep_handle *h, int64_t B, const int64_t *n, const int64_t *off_d,
                         const int64_t *off_q, const double *dv, const double *ev,
                         int32_t *choice);
/* Force the batched apply's fused depth and rows per lane (16/1, 8/1, 8/2 or 4/1) for later
 * solves on this handle, overriding the autotuner; 0/0 restores it. */
int sep_set_batched_config(sep_handle *h, int32_t fused_depth, int32_t rows_per_lane);

/* Accept odd orders in the solves of this handle (allow = 1; 0 restores the default).
 * No counterpart in the reference, which asserts iseven(n) as a "limitation for now"
 * (src/bulge_chasing.jl:155): the assertion protects only its wave kernels'
 * preconditions (src/wave.jl:65-66), which this library's apply does not have.  The
 * default keeps the reference's behaviour (SEP_ERR_ODD_N). */
int sep_set_odd_order(sep_handle *h, int32_t allow);

/* Rotation arithmetic of the Float64 solves on this handle (sep_qr_algorithm_f64* and
 * sep_qr_batched_f64_dev):
 * SEP_CHASE_EXACT (the default) computes every chase rotation as LinearAlgebra's
 * givensAlgorithm does (src/simple_givens.jl:20-23, src/bulge_chasing.jl:97), so D, E, the
 * (c,s) grid and Q equal the reference's bit for bit; SEP_CHASE_FAST computes the bulge
 * rotations from one reciprocal square root (c = f/r, s = g/r, r = sqrt(f^2+g^2) each
 * within about an ulp, c >= 0), a shorter dependency chain per chase step, so results
 * agree with the reference to rounding (eigenvalues, residual and orthogonality at the
 * same tolerances) but not bitwise.  No counterpart in the reference.  Float32 solves
 * always use the exact rotation. */
enum { SEP_CHASE_EXACT = 0, SEP_CHASE_FAST = 1 };
int sep_set_chase_rotations(sep_handle *h, int32_t mode);

/* ---- Backward of the eigendecomposition (no counterpart in the reference) -----
 * Given eigenvalues lam (n), eigenvectors Q (n x n, column-major, ldq) of the
 * tridiagonal T(d, e) and the cotangents lam_bar (n) and Q_bar (n x n, ld ldq),
 * computes d_bar (n) and e_bar (n-1):
 *   M = diag(lam_bar) + F o (Q^T Q_bar),  F_ij = 1/(lam_j - lam_i) (0 on the diagonal)
 *   d_bar_k = (Q M Q^T)_kk,  e_bar_k = (Q M Q^T)_{k,k+1} + (Q M Q^T)_{k+1,k}.
 * Q^T Q_bar and Q M run on FP64 MFMA; the row projection is fused into the second
 * product.  All pointers are device pointers. */
int sep_eig_backward_f64_dev(sep_handle *h, int64_t n, const double *lam, const double *Q,
                             int64_t ldq, const double *lam_bar, const double *Q_bar,
                             double *d_bar, double *e_bar);
/* Batched backward: problem p has order n[p], lam/lam_bar/d_bar at off_d[p], e_bar at
 * off_d[p] (n[p]-1 used), Q/Q_bar at off_q[p] with ld n[p].  n/off_d/off_q are HOST
 * arrays; the rest are device pointers. */
int sep_eig_backward_batched_f64_dev(sep_handle *h, int64_t B, const int64_t *n,
                                     const int64_t *off_d, const int64_t *off_q,
                                     const double *lam, const double *Q, const double *lam_bar,
                                     const double *Q_bar, double *d_bar, double *e_bar);

/* ---- Sorted output (the convenience beside eigen_decomp!) ------------------------
 * qr_algorithm! leaves the eigenvalues in H.dv in deflation order (src/bulge_chasing.jl:
 * 146-288) and eigen_decomp! returns them so (:290-299).  This puts them in ascending
 * order: lam_out[r] = lam[perm[r]] and column r of Q_out = column perm[r] of Q, ties
 * broken by index (NaNs last, negative NaNs first), so A Q_out = Q_out diag(lam_out)
 * holds whenever A Q = Q diag(lam).  Out of place (Q_out must not overlap Q); Q = NULL
 * sorts the eigenvalues only; perm (n int64, 0-based) may be NULL.  Device pointers, on
 * the handle's stream. */
int sep_sort_eigen_f64_dev(sep_handle *h, int64_t n, const double *lam, const double *Q,
                           int64_t ldq, double *lam_out, double *Q_out, int64_t ldo,
                           int64_t *perm);
int sep_sort_eigen_f32_dev(sep_handle *h, int64_t n, const float *lam, const float *Q,
                           int64_t ldq, float *lam_out, float *Q_out, int64_t ldo, int64_t *perm);

/* ---- Dense symmetric -> tridiagonal (no counterpart in the reference) ----------------
 * The reference's qr_algorithm! starts from a SymTridiagonal (README.md:3); these two
 * entries put the reduction in front of it for a dense symmetric A, with LAPACK's
 * conventions (dsytrd / dorgtr, uplo = 'L'), so that
 *     sep_sytrd_f64_dev -> sep_orgtr_f64_dev -> sep_qr_algorithm_f64_dev(d, e, Q)
 * leaves the eigenvalues in d and the eigenvectors of the dense A in Q.
 * sytrd: only the lower triangle of A (n x n, column-major, ld lda) is referenced; on
 * return d (n) and e (n-1) hold T, tau (n-1) and the strictly lower part of A below the
 * subdiagonal hold the reflectors H(i) = I - tau_i v_i v_i^T (v_i[i+1] = 1), and A[i+1, i]
 * = e[i].  orgtr: Q (n x n, ld ldq) <- H(0) H(1) ... H(n-2) from that A and tau.
 * All pointers are device pointers on the handle's stream; n <= 65536. */
int sep_sytrd_f64_dev(sep_handle *h, int64_t n, double *A, int64_t lda, double *d, double *e,
                      double *tau);
int sep_orgtr_f64_dev(sep_handle *h, int64_t n, const double *A, int64_t lda, const double *tau,
                      double *Q, int64_t ldq);

#ifdef __cplusplus
}
#endif

#endif /* SEP_H */

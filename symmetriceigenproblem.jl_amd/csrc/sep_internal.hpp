// sep_internal.hpp -- shared declarations of libsep_amd (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/sep.h"

namespace sep {

// Thread-local message behind sep_last_error().
void set_error(const char* fmt, ...);

#define SEP_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::sep::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,               \
                             hipGetErrorString(e_));                                    \
            return SEP_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

#define SEP_CHECK_LAUNCH()                                                              \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            ::sep::set_error("%s:%d kernel launch -> %s", __FILE__, __LINE__,           \
                             hipGetErrorString(e_));                                    \
            return SEP_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for every
// outstanding global load and store of the wave (vmcnt(0)), which would drain the
// multi-interval prefetches of the apply kernels at each interval; their waves exchange
// data through LDS only (each column of Q is read by the first wave strictly before the
// last wave writes it), so an LDS fence is all that is needed.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A growable device buffer owned by the handle.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int reserve(size_t want);  // keeps contents only if no growth
    void release();
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// Apply-kernel geometry: a pass applies KG layers in each of G waves (G*KG layers),
// S time steps per barrier interval.
struct ApplyGeom {
    int KG, G, S;
    int layers() const { return KG * G; }
};

// Per-launch workspace of the systolic apply: the relaid (c,s) sweep table.
struct ApplyWork {
    DevBuf table;
    DevBuf aux;   // column scales, pass scale factors, normalisation flag
    DevBuf sync;  // the column-scale kernel's completion words (zeroed once, left zero)
};

}  // namespace sep

struct sep_handle {
    int device = 0;
    int num_cus = 256;  // compute units of the device (hipDeviceProp_t::multiProcessorCount)
    int num_xcds = 8;   // XCDs (each with its own L2); workgroups go to them in turn
    hipStream_t stream = nullptr;        // apply stream (user-visible)
    hipStream_t chase_stream = nullptr;  // chase stream (overlap with apply)
    bool own_stream = true;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // single solves: pinned copies of the chase's state words (one per grid buffer), events
    // applied[2] and chased[2]
    int64_t* solve_pinned = nullptr;
    hipEvent_t solve_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // scratch
    sep::DevBuf host_io_A, host_io_cs, host_io_d, host_io_e;
    sep::ApplyWork apply_work[2];
    sep::DevBuf grid[2], state, counters, lt;
    sep::DevBuf batch_meta, batch_scratch;
    // backward: problem table (device + the host copy it was uploaded from), M, partials
    sep::DevBuf bwd_meta, bwd_M, bwd_part;
    // dense -> tridiagonal: reflector / update vectors and per-tile partial sums
    sep::DevBuf trd_vec, trd_part, trd_blk;
    sep::DevBuf sort_rank;  // sorted output: rank of each eigenvalue
    void* blas = nullptr;  // rocblas_handle of the blocked orthogonal-factor GEMMs (lazy)
    std::vector<char> bwd_meta_host;
    void* bwd_meta_dev = nullptr;
    // batched solver: host copy of the per-problem state words, autotuner choice
    std::vector<int64_t> batch_host_st;
    int* batch_pinned = nullptr;             // pinned finished-problem count
    // chased[2], applied[2], fork and join (used while a graph is captured), end of replay
    hipEvent_t batch_ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // captured rounds of the batched solver, reused while the key (buffers, layout and
    // configuration it was captured with) matches
    hipGraphExec_t batch_graph = nullptr;
    std::vector<int64_t> batch_graph_key;
    bool batch_eager = false;  // SEP_BATCHED_GRAPH=0: launch the rounds without a graph
    // the batched solver's own streams (apply, chase): non-blocking, so that the rounds can
    // be captured whatever stream the caller set; the solver synchronises the caller's
    // stream before the rounds and its own streams after them
    hipStream_t batch_stream[2] = {nullptr, nullptr};
    std::vector<int64_t> tune_key;
    int tune_choice = -1;
    int force_kf = 0, force_r = 0;  // sep_set_batched_config (0: autotuned / default)
    int apply_geom = 0;  // apply geometry forced by SEP_APPLY (0: chosen by size)
    bool allow_odd = false;  // sep_set_odd_order: solves accept odd n
    bool chase_fast = false;  // sep_set_chase_rotations: Float64 solves chase with the
                              // fast (rounding-level) rotation
    // kernel timing (sep_set_timing): event pairs recorded around apply / prep kernels
    bool timing = false;
    std::vector<hipEvent_t> tev;            // pool, reused
    std::vector<std::pair<int, int>> tapply, tprep;  // indices into tev
    int tev_used = 0;
};

namespace sep {
// Record a timing event on `st` (returns its index, or -1 when timing is off).
int timing_mark(sep_handle* h, hipStream_t st);
inline void timing_pair(sep_handle* h, bool apply, int a, int b)
{
    if (a >= 0 && b >= 0) (apply ? h->tapply : h->tprep).emplace_back(a, b);
}
}  // namespace sep

namespace sep {

// apply_dpp.hip (T = double or float).  `rotations` promises c^2 + s^2 = 1 for every
// pair (skips the normalisation check).
template <typename T>
int bulk_apply_dev(sep_handle* h, hipStream_t st, ApplyWork& w, int64_t m, int64_t n, T* A,
                   int64_t lda, const T* cs, int64_t ld_cs, int64_t k, bool rotations = false);
// The allocation-free Float64 apply: the same kernels with their table and auxiliary words in
// a fixed device array of kNoAllocScratchBytes per GPU (48 MiB: n <= 16384 for every
// geometry; the largest need, the 32-row geometry's table and the column scales, is
// 42.7 MB there).  scratch_init records the array's address on `device` (the current
// device).
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;
int scratch_init(int device);
int bulk_apply_noalloc_f64(sep_handle* h, int64_t m, int64_t n, double* A, int64_t lda,
                           const double* cs, int64_t k);

// chase.hip
template <typename T>
int qr_algorithm_dev(sep_handle* h, int64_t n, T* dv, T* ev, T* Q, int64_t ldq,
                     int64_t max_layers, T tol, int64_t maxiter, int32_t* converged,
                     int64_t* iter, int64_t* n_rot);

// batched.hip
int qr_batched_f64(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                   const int64_t* off_q, double* dv, double* ev, double* Q, int64_t max_layers,
                   int32_t* converged, int64_t* iters);
int tune_batched_f64(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                     const int64_t* off_q, const double* dv, const double* ev, int32_t* choice);
int set_batched_config(sep_handle* h, int fused_depth, int rows_per_lane);

// backward.hip
int eig_backward_f64(sep_handle* h, int64_t n, const double* lam, const double* Q, int64_t ldq,
                     const double* lam_bar, const double* Q_bar, double* d_bar, double* e_bar);
int eig_backward_batched_f64(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                             const int64_t* off_q, const double* lam, const double* Q,
                             const double* lam_bar, const double* Q_bar, double* d_bar,
                             double* e_bar);

// sort.hip
template <typename T>
int sort_eigen_dev(sep_handle* h, int64_t n, const T* lam, const T* Q, int64_t ldq, T* lam_out,
                   T* Q_out, int64_t ldo, int64_t* perm);

// tridiag.hip
int sytrd_f64(sep_handle* h, int64_t n, double* A, int64_t lda, double* d, double* e,
              double* tau);
int orgtr_f64(sep_handle* h, int64_t n, const double* A, int64_t lda, const double* tau,
              double* Q, int64_t ldq);
void release_blas(sep_handle* h);

}  // namespace sep

/*
 * sep_oracle.c -- CPU restatement of SymmetricEigenProblem.jl's eigensolver hot path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker or as
 * the timed CPU baseline; the HIP library never calls it.
 *
 * What it restates (all paths relative to the reference repository):
 *   orc_givens_*            LinearAlgebra.givensAlgorithm via src/simple_givens.jl:20-23
 *   orc_single_shift_*      src/bulge_chasing.jl:11-144
 *   orc_wilkinson_shift_*   src/bulge_chasing.jl:239-265
 *   orc_qr_algorithm_*      src/bulge_chasing.jl:146-288
 *   orc_chase_block_*       src/bulge_chasing.jl:166-284 (one outer iteration, no apply)
 *   orc_bulk_rmul_seq_*     layer-major Rotation2 rmul!, src/simple_givens.jl:53-64
 *   orc_bulk_rmul_seq_fma_* the same with the rotation as two FMAs (src/bulk_givens.jl:44-57)
 *   orc_bulk_wave_*         src/wave.jl:62-162 with src/bulk_givens.jl:4-279
 *   orc_bulk_wave_parallel_* src/wave.jl:36-60
 *
 * Pinning (what checks each function beyond retyping the formula) lives in
 * tests/test_oracle.py: DLARTG identities and sign rule, LAPACK eigenvalues,
 * Toeplitz closed forms, the 2x2 closed form, a dense shifted QR step, bitwise
 * layer-major == wave order, residual and orthogonality bounds.
 *
 * Parity unpinned against the Julia reference itself (it cannot run here and holds no
 * golden vectors): the exact iteration count, the deflation order of the eigenvalues,
 * the bitwise grid contents and the rounding under @fastmath contraction are what this
 * restatement reads in the reference's source, not values checked against it.
 *
 * Build: oracle/Makefile (two variants of the same source: -ffp-contract=off for
 * parity, and FMA contraction for the timed CPU baseline, like the reference's
 * @fastmath SIMD kernels).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_PREFIX_FAST
#define PFX(x) x
#else
#define PFX(x) fast_##x
#endif

/* ---- Float64 ---- */
#define T double
#define SUF _f64
#define FABS fabs
#define SQRT sqrt
#define FMA fma
#define SAFMN2 0x1p-485
#define VEC_W 8
#define orc_givens PFX(orc_givens)
#define orc_single_shift PFX(orc_single_shift)
#define orc_wilkinson_shift PFX(orc_wilkinson_shift)
#define orc_bulk_rmul_seq PFX(orc_bulk_rmul_seq)
#define orc_bulk_rmul_seq_fma PFX(orc_bulk_rmul_seq_fma)
#define orc_apply_log PFX(orc_apply_log)
#define orc_bulk_wave PFX(orc_bulk_wave)
#define orc_bulk_wave_parallel PFX(orc_bulk_wave_parallel)
#define orc_qr_algorithm PFX(orc_qr_algorithm)
#define orc_chase_block PFX(orc_chase_block)
#include "sep_oracle_impl.h"
#undef T
#undef SUF
#undef FABS
#undef SQRT
#undef FMA
#undef SAFMN2
#undef VEC_W

/* ---- Float32 ---- */
#define T float
#define SUF _f32
#define FABS fabsf
#define SQRT sqrtf
#define FMA fmaf
#define SAFMN2 0x1p-51f
#define VEC_W 16
#include "sep_oracle_impl.h"
#undef T
#undef SUF
#undef FABS
#undef SQRT
#undef FMA
#undef SAFMN2
#undef VEC_W

/* The rotation counts the chase produces for a finished solve are what the
 * algorithmic flop count 6 * m * R uses; expose the flop formula of
 * benchmark/benchmark.jl:10-11 for the bench grid as well. */
double PFX(orc_bench_flops)(int64_t m, int64_t n, int64_t k) { return 6.0 * (double)k * (double)(n - 1) * (double)m; }

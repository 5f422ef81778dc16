// apply.hip -- bulk Givens application on gfx950 (the O(n^3) part of the solver).
//
// Computes A <- A * prod_{L=0..k-1} prod_{i=0..n-2} G(i,L)^T, the product that
// bulk_wave_order_2x2_rmul! (src/wave.jl:62-162) evaluates with the fused 4-rotation
// SIMD kernels of src/bulk_givens.jl:4-279.  Each rotation maps the column pair
// (i, i+1) of every row as a' = c a + s b, b' = c b - s a (Rotation2 rmul!,
// src/simple_givens.jl:53-64).
//
// Design (one workgroup = 64 rows of A, one row per lane):
//  * Rows are independent, so a lane streams its row through all columns.  Rotation
//    (i, L) is applied at "time" t = i + 2L: the rotations of one time step touch
//    disjoint column pairs and every dependency points to an earlier step (this is the
//    anti-diagonal wave order of src/wave.jl, one column per step instead of two).
//  * The G waves of a workgroup form a systolic pipeline over layers: wave g applies
//    the KG layers [g*KG, (g+1)*KG) of the pass to a window of 2*KG columns held in
//    VGPRs.  Each step it takes one new column from wave g-1 (wave 0: from HBM) and
//    hands the column that has left its window to wave g+1 (last wave: to HBM), so Q
//    crosses HBM once per G*KG layers.  Columns move between waves through a
//    double-buffered LDS ring, S steps per barrier interval; wave g runs S-1 steps
//    behind wave g-1 so that it consumes exactly what g-1 wrote in the previous
//    interval.
//  * (c,s) are wave-uniform: a prologue kernel relays the grid into "sweep table"
//    order (per wave, per time step, KG contiguous pairs, identity padding outside
//    the grid), which the main kernel reads with scalar loads into SGPRs.  The inner
//    loop is then 4 VALU ops per rotation and row (2 mul + 2 fma) and nothing else.
//  * Both element types now go to the DPP-fed kernels of apply_dpp.hip by default (2.9x
//    this kernel's Float32 rate at n = 16384); these kernels stay selectable with
//    SEP_APPLY=sgpr.
#include <type_traits>

#include "sep_internal.hpp"

namespace sep {
namespace {

constexpr int kLanes = 64;

// Sweep table layout for one pass: tab[g][u][l][2], u = t + S for time t in
// [-S, T_out + S), l in [0, KG).  Rotation index of (g, t, l) is
// i = t - 2 (g KG + l), layer L = L0 + g KG + l.
template <typename T, int KG, int G, int S>
__global__ void __launch_bounds__(256) k_relayout(T* __restrict__ tab, const T* __restrict__ cs,
                                                 int64_t ld_cs, int n, int k, int L0, int U)
{
    const int64_t total = (int64_t)G * U * KG;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int l = (int)(q % KG);
        const int64_t r = q / KG;
        const int u = (int)(r % U);
        const int g = (int)(r / U);
        const int lr = g * KG + l;
        const int L = L0 + lr;
        const int64_t i = (int64_t)(u - S) - 2 * (int64_t)lr;
        T c = T(1), s = T(0);
        if (L < k && i >= 0 && i <= n - 2) {
            const T* p = cs + 2 * (i + ld_cs * (int64_t)L);
            c = p[0];
            s = p[1];
        }
        tab[2 * q] = c;
        tab[2 * q + 1] = s;
    }
}

// One rotation on a register pair: x' = c x + s y, y' = c y - s x.
template <typename T>
__device__ __forceinline__ void rot(T& x, T& y, T c, T s)
{
    const T xn = c * x + s * y;
    const T yn = c * y - s * x;
    x = xn;
    y = yn;
}

// One barrier interval of wave g: S time steps t0 .. t0+S-1.  CHECKED selects the
// body with per-step column-range tests, used only for the intervals at the two ends
// of the sweep; interior intervals run the test-free body.
template <typename T, int KG, int G, int S, bool FIRST, bool LAST, bool CHECKED>
__device__ __forceinline__ void systolic_interval(T* __restrict__ Ar, int64_t lda, int n,
                                                  const T* __restrict__ tp, T (&win)[2 * KG],
                                                  T (&pf)[FIRST ? S : 1], const T* lds_in,
                                                  T* lds_out, int t0, int g)
{
    constexpr int W = 2 * KG;
    constexpr int Kp = KG * G;
    T in[S];
    if constexpr (FIRST) {
#pragma unroll
        for (int s = 0; s < S; ++s) in[s] = pf[s];
        // prefetch the next interval's columns t0+S+1 .. t0+2S
        const int c0 = t0 + S + 1;
        const T* __restrict__ p = Ar + (int64_t)c0 * lda;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if constexpr (CHECKED)
                pf[s] = (c0 + s < n) ? p[(int64_t)s * lda] : T(0);
            else
                pf[s] = p[(int64_t)s * lda];
        }
    } else {
        const int c0 = t0 - 2 * g * KG + 1;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if constexpr (CHECKED)
                in[s] = (c0 + s >= 0 && c0 + s < n) ? lds_in[s * kLanes] : T(0);
            else
                in[s] = lds_in[s * kLanes];
        }
    }
    // (c,s) of the current step live in SGPRs; the next step's are loaded one step
    // ahead (scheduling fences keep the compiler from hoisting all S steps' loads,
    // which would not fit the SGPR file).
    T cur[2 * KG], nxt[2 * KG];
#pragma unroll
    for (int q = 0; q < 2 * KG; ++q) cur[q] = tp[q];
    const int o0 = t0 - 2 * Kp + 2;  // output column of step 0 (last wave)
    T* __restrict__ po = Ar + (int64_t)o0 * lda;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        __builtin_amdgcn_sched_barrier(0);
        if (s + 1 < S) {
#pragma unroll
            for (int q = 0; q < 2 * KG; ++q) nxt[q] = tp[(s + 1) * (2 * KG) + q];
        }
        win[(s + 1) % W] = in[s];
#pragma unroll
        for (int l = 0; l < KG; ++l) {
            const int a = ((s - 2 * l) % W + W) % W;
            rot(win[a], win[(a + 1) % W], cur[2 * l], cur[2 * l + 1]);
        }
        if (s + 1 < S) {
#pragma unroll
            for (int q = 0; q < 2 * KG; ++q) cur[q] = nxt[q];
        }
        const T out = win[(s + 2) % W];
        if constexpr (LAST) {
            if constexpr (CHECKED) {
                if (o0 + s >= 0 && o0 + s < n) po[(int64_t)s * lda] = out;
            } else {
                po[(int64_t)s * lda] = out;
            }
        } else {
            lds_out[s * kLanes] = out;
        }
    }
}

template <typename T, int KG, int G, int S, bool FIRST, bool LAST>
__device__ __forceinline__ void systolic_wave(T* __restrict__ A, int64_t lda, int m, int n,
                                              const T* __restrict__ tabg, int T_out, int NI,
                                              T (*ring)[2][S][kLanes], int g, int lane)
{
    constexpr int W = 2 * KG;  // window: columns [t - 2(g+1)KG + 2, t - 2gKG + 1]
    constexpr int Kp = KG * G;
    // Lanes past the last row alias row m-1: they load, compute and store exactly the
    // values of that row's own lane, so no per-lane predication is needed.
    const int64_t row = min((int64_t)blockIdx.x * kLanes + lane, (int64_t)m - 1);
    T* __restrict__ Ar = A + row;

    T win[W];
#pragma unroll
    for (int w = 0; w < W; ++w) win[w] = T(0);

    T pf[FIRST ? S : 1];
    if constexpr (FIRST) {
        // column 0 enters before step 0 (slot 0); steps 0..S-1 take columns 1..S
        win[0] = Ar[0];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int col = s + 1;
            pf[s] = (col < n) ? Ar[(int64_t)col * lda] : T(0);
        }
    }

    for (int b = 0; b < NI; ++b) {
        const int t0 = (b - g) * S + g;
        const bool active = (t0 + S > 0) && (t0 < T_out);
        if (active) {
            const T* __restrict__ tp = tabg + (int64_t)(t0 + S) * (2 * KG);
            const T* lds_in = FIRST ? nullptr : &ring[g > 0 ? g - 1 : 0][(b - 1) & 1][0][lane];
            T* lds_out = LAST ? nullptr : &ring[g][b & 1][0][lane];
            bool interior;
            if constexpr (FIRST) {
                interior = t0 + 2 * S < n;
            } else {
                const int c0 = t0 - 2 * g * KG + 1;
                interior = c0 >= 0 && c0 + S <= n;
            }
            if constexpr (LAST) {
                const int o0 = t0 - 2 * Kp + 2;
                interior = interior && o0 >= 0 && o0 + S <= n;
            }
            if (interior)
                systolic_interval<T, KG, G, S, FIRST, LAST, false>(Ar, lda, n, tp, win, pf,
                                                                   lds_in, lds_out, t0, g);
            else
                systolic_interval<T, KG, G, S, FIRST, LAST, true>(Ar, lda, n, tp, win, pf,
                                                                  lds_in, lds_out, t0, g);
        }
        lds_barrier();
    }
}

template <typename T, int KG, int G, int S>
__global__ void __launch_bounds__(kLanes * G)
    k_apply_systolic(T* __restrict__ A, int64_t lda, int m, int n, const T* __restrict__ tab,
                     int U, int T_out, int NI)
{
    static_assert(S % (2 * KG) == 0, "interval must be a multiple of the window");
    __shared__ T ring[G > 1 ? G - 1 : 1][2][S][kLanes];
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int lane = threadIdx.x % kLanes;
    const T* __restrict__ tabg = tab + (int64_t)g * U * (2 * KG);
    if (g == 0) {
        systolic_wave<T, KG, G, S, true, G == 1>(A, lda, m, n, tabg, T_out, NI, ring, g, lane);
    } else if (g == G - 1) {
        systolic_wave<T, KG, G, S, false, true>(A, lda, m, n, tabg, T_out, NI, ring, g, lane);
    } else {
        systolic_wave<T, KG, G, S, false, false>(A, lda, m, n, tabg, T_out, NI, ring, g, lane);
    }
}

template <typename T, int KG, int G, int S>
int run_passes(sep_handle* h, hipStream_t st, ApplyWork& w, int64_t m, int64_t n, T* A,
               int64_t lda, const T* cs, int64_t ld_cs, int64_t k)
{
    constexpr int Kp = KG * G;
    const int T_out = (int)(n + 2 * Kp - 2);
    const int U = T_out + 2 * S;
    const int NI = G + (T_out + S - 1) / S + 1;
    const size_t tab_elems = (size_t)G * U * KG * 2;
    if (int rc = w.table.reserve(tab_elems * sizeof(T))) return rc;
    T* tab = w.table.as<T>();
    const int64_t rb = (m + kLanes - 1) / kLanes;
    for (int64_t L0 = 0; L0 < k; L0 += Kp) {
        const int64_t total = (int64_t)G * U * KG;
        const int rblocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
        const int tp = timing_mark(h, st);
        hipLaunchKernelGGL((k_relayout<T, KG, G, S>), dim3(rblocks), dim3(256), 0, st, tab, cs,
                           ld_cs, (int)n, (int)k, (int)L0, U);
        SEP_CHECK_LAUNCH();
        const int ta = timing_mark(h, st);
        timing_pair(h, false, tp, ta);
        hipLaunchKernelGGL((k_apply_systolic<T, KG, G, S>), dim3((unsigned)rb),
                           dim3(kLanes * G), 0, st, A, lda, (int)m, (int)n, tab, U, T_out, NI);
        SEP_CHECK_LAUNCH();
        timing_pair(h, true, ta, timing_mark(h, st));
    }
    return SEP_OK;
}

}  // namespace

template <typename T>
int bulk_apply_dev(sep_handle* h, hipStream_t st, ApplyWork& w, int64_t m, int64_t n, T* A,
                   int64_t lda, const T* cs, int64_t ld_cs, int64_t k, bool rotations)
{
    (void)h;
    if (m <= 0 || n <= 1 || k <= 0) return SEP_OK;
    if (m > INT32_MAX || n > (INT32_MAX / 4)) {
        set_error("bulk apply: m=%lld n=%lld beyond the 32-bit index range",
                  (long long)m, (long long)n);
        return SEP_ERR_ARG;
    }
    if (h == nullptr || !h->force_sgpr_apply)
        return bulk_apply_dpp<T>(h, st, w, m, n, A, lda, cs, ld_cs, k, rotations);
    if (k <= 32) return run_passes<T, 8, 4, 16>(h, st, w, m, n, A, lda, cs, ld_cs, k);
    return run_passes<T, 8, 8, 16>(h, st, w, m, n, A, lda, cs, ld_cs, k);
}

template int bulk_apply_dev<double>(sep_handle*, hipStream_t, ApplyWork&, int64_t, int64_t,
                                    double*, int64_t, const double*, int64_t, int64_t, bool);
template int bulk_apply_dev<float>(sep_handle*, hipStream_t, ApplyWork&, int64_t, int64_t, float*,
                                   int64_t, const float*, int64_t, int64_t, bool);

}  // namespace sep

// batched.hip -- thousands of independent qr_algorithm! solves (Float64) per call (no
// counterpart in the reference, which solves one matrix per call).
//
// The outer loop of qr_algorithm! (src/bulge_chasing.jl:166-285) runs for all problems at
// once, in rounds.  Round r is two launches over the whole batch:
//   k_chase_all:  one wavefront per problem runs one block of sweeps with the single
//                 solve's device chase (chase.hpp, src/bulge_chasing.jl:166-276), into the
//                 reference's (n-1) x max_layers grid layout; a problem that has finished
//                 (no sweep in its last block, or maxiter exceeded) returns at once.  It
//                 leaves the depth to apply (0 when nothing is to be applied) and the
//                 column range the block touched, per problem;
//   k_apply_all:  the block's rotations are applied to each problem's Q in wave order
//                 t = i + 2L (src/wave.jl:62-162), KF layers per pass over Q, R rows of Q
//                 per lane with the window of 2 KF columns in VGPRs, Q columns prefetched
//                 8 steps ahead, and the (c, s) of 16 time steps staged in LDS (double
//                 buffered, loaded one stage ahead).  Rotations beyond a column's layer
//                 count read as the identity, so the grid never needs an identity fill.
//                 Each rotation (Rotation2 rmul!, src/simple_givens.jl:53-64) is applied
//                 as three shears with DPP-broadcast coefficients, as in the bulk apply;
//                 Q agrees with the layer-major product to rounding (eigenvalues, the
//                 off-diagonal and the iteration counts come from the chase alone and
//                 equal the single solve's bit for bit).
// The grid, layer counts and depths are double buffered by round parity, so the apply of
// round r (apply stream) overlaps the chase of round r+1 (chase stream).  The host reads a
// device count of finished problems after each chase; every round either finishes a
// problem or advances its iteration count, so the number of rounds is at most
// max(maxiter) + 2.  The autotuner picks max_layers, the fused depth KF and the rows per
// lane R.
#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <type_traits>

#include "chase.hpp"
#include "sep_internal.hpp"
#include "shear.hpp"

namespace sep {
namespace {

constexpr int kLanes = 64;
constexpr int kStageSteps = 16;  // wave-order time steps staged per LDS fill
constexpr int kQAhead = 8;       // Q columns prefetched ahead
constexpr int kMaxBatchedN = 8192;
constexpr int kGraphRounds = 4;  // rounds per captured graph (even: buffer parity repeats)
// The batched chase shares each SIMD with other chase waves and the apply: it takes the
// chase's shared-SIMD form (per-step scalar range test, chunk inputs read from LDS).
constexpr bool kBatchedSolo = false;

// state word (beyond the chase's own) marking a finished problem
constexpr int ST_DONE = 10;
static_assert(ST_DONE < chase::ST_WORDS, "state words");

struct BProb {
    int n;
    int pad;
    int64_t off_d;   // dv / ev
    int64_t off_q;   // Q (n x n, ld n)
    int64_t off_lt;  // layer counts (n ints)
    int64_t off_g;   // grid ((n-1) x max_layers pairs)
};

struct BWork {
    int prob;
    int row0;
};

// The caller's arrays of one solve, read by the kernels from device memory (uploaded with
// the call's other metadata) rather than passed as kernel arguments, so that the captured
// round graph does not depend on them: callers that solve into new buffers every call
// (fresh framework tensors) replay the same graph instead of capturing and instantiating a
// new one.
struct BatchBufs {
    double* dv;
    double* ev;
    double* Q;
};

// A problem's dv, ev and layer counts in LDS take this many bytes per order (beside the
// LDS share that caps the chase waves per SIMD); problems up to kLdsChaseMaxN chase there.
constexpr unsigned kLdsBytesPerN = 2 * sizeof(double) + sizeof(int);
constexpr int kLdsChaseMaxN = 3072;

// One block of sweeps for every unfinished problem (one wavefront each).  lt / grid /
// depth are this round's buffers.  With in_lds the wave first copies its problem's dv and
// ev into dynamic LDS (nmax entries each, then nmax layer counts) and runs the block there:
// the deflation scans, thickness maxima and the sweeps' loads and stores then never wait
// on a global memory that the concurrent apply keeps busy.  The block's results are
// copied back (dv, ev and the layer counts the apply reads); the grid goes straight to
// global memory.
template <bool kFast>
__global__ void __launch_bounds__(kLanes)
    k_chase_all(const BProb* __restrict__ probs, const int* __restrict__ order,
                const BatchBufs* __restrict__ bufs, int* lt, double* grid, int4* info,
                int64_t* st_all, int ML, int nmax, int* done_count, int in_lds)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    double* const dv = bufs->dv;
    double* const ev = bufs->ev;
    // blocks are dispatched in index order as wave slots free up; block b takes the b-th
    // largest problem, so the longest blocks start first (longest-first list scheduling)
    const int p = order[blockIdx.x];
    const int lane = threadIdx.x;
    int64_t* st = st_all + (int64_t)p * chase::ST_WORDS;
    if (__builtin_amdgcn_readfirstlane((int)st[ST_DONE])) {
        if (lane == 0) info[p] = make_int4(0, 0, 0, 0);
        return;
    }
    const BProb pr = probs[p];
    // The serial chase is the critical path, and a round lasts as long as its largest
    // problems' blocks: chase waves issue ahead of the concurrent apply, larger problems
    // ahead of smaller ones on a shared SIMD.
    const int pri = __builtin_amdgcn_readfirstlane((4 * pr.n) / (nmax + 1));
    if (pri >= 3)
        __builtin_amdgcn_s_setprio(3);
    else if (pri == 2)
        __builtin_amdgcn_s_setprio(2);
    else
        __builtin_amdgcn_s_setprio(1);
    chase::Problem<double> P{dv + pr.off_d, ev + pr.off_d, lt + pr.off_lt, grid + 2 * pr.off_g,
                             pr.n, ML, (int64_t)(pr.n - 1)};
    const double tol = 2.220446049250313e-16;  // eps(Float64)
    if (in_lds) {
        double* ld = reinterpret_cast<double*>(dyn_lds);
        double* le = ld + nmax;
        int* ll = reinterpret_cast<int*>(le + nmax);
        for (int q = lane; q < pr.n; q += kLanes) {
            ld[q] = P.dv[q];
            if (q + 1 < pr.n) le[q] = P.ev[q];
        }
        __syncthreads();
        chase::Problem<double> PL{ld, le, ll, P.grid, pr.n, ML, (int64_t)(pr.n - 1)};
        chase::chase_block<double, kBatchedSolo, kFast>(PL, tol, 100 * (int64_t)pr.n, st, lane);
        __syncthreads();
        for (int q = lane; q < pr.n; q += kLanes) {
            P.dv[q] = ld[q];
            if (q + 1 < pr.n) P.ev[q] = le[q];
            P.lt[q] = ll[q];
        }
    } else {
        chase::chase_block<double, kBatchedSolo, kFast>(P, tol, 100 * (int64_t)pr.n, st, lane);
    }
    __threadfence_block();
    if (lane == 0) {
        // maxiter exceeded: the reference returns without applying this block
        // (src/bulge_chasing.jl:180-182); no sweep: nothing to apply and done (:279)
        const bool early = st[chase::ST_EARLY] != 0, did = st[chase::ST_DID] != 0;
        // apply depth and the column range the block touched: rotation indices
        // [ilo, ihi] (1-based) act on columns ilo - 1 .. ihi (0-based)
        const int ilo = (int)st[chase::ST_ILO], ihi = (int)st[chase::ST_IHI];
        info[p] = (did && !early)
                      ? make_int4((int)st[chase::ST_MAXLAYER], ilo - 1, ihi - ilo + 2, 0)
                      : make_int4(0, 0, 0, 0);
        if (early || !did) {
            st[ST_DONE] = 1;
            atomicAdd(done_count, 1);
        }
    }
}

template <int N, int i0>
struct Unroll {
    template <class F>
    __device__ __forceinline__ static void run(F&& f)
    {
        if constexpr (i0 < N) {
            f(std::integral_constant<int, i0>{});
            Unroll<N, i0 + 1>::run(f);
        }
    }
};

// Buffer resource over `bytes` bytes at p: loads beyond return 0 and stores beyond are
// dropped by the hardware range check (raw buffer, word 3 as for every gfx9 part).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, double v)
{
    using u2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, off, 0, 0);
}

// Apply this round's block to Q for rows [row0, row0 + 64 R) of one problem, over the
// columns c0 .. c0 + nc - 1 the block touched (the others are unchanged by it).  Q is
// addressed through a buffer resource spanning exactly those columns, so the wave-order
// loads of columns past the range read 0 and the stores of columns outside it (and of
// lanes past the last row) are dropped without a branch.
//
// Each rotation is three in-place FMAs, a += p b; b += q a; a += p b with p = s / (1 + c),
// q = -s (the shear form of the bulk apply, shear.hpp), fed by DPP broadcasts from one
// coefficient register per 8 layers and step.  The chase's rotations come from DLARTG,
// whose sign rule leaves c > 0 whenever |f| > |g| and |c| <= 1/sqrt 2 otherwise, so
// 1 + c >= 1 - 1/sqrt 2 and |p| <= 1 + sqrt 2: no column signs need tracking.  Slots past
// a column's layer count, outside the block or past its depth hold p = q = 0.
template <int KF, int R>
__device__ __forceinline__ void apply_item(const BWork wk, const BProb* __restrict__ probs,
                                           const int* __restrict__ lt,
                                           const double* __restrict__ grid,
                                           const int4* __restrict__ info, double* __restrict__ Q)
{
    constexpr int W = 2 * KF;
    // a stage is a whole number of window turns (the window slots are compile-time)
    constexpr int SB = W > kStageSteps ? W : kStageSteps;
    constexpr int PA = kQAhead;
    constexpr int NE = SB * KF / kLanes;  // staged entries per lane
    constexpr int NCR = KF >= 8 ? KF / 8 : 1;  // coefficient registers per step
    static_assert(SB % W == 0 && SB % PA == 0 && NE * kLanes == SB * KF, "stage geometry");
    static_assert(KF == 4 || KF % 8 == 0, "KF is 4 or a multiple of 8");
    __shared__ double stc[2][SB][NCR][16];

    const int4 inf = info[wk.prob];
    const int depth = __builtin_amdgcn_readfirstlane(inf.x);
    if (depth == 0) return;  // nothing to apply for this item
    const int c0 = __builtin_amdgcn_readfirstlane(inf.y);
    const int nc = __builtin_amdgcn_readfirstlane(inf.z);
    const BProb pr = probs[wk.prob];
    const int n = pr.n;
    const int lane = threadIdx.x;
    const int* __restrict__ ltp = lt + pr.off_lt;
    const double* __restrict__ g = grid + 2 * pr.off_g;
    const __amdgpu_buffer_rsrc_t qr =
        make_rsrc(Q + pr.off_q + (int64_t)c0 * n, (uint32_t)n * (uint32_t)nc * 8u);
    const uint32_t colb = (uint32_t)n * 8u;  // bytes per column
    // byte offset of this lane's row; lanes past the last row get an offset beyond any
    // range (2^31 > 8 n nc + every column offset used), so they load 0 and store nothing
    uint32_t rowb[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = wk.row0 + r * kLanes + lane;
        rowb[r] = row < n ? 8u * (uint32_t)row : 0x80000000u;
    }
    // staged entry e of this lane: step s = q / KF, layer l = q % KF with q = lane + 64 e
    double2 sv[NE];
    int slt[NE];
    bool sin_[NE];
    auto stage_load = [&](int t0, int L0) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int q = lane + kLanes * e;
            const int s = q / KF, l = q % KF;
            const int il = t0 + s - 2 * l, L = L0 + l;
            sin_[e] = il >= 0 && il <= nc - 2 && L < depth;
            const int ic = c0 + min(max(il, 0), nc - 2), Lc = min(L, depth - 1);
            slt[e] = ltp[ic];
            sv[e] = *reinterpret_cast<const double2*>(g + 2 * (ic + (int64_t)(n - 1) * Lc));
        }
    };
    auto stage_store = [&](int buf, int L0) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int q = lane + kLanes * e;
            const int s = q / KF, l = q % KF;
            const bool on = sin_[e] && L0 + l < slt[e];
            // p = s / (1 + c) from the hardware reciprocal, one Newton step and one
            // remainder correction (a rounding or two from the IEEE quotient: 1 + c lies in
            // [1 - 1/sqrt 2, 2] for every rotation the chase produces); stale entries past a
            // column's layer count may give anything here and are replaced by 0
            const double den = 1.0 + sv[e].x;
            double rc = __builtin_amdgcn_rcp(den);
            rc = __builtin_fma(rc, __builtin_fma(-den, rc, 1.0), rc);
            const double q0 = sv[e].y * rc;
            const double p = on ? __builtin_fma(__builtin_fma(-den, q0, sv[e].y), rc, q0) : 0.0;
            const double qq = on ? -sv[e].y : 0.0;
            stc[buf][s][l / 8][2 * (l % 8)] = p;
            stc[buf][s][l / 8][2 * (l % 8) + 1] = qq;
        }
    };

    // time steps of one pass (the last ones only store), rounded up to whole stages: steps
    // past the end meet identity rotations and columns beyond the range
    const int nblk = (nc + 2 * KF - 2 + SB - 1) / SB;
    for (int L0 = 0; L0 < depth; L0 += KF) {
        double win[R][W];
        double qn[R][PA];
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int w = 0; w < W; ++w) win[r][w] = 0.0;
            win[r][0] = buf_load(qr, rowb[r]);
#pragma unroll
            for (int k = 0; k < PA; ++k) qn[r][k] = buf_load(qr, rowb[r] + (uint32_t)(k + 1) * colb);
        }
        __syncthreads();  // the previous pass has read both stage buffers
        stage_load(0, L0);
        for (int b = 0; b < nblk; ++b) {
            const int buf = b & 1;
            const int t0 = b * SB;
            stage_store(buf, L0);
            __syncthreads();
            if (b + 1 < nblk) stage_load(t0 + SB, L0);  // in flight during this stage
            // byte offsets of the columns loaded / stored at step 0 of this stage (a
            // negative column wraps to an offset beyond the range), advanced by one column
            // per step: the asm keeps them running vector adds, so that the compiler does not
            // fold them into per-step scalar constants (which spill at this register count)
            uint32_t lo[R], so[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lo[r] = rowb[r] + (uint32_t)(t0 + 1 + PA) * colb;
                so[r] = rowb[r] + (uint32_t)(t0 - 2 * KF + 2) * colb;
            }
            // the coefficients of step s are read from LDS one step ahead
            double cfn[NCR];
#pragma unroll
            for (int h = 0; h < NCR; ++h) cfn[h] = stc[buf][0][h][lane & 15];
            Unroll<SB, 0>::run([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                constexpr int k = s % PA;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    win[r][(s + 1) % W] = qn[r][k];  // column t + 1 enters the window
                    qn[r][k] = buf_load(qr, lo[r]);
                    lo[r] += colb;
                    asm volatile("" : "+v"(lo[r]));
                }
                double cf[NCR];
#pragma unroll
                for (int h = 0; h < NCR; ++h) {
                    cf[h] = cfn[h];
                    if constexpr (s + 1 < SB) cfn[h] = stc[buf][s + 1][h][lane & 15];
                }
                // the next step's reads stay ahead of this step's FMAs
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (KF == 4) {
                        step_shear<double, 4, s, 0, W>(win[r], cf[0]);
                    } else {
                        Unroll<NCR, 0>::run([&](auto hc) {
                            constexpr int h = decltype(hc)::value;
                            // layers 8h .. 8h + 7: slots s - 2 (8h + l)
                            step_shear<double, 8, s - 16 * h, 0, W>(win[r], cf[h]);
                        });
                    }
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {  // column t - 2 KF + 2 leaves the window
                    buf_store(qr, so[r], win[r][(s + 2) % W]);
                    so[r] += colb;
                    asm volatile("" : "+v"(so[r]));
                }
                __builtin_amdgcn_sched_barrier(0);
            });
        }
    }
}

// The block's application for every work item, largest problems first.  The grid may be
// smaller than the work list (each wave then takes items blockIdx.x, + gridDim.x, ...):
// a capped grid leaves wave slots free for the next round's chase, which runs beside it.
template <int KF, int R>
__global__ void __launch_bounds__(kLanes)
    k_apply_all(const BWork* __restrict__ work, int nwork, const BProb* __restrict__ probs,
                const int* __restrict__ lt, const double* __restrict__ grid,
                const int4* __restrict__ info, const BatchBufs* __restrict__ bufs)
{
    double* const Q = bufs->Q;
    for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const BWork wk = work[wi];
        if (wk.prob >= 0) apply_item<KF, R>(wk, probs, lt, grid, info, Q);  // -1: padding
    }
}

// max_layers, fused depth, rows per lane, and chase waves per SIMD (the chase kernel
// reserves LDS so that at most CW of its waves share a SIMD: a chase wave's serial
// recurrence slows down in proportion to the chase waves beside it)
struct Config {
    int ML, KF, R, CW;
    int AW = 0;  // apply waves per SIMD (0: one wave per work item, uncapped)
};
constexpr Config kConfigs[] = {
    // measured on the 4096-problem workload (n in [32, 512]): 16 fused layers (one pass
    // over Q per block of 16) beat 8 (81-95 ms) and 4 (122-127 ms); one chase wave per
    // SIMD (CW 1) costs 103-108 ms; capping the apply grid (AW) does not help
    {16, 16, 1, 2}, {16, 16, 1, 4}, {32, 16, 1, 2}, {32, 16, 1, 4}, {24, 16, 1, 4},
    {16, 16, 1, 4, 1}, {16, 8, 1, 4},  {16, 8, 1, 2},  {8, 8, 1, 2},   {32, 8, 2, 4},
    {32, 8, 1, 2},  {64, 8, 2, 2},
};
constexpr int kLdsPerCU = 160 * 1024;
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);

struct Layout {
    std::vector<BProb> probs;
    std::vector<int> order;  // problems by decreasing order n (stable)
    std::vector<BWork> work;
    int nmax = 0;
    int64_t tot_lt = 0, tot_g = 0;
};

int check_batch(int64_t B, const int64_t* n, const int64_t* off_d, const int64_t* off_q,
                bool allow_odd)
{
    if (B < 0 || (B > 0 && (!n || !off_d || !off_q))) {
        set_error("batched: invalid batch description");
        return SEP_ERR_ARG;
    }
    if (B > (int64_t)1 << 30) {
        set_error("batched: too many problems (%lld)", (long long)B);
        return SEP_ERR_ARG;
    }
    for (int64_t p = 0; p < B; ++p) {
        if (n[p] < 2 || n[p] > kMaxBatchedN || off_d[p] < 0 || off_q[p] < 0) {
            set_error("batched: problem %lld has order %lld (2 <= n <= %d) or bad offsets",
                      (long long)p, (long long)n[p], kMaxBatchedN);
            return SEP_ERR_ARG;
        }
        if (n[p] % 2 && !allow_odd) {
            set_error("batched: problem %lld has odd order %lld (the reference asserts iseven(n))",
                      (long long)p, (long long)n[p]);
            return SEP_ERR_ODD_N;
        }
    }
    return SEP_OK;
}

// Work items by XCD: consecutive workgroups of a launch go to the device's XCDs in turn
// (workgroup b to XCD b mod X), and each XCD has its own L2.  All row blocks of one problem
// read the same staged (c, s) grid, so they are placed on one XCD (positions x, x + X,
// x + 2X, ... of the list), where the grid is fetched from HBM once and then hits in L2.
// Problems go, largest first, to the XCD with the least work so far (rows x columns); the
// shorter queues are padded with empty items (prob = -1).
void build_layout(int64_t B, const int64_t* n, const int64_t* off_d, const int64_t* off_q,
                  const Config& cfg, int xcds, Layout& L)
{
    L.probs.resize(B);
    L.work.clear();
    L.nmax = 0;
    L.tot_lt = L.tot_g = 0;
    for (int64_t p = 0; p < B; ++p) {
        BProb& q = L.probs[p];
        q.n = (int)n[p];
        q.pad = 0;
        q.off_d = off_d[p];
        q.off_q = off_q[p];
        q.off_lt = L.tot_lt;
        q.off_g = L.tot_g;
        L.tot_lt += n[p];
        L.tot_g += (n[p] - 1) * cfg.ML;
        L.nmax = std::max(L.nmax, (int)n[p]);
    }
    // largest problems first: their apply is the longest
    std::vector<int>& order = L.order;
    order.resize(B);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    const int rows = kLanes * cfg.R;
    const int X = std::max(1, xcds);
    std::vector<std::vector<BWork>> q(X);
    std::vector<double> load(X, 0.0);
    for (int p : order) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        for (int r0 = 0; r0 < (int)n[p]; r0 += rows) q[x].push_back(BWork{p, r0});
        load[x] += (double)n[p] * (double)n[p];
    }
    size_t len = 0;
    for (auto& v : q) len = std::max(len, v.size());
    L.work.assign(len * X, BWork{-1, 0});
    for (int x = 0; x < X; ++x)
        for (size_t j = 0; j < q[x].size(); ++j) L.work[j * X + x] = q[x][j];
}

template <int KF, int R>
void launch_apply(int nwork, int grid_cap, hipStream_t s, const BWork* w, const BProb* p,
                  const int* lt, const double* g, const int4* depth, const BatchBufs* bufs)
{
    const int blocks = grid_cap > 0 ? std::min(nwork, grid_cap) : nwork;
    hipLaunchKernelGGL((k_apply_all<KF, R>), dim3(blocks), dim3(kLanes), 0, s, w, nwork, p, lt, g,
                       depth, bufs);
}

// Events and pinned counter of the round pipeline, created on first use.
int ensure_batch_sync(sep_handle* h)
{
    if (!h->batch_pinned)
        SEP_HIP(hipHostMalloc((void**)&h->batch_pinned, sizeof(int) * 4, hipHostMallocDefault));
    for (auto& bs : h->batch_stream)
        if (!bs) SEP_HIP(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
    for (int i = 0; i < 7; ++i)
        if (!h->batch_ev[i]) SEP_HIP(hipEventCreateWithFlags(&h->batch_ev[i], hipEventDisableTiming));
    return SEP_OK;
}

// The whole batched solve; returns after completion.  Round r uses buffer set r & 1:
// chase(r) on the chase stream waits for apply(r - 2) (same buffers), apply(r) on the
// apply stream waits for chase(r).
int run_batched(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                const int64_t* off_q, double* dv, double* ev, double* Q, const Config& cfg,
                int32_t* converged, int64_t* iters)
{
    Layout L;
    build_layout(B, n, off_d, off_q, cfg, h->num_xcds, L);
    if (int rc = ensure_batch_sync(h)) return rc;
    const size_t probs_b = sizeof(BProb) * (size_t)B;
    const size_t work_b = sizeof(BWork) * L.work.size();
    const size_t st_b = sizeof(int64_t) * chase::ST_WORDS * (size_t)B;
    const size_t depth_b = sizeof(int4) * (size_t)B;
    const size_t order_b = sizeof(int) * (size_t)B;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    if (int rc = h->batch_meta.reserve(al(probs_b) + al(work_b) + al(st_b) + 2 * al(depth_b) +
                                       256 + al(order_b)))
        return rc;
    const size_t lt_set = al(sizeof(int) * (size_t)L.tot_lt);
    const size_t g_set = al(sizeof(double) * 2 * (size_t)L.tot_g);
    if (int rc = h->batch_scratch.reserve(2 * (lt_set + g_set))) return rc;
    char* m = static_cast<char*>(h->batch_meta.p);
    BProb* d_probs = reinterpret_cast<BProb*>(m);
    BWork* d_work = reinterpret_cast<BWork*>(m + al(probs_b));
    int64_t* d_st = reinterpret_cast<int64_t*>(m + al(probs_b) + al(work_b));
    int4* d_depth[2];
    d_depth[0] = reinterpret_cast<int4*>(m + al(probs_b) + al(work_b) + al(st_b));
    d_depth[1] = reinterpret_cast<int4*>(reinterpret_cast<char*>(d_depth[0]) + al(depth_b));
    int* d_count = reinterpret_cast<int*>(reinterpret_cast<char*>(d_depth[1]) + al(depth_b));
    int* d_order = reinterpret_cast<int*>(reinterpret_cast<char*>(d_count) + 256);
    // the caller's arrays, in the count's 256-byte slot
    BatchBufs* d_bufs = reinterpret_cast<BatchBufs*>(reinterpret_cast<char*>(d_count) + 64);
    const BatchBufs host_bufs{dv, ev, Q};
    char* sc = static_cast<char*>(h->batch_scratch.p);
    int* d_lt[2] = {reinterpret_cast<int*>(sc), reinterpret_cast<int*>(sc + lt_set)};
    double* d_grid[2] = {reinterpret_cast<double*>(sc + 2 * lt_set),
                         reinterpret_cast<double*>(sc + 2 * lt_set + g_set)};

    // host staging lives in the handle; every copy below completes before it is reused
    h->batch_host_st.assign((size_t)chase::ST_WORDS * B, 0);
    for (int64_t p = 0; p < B; ++p) {
        int64_t* s = &h->batch_host_st[(size_t)p * chase::ST_WORDS];
        s[chase::ST_TO] = n[p];
        s[chase::ST_LBEA] = n[p];
        s[chase::ST_ITER] = 0;
    }
    hipStream_t us = h->stream;  // the caller's stream: uploads, results
    hipStream_t as = h->batch_stream[0], cs = h->batch_stream[1];
    int* pin = h->batch_pinned;
    pin[0] = 0;
    SEP_HIP(hipMemcpyAsync(d_probs, L.probs.data(), probs_b, hipMemcpyHostToDevice, us));
    SEP_HIP(hipMemcpyAsync(d_work, L.work.data(), work_b, hipMemcpyHostToDevice, us));
    SEP_HIP(hipMemcpyAsync(d_order, L.order.data(), order_b, hipMemcpyHostToDevice, us));
    SEP_HIP(hipMemcpyAsync(d_st, h->batch_host_st.data(), st_b, hipMemcpyHostToDevice, us));
    SEP_HIP(hipMemcpyAsync(d_count, pin, sizeof(int), hipMemcpyHostToDevice, us));
    SEP_HIP(hipMemcpyAsync(d_bufs, &host_bufs, sizeof(BatchBufs), hipMemcpyHostToDevice, us));
    SEP_HIP(hipStreamSynchronize(us));  // also orders the caller's dv / ev / Q before the chase

    hipEvent_t* chased = h->batch_ev;     // [2] by round parity
    hipEvent_t* applied = h->batch_ev + 2;
    const int nwork = (int)L.work.size();
    const int64_t max_rounds = 100 * (int64_t)L.nmax + 3;
    // (the chase's own static LDS counts towards the share).  Beside the apply, problems of
    // order up to kLdsChaseMaxN are chased in LDS, which needs kLdsBytesPerN bytes per order
    // of the largest problem (4096 problems of order 32..512: 68-70 ms against 70-74 ms
    // from global memory); for eigenvalues only, with nothing else loading memory, the
    // chase reads global memory (36 ms against 39-45 ms through LDS)
    const int in_lds = (Q != nullptr && L.nmax <= kLdsChaseMaxN) ? 1 : 0;
    const unsigned occ =
        cfg.CW >= 4 ? 0u : (unsigned)(kLdsPerCU / (4 * cfg.CW)) - chase::kChaseLdsBytes;
    const unsigned need = in_lds ? (kLdsBytesPerN * (unsigned)L.nmax + 15u) & ~15u : 0u;
    const unsigned reserve = std::max(occ, need);
    const int acap = cfg.AW > 0 ? cfg.AW * h->num_cus * 4 : 0;  // apply waves (AW per SIMD)
    // Round r: the chase (chase stream) waits for the apply of round r - 2, which used the
    // same buffers; the apply (apply stream) waits for the chase; the finished count is
    // copied to pinned host memory after the chase.  `first` marks the rounds whose
    // buffers no earlier round of the same sequence used.
    auto enqueue_chase = [&](int64_t r, bool first) -> hipError_t {
        const int b = (int)(r & 1);
        hipError_t e = hipSuccess;
        if (!first && (e = hipStreamWaitEvent(cs, applied[b], 0)) != hipSuccess) return e;
        if (h->chase_fast)
            hipLaunchKernelGGL(k_chase_all<true>, dim3((unsigned)B), dim3(kLanes), reserve, cs,
                               d_probs, d_order, d_bufs, d_lt[b], d_grid[b], d_depth[b], d_st,
                               cfg.ML, L.nmax, d_count, in_lds);
        else
            hipLaunchKernelGGL(k_chase_all<false>, dim3((unsigned)B), dim3(kLanes), reserve, cs,
                               d_probs, d_order, d_bufs, d_lt[b], d_grid[b], d_depth[b], d_st,
                               cfg.ML, L.nmax, d_count, in_lds);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(pin, d_count, sizeof(int), hipMemcpyDeviceToHost, cs)) !=
            hipSuccess)
            return e;
        return hipEventRecord(chased[b], cs);
    };
    // The apply of round r, after its chase (chase_done: that chase's event needs no wait,
    // because it completed before this work was enqueued).
    auto enqueue_apply = [&](int64_t r, bool chase_done) -> hipError_t {
        const int b = (int)(r & 1);
        hipError_t e = hipSuccess;
        if (!chase_done && (e = hipStreamWaitEvent(as, chased[b], 0)) != hipSuccess) return e;
        if (!Q) {
            // eigenvalues only: nothing to apply (the ordering events stay the same)
        } else if (cfg.KF == 16) {
            launch_apply<16, 1>(nwork, acap, as, d_work, d_probs, d_lt[b], d_grid[b], d_depth[b], d_bufs);
        } else if (cfg.KF == 8 && cfg.R == 2) {
            launch_apply<8, 2>(nwork, acap, as, d_work, d_probs, d_lt[b], d_grid[b], d_depth[b], d_bufs);
        } else if (cfg.KF == 8) {
            launch_apply<8, 1>(nwork, acap, as, d_work, d_probs, d_lt[b], d_grid[b], d_depth[b], d_bufs);
        } else {
            launch_apply<4, 1>(nwork, acap, as, d_work, d_probs, d_lt[b], d_grid[b], d_depth[b], d_bufs);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        return hipEventRecord(applied[b], as);
    };
    int64_t round = 0;
    int done = 0;
    int rc = SEP_OK;
    if (h->batch_eager) {
        // one round at a time; the host reads the finished count after each chase
        for (; round < max_rounds && done < B; ++round) {
            if (enqueue_chase(round, round < 2) != hipSuccess ||
                enqueue_apply(round, false) != hipSuccess ||
                hipEventSynchronize(chased[round & 1]) != hipSuccess) {
                rc = SEP_ERR_HIP;
                break;
            }
            done = pin[0];
        }
    } else {
        // The chase of round 0 runs on its own; then kGraphRounds pairs (the apply of round
        // r beside the chase of round r + 1) are captured once into a graph (the chase
        // stream forks from and joins the apply stream inside the capture) and replayed
        // until the finished count reaches B.  A replay therefore ends between two pairs,
        // which are synchronised anyway (the chase of round r + 2 waits for the apply of
        // round r, which used the same buffers), so no apply runs alone at a replay
        // boundary, as it would if a replay held whole rounds (chase, then apply): replays
        // on one stream run one after the other.  When the count reaches B after the chase
        // of round c, no problem was left with rotations in round c, so the missing apply
        // of the last replay's last chase has nothing to do.  Finished problems return at
        // once from the chase and leave nothing to apply, so rounds past the end cost two
        // near-empty launches.  Successive replays on the apply stream are ordered, which
        // supplies the cross-replay dependencies of a replay's first pair.
        const std::vector<int64_t> key = {
            (int64_t)(Q != nullptr), (int64_t)d_probs, (int64_t)d_work,
            (int64_t)d_st, (int64_t)d_depth[0], (int64_t)d_lt[0], (int64_t)d_grid[0],
            (int64_t)d_order,
            (int64_t)d_lt[1], (int64_t)d_grid[1], (int64_t)d_count, (int64_t)pin, B, nwork,
            L.nmax, cfg.ML, cfg.KF, cfg.R, cfg.CW, cfg.AW, (int64_t)cs, (int64_t)as,
            (int64_t)h->chase_fast};  // the chase kernel the rounds launch
        if (!h->batch_graph || h->batch_graph_key != key) {
            if (h->batch_graph) {
                (void)hipGraphExecDestroy(h->batch_graph);
                h->batch_graph = nullptr;
            }
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamBeginCapture(as, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) e = hipEventRecord(h->batch_ev[4], as);
            if (e == hipSuccess) e = hipStreamWaitEvent(cs, h->batch_ev[4], 0);
            // pair j: the chase of round j + 1 (waiting for the apply of round j - 1 of this
            // replay, if any) and the apply of round j (its chase ran in this replay's
            // previous pair, or before the replay when j = 0); rounds relative to the
            // replay's first apply, whose round is even
            for (int j = 0; j < kGraphRounds && e == hipSuccess; ++j) {
                e = enqueue_chase(j + 1, j == 0);
                if (e == hipSuccess) e = enqueue_apply(j, j == 0);
            }
            if (e == hipSuccess) e = hipEventRecord(h->batch_ev[5], cs);
            if (e == hipSuccess) e = hipStreamWaitEvent(as, h->batch_ev[5], 0);
            const hipError_t ee = hipStreamEndCapture(as, &graph);
            if (e == hipSuccess) e = ee;
            if (e == hipSuccess) e = hipGraphInstantiate(&h->batch_graph, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                h->batch_graph = nullptr;
                set_error("batched: graph capture failed: %s", hipGetErrorString(e));
                return SEP_ERR_HIP;
            }
            h->batch_graph_key = key;
        }
        // the chase of round 0, then one replay in flight ahead of the one whose count the
        // host reads
        hipEvent_t ran = h->batch_ev[6];  // end-of-replay marker (never captured)
        if (enqueue_chase(0, true) != hipSuccess ||
            hipStreamWaitEvent(as, chased[0], 0) != hipSuccess ||
            hipGraphLaunch(h->batch_graph, as) != hipSuccess)
            rc = SEP_ERR_HIP;
        round = 1 + kGraphRounds;  // chases enqueued so far
        while (rc == SEP_OK && round < max_rounds) {
            if (hipEventRecord(ran, as) != hipSuccess ||
                hipGraphLaunch(h->batch_graph, as) != hipSuccess ||
                hipEventSynchronize(ran) != hipSuccess) {
                rc = SEP_ERR_HIP;
                break;
            }
            round += kGraphRounds;
            done = pin[0];  // at least the count after the replay that just ended
            if (done >= B) break;
        }
    }
    if (hipStreamSynchronize(as) != hipSuccess || hipStreamSynchronize(cs) != hipSuccess)
        rc = SEP_ERR_HIP;
    if (rc) {
        set_error("batched: a HIP call failed in round %lld", (long long)round);
        return rc;
    }
    done = pin[0];
    if (done < B) {
        set_error("batched: %d of %lld problems unfinished after %lld rounds", (int)(B - done),
                  (long long)B, (long long)round);
        return SEP_ERR_HIP;
    }
    SEP_HIP(hipMemcpyAsync(h->batch_host_st.data(), d_st, st_b, hipMemcpyDeviceToHost, us));
    SEP_HIP(hipStreamSynchronize(us));
    std::vector<int32_t> conv((size_t)B);
    std::vector<int64_t> its((size_t)B);
    int all = 1;
    for (int64_t p = 0; p < B; ++p) {
        const int64_t* s = &h->batch_host_st[(size_t)p * chase::ST_WORDS];
        conv[p] = s[chase::ST_EARLY] ? 0 : 1;
        its[p] = s[chase::ST_ITER];
        all &= conv[p];
    }
    if (converged)
        SEP_HIP(hipMemcpyAsync(converged, conv.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice,
                               us));
    if (iters)
        SEP_HIP(hipMemcpyAsync(iters, its.data(), sizeof(int64_t) * B, hipMemcpyHostToDevice, us));
    SEP_HIP(hipStreamSynchronize(us));
    return all ? SEP_OK : SEP_NOT_CONVERGED;
}

std::vector<int64_t> layout_key(int64_t B, const int64_t* n, const int64_t* off_d,
                                const int64_t* off_q)
{
    std::vector<int64_t> k;
    k.reserve(3 * B + 1);
    k.push_back(B);
    k.insert(k.end(), n, n + B);
    k.insert(k.end(), off_d, off_d + B);
    k.insert(k.end(), off_q, off_q + B);
    return k;
}

__global__ void k_identity_batch(double* Q, const int64_t* __restrict__ nq,
                                 const int64_t* __restrict__ off_q, int B)
{
    const int p = blockIdx.y;
    if (p >= B) return;
    const int64_t n = nq[p];
    double* q = Q + off_q[p];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x)
        q[e] = (e % (n + 1) == 0) ? 1.0 : 0.0;
}

}  // namespace

int qr_batched_f64(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                   const int64_t* off_q, double* dv, double* ev, double* Q, int64_t max_layers,
                   int32_t* converged, int64_t* iters)
{
    if (int rc = check_batch(B, n, off_d, off_q, h->allow_odd)) return rc;
    if (B == 0) return SEP_OK;
    if (!dv || !ev) {  // Q == NULL: eigenvalues only
        set_error("batched: null device pointer");
        return SEP_ERR_ARG;
    }
    if (max_layers < 0 || max_layers > 256) {
        set_error("batched: max_layers=%lld outside [0, 256] (0 = tuned / default)",
                  (long long)max_layers);
        return SEP_ERR_BAD_LAYERS;
    }
    Config cfg = kConfigs[0];
    if (h->tune_choice >= 0 && h->tune_key == layout_key(B, n, off_d, off_q))
        cfg = kConfigs[h->tune_choice];
    if (h->force_kf > 0) {
        cfg.KF = h->force_kf;
        cfg.R = h->force_r;
    }
    if (max_layers > 0) cfg.ML = (int)max_layers;
    return run_batched(h, B, n, off_d, off_q, dv, ev, Q, cfg, converged, iters);
}

int set_batched_config(sep_handle* h, int fused_depth, int rows_per_lane)
{
    if (fused_depth == 0 && rows_per_lane == 0) {
        h->force_kf = h->force_r = 0;
        return SEP_OK;
    }
    const bool ok = (fused_depth == 8 && (rows_per_lane == 1 || rows_per_lane == 2)) ||
                    ((fused_depth == 4 || fused_depth == 16) && rows_per_lane == 1);
    if (!ok) {
        set_error("batched: unsupported configuration fused_depth=%d rows_per_lane=%d "
                  "(supported: 16/1, 8/1, 8/2, 4/1; 0/0 restores the autotuned choice)",
                  fused_depth, rows_per_lane);
        return SEP_ERR_ARG;
    }
    h->force_kf = fused_depth;
    h->force_r = rows_per_lane;
    return SEP_OK;
}

int tune_batched_f64(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                     const int64_t* off_q, const double* dv, const double* ev, int32_t* choice)
{
    if (int rc = check_batch(B, n, off_d, off_q, h->allow_odd)) return rc;
    if (B == 0) {
        if (choice) *choice = 0;
        return SEP_OK;
    }
    if (!dv || !ev) {
        set_error("batched: null device pointer");
        return SEP_ERR_ARG;
    }
    // scratch copies of the inputs (the caller's arrays are not modified)
    int64_t nd = 0, nq = 0;
    for (int64_t p = 0; p < B; ++p) {
        nd = std::max(nd, off_d[p] + n[p]);
        nq = std::max(nq, off_q[p] + n[p] * n[p]);
    }
    DevBuf cd, ce, cq, cmeta;
    auto release = [&]() {
        cd.release();
        ce.release();
        cq.release();
        cmeta.release();
    };
    int rc = SEP_OK;
    if ((rc = cd.reserve(sizeof(double) * nd)) || (rc = ce.reserve(sizeof(double) * nd)) ||
        (rc = cq.reserve(sizeof(double) * nq)) || (rc = cmeta.reserve(sizeof(int64_t) * 2 * B))) {
        release();
        return rc;
    }
    int64_t* d_n = cmeta.as<int64_t>();
    int64_t* d_offq = d_n + B;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best = 1e30f;
    int best_i = 0;
    hipError_t he = hipMemcpyAsync(d_n, n, sizeof(int64_t) * B, hipMemcpyHostToDevice, h->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(d_offq, off_q, sizeof(int64_t) * B, hipMemcpyHostToDevice, h->stream);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    // one timed solve of configuration c on fresh copies of the inputs (ms, or < 0 on error)
    auto time_config = [&](int c) -> float {
        he = hipMemcpyAsync(cd.p, dv, sizeof(double) * nd, hipMemcpyDeviceToDevice, h->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(ce.p, ev, sizeof(double) * nd, hipMemcpyDeviceToDevice, h->stream);
        if (he != hipSuccess) return -1.0f;
        hipLaunchKernelGGL(k_identity_batch, dim3(64, (unsigned)B), dim3(256), 0, h->stream,
                           cq.as<double>(), d_n, d_offq, (int)B);
        if ((he = hipGetLastError()) != hipSuccess) return -1.0f;
        if ((he = hipEventRecord(e0, h->stream)) != hipSuccess) return -1.0f;
        rc = run_batched(h, B, n, off_d, off_q, cd.as<double>(), ce.as<double>(), cq.as<double>(),
                         kConfigs[c], nullptr, nullptr);
        if (rc == SEP_NOT_CONVERGED) rc = SEP_OK;
        if (rc) return -1.0f;
        if ((he = hipEventRecord(e1, h->stream)) != hipSuccess) return -1.0f;
        if ((he = hipEventSynchronize(e1)) != hipSuccess) return -1.0f;
        float ms = 0;
        if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return -1.0f;
        if (std::getenv("SEP_TUNE_VERBOSE"))
            std::fprintf(stderr, "[sep tune] config %d (ML %d KF %d R %d CW %d AW %d): %.2f ms\n",
                         c, kConfigs[c].ML, kConfigs[c].KF, kConfigs[c].R, kConfigs[c].CW,
                         kConfigs[c].AW, ms);
        return ms;
    };
    // Every configuration once (its first run also captures its round graph), then the three
    // fastest twice more, each keeping its fastest run: the leading configurations are often
    // within a few percent of each other, about the run-to-run spread of one solve.
    float t[kNumConfigs];
    bool ok = he == hipSuccess;
    for (int c = 0; c < kNumConfigs && ok; ++c) ok = (t[c] = time_config(c)) >= 0.0f;
    int top[3] = {-1, -1, -1};
    for (int j = 0; j < 3 && ok; ++j)
        for (int c = 0; c < kNumConfigs; ++c)
            if (c != top[0] && c != top[1] && (top[j] < 0 || t[c] < t[top[j]])) top[j] = c;
    for (int rep = 0; rep < 2 && ok; ++rep)
        for (int j = 0; j < 3 && ok; ++j) {
            const float ms = time_config(top[j]);
            ok = ms >= 0.0f;
            if (ok) t[top[j]] = std::min(t[top[j]], ms);
        }
    for (int j = 0; j < 3 && ok; ++j)
        if (t[top[j]] < best) {
            best = t[top[j]];
            best_i = top[j];
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(h->stream);
    release();
    if (he != hipSuccess) {
        set_error("batched tuner: %s", hipGetErrorString(he));
        return SEP_ERR_HIP;
    }
    if (rc) return rc;
    h->tune_key = layout_key(B, n, off_d, off_q);
    h->tune_choice = best_i;
    if (choice) *choice = best_i;
    return SEP_OK;
}

}  // namespace sep

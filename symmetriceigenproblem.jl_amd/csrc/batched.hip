// batched.hip -- thousands of independent qr_algorithm! solves (Float64) per call (no
// counterpart in the reference, which solves one matrix per call).
//
// The outer loop of qr_algorithm! (src/bulge_chasing.jl:166-285) runs for all problems at
// once.  One round is two launches over the whole batch:
//   k_chase_all:  one wavefront per problem runs one block of sweeps with the single
//                 solve's device chase (chase.hpp, src/bulge_chasing.jl:166-276), in the
//                 reference's (n-1) x max_layers grid layout; a problem that has finished
//                 (no sweep in its last block, or maxiter exceeded) returns at once;
//   k_apply_all:  the block's rotations are applied to each active problem's Q in wave
//                 order t = i + 2L (src/wave.jl:62-162), KF layers per pass over Q, R rows
//                 of Q per lane, the window of 2 KF columns per row in VGPRs, and the
//                 (c, s) of 16 time steps staged in LDS at a time.  Rotations beyond a
//                 column's layer count are read as the identity, so the grid never needs
//                 an identity fill.  The rotation is Rotation2 rmul! (src/simple_givens.jl:
//                 53-64) operation for operation, so Q equals the layer-major product of
//                 the C restatement bit for bit.
// The host replays rounds until a device counter says every problem is done (every
// round either finishes a problem or advances its iteration count, so the number of
// rounds is bounded by max maxiter + 2).  The autotuner picks max_layers, the fused depth
// KF and the rows per lane R.
#include <algorithm>
#include <numeric>
#include <type_traits>

#include "chase.hpp"
#include "sep_internal.hpp"

namespace sep {
namespace {

constexpr int kLanes = 64;
constexpr int kStageSteps = 16;  // wave-order time steps staged per LDS fill
constexpr int kMaxBatchedN = 8192;

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

// One block of sweeps for every unfinished problem (one wavefront each).
__global__ void __launch_bounds__(kLanes)
    k_chase_all(const BProb* __restrict__ probs, double* dv, double* ev, int* lt, double* grid,
                int64_t* st_all, int ML, int* done_count)
{
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    int64_t* st = st_all + (int64_t)p * chase::ST_WORDS;
    if (__builtin_amdgcn_readfirstlane((int)st[ST_DONE])) return;
    const BProb pr = probs[p];
    chase::Problem<double> P{dv + pr.off_d, ev + pr.off_d, lt + pr.off_lt, grid + 2 * pr.off_g,
                             pr.n, ML, (int64_t)(pr.n - 1), 0};
    const double tol = 2.220446049250313e-16;  // eps(Float64)
    chase::chase_block(P, tol, 100 * (int64_t)pr.n, st, lane);
    __threadfence_block();
    if (lane == 0) {
        const bool done = st[chase::ST_EARLY] != 0 || st[chase::ST_DID] == 0;
        if (done) {
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

#pragma clang fp contract(off)

// Apply the block's rotations to Q for rows [row0, row0 + 64 R) of one problem.
template <int KF, int R>
__global__ void __launch_bounds__(kLanes)
    k_apply_all(const BWork* __restrict__ work, const BProb* __restrict__ probs,
                const int* __restrict__ lt, const double* __restrict__ grid,
                const int64_t* __restrict__ st_all, double* __restrict__ Q)
{
    constexpr int W = 2 * KF;
    constexpr int SB = kStageSteps;
    static_assert(SB % W == 0, "staged steps must be a multiple of the window");
    __shared__ double2 stg[SB][KF];

    const BWork wk = work[blockIdx.x];
    const int64_t* st = st_all + (int64_t)wk.prob * chase::ST_WORDS;
    // nothing to apply: finished (no sweep, or maxiter exceeded: the reference returns
    // without applying that block, src/bulge_chasing.jl:180-182)
    if (__builtin_amdgcn_readfirstlane((int)st[ST_DONE])) return;
    const int depth = __builtin_amdgcn_readfirstlane((int)st[chase::ST_MAXLAYER]);
    const BProb pr = probs[wk.prob];
    const int n = pr.n;
    const int lane = threadIdx.x;
    const int* __restrict__ ltp = lt + pr.off_lt;
    const double* __restrict__ g = grid + 2 * pr.off_g;
    double* __restrict__ Qp = Q + pr.off_q;
    int rowi[R];
    bool act[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = wk.row0 + r * kLanes + lane;
        act[r] = row < n;
        rowi[r] = min(row, n - 1);
    }
    const int Tn = n + 2 * KF - 2;  // time steps of one pass, the last one only stores
    for (int L0 = 0; L0 < depth; L0 += KF) {
        double win[R][W];
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int w = 0; w < W; ++w) win[r][w] = 0.0;
            win[r][0] = act[r] ? Qp[rowi[r]] : 0.0;
        }
        for (int t0 = 0; t0 < Tn; t0 += SB) {
            __syncthreads();  // the previous fill has been read
            for (int q = lane; q < SB * KF; q += kLanes) {
                const int s = q / KF, l = q % KF;
                const int i = t0 + s - 2 * l, L = L0 + l;
                double2 v = make_double2(1.0, 0.0);
                if (i >= 0 && i <= n - 2 && L < depth && L < ltp[i])
                    v = *reinterpret_cast<const double2*>(g + 2 * (i + (int64_t)(n - 1) * L));
                stg[s][l] = v;
            }
            __syncthreads();
            Unroll<SB, 0>::run([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                const int t = t0 + s;
                if (t < Tn) {
                    const int jn = t + 1;  // column entering the window
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        win[r][(s + 1) % W] =
                            (act[r] && jn < n) ? Qp[rowi[r] + (int64_t)jn * n] : 0.0;
                    Unroll<KF, 0>::run([&](auto lc) {
                        constexpr int l = decltype(lc)::value;
                        constexpr int ia = ((s - 2 * l) % W + W) % W;
                        constexpr int ib = (ia + 1) % W;
                        const double2 cs = stg[s][l];
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const double x = win[r][ia], y = win[r][ib];
                            win[r][ia] = x * cs.x + y * cs.y;
                            win[r][ib] = x * -cs.y + y * cs.x;
                        }
                    });
                    const int jo = t - 2 * KF + 2;  // column leaving the window
                    if (jo >= 0 && jo < n) {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            if (act[r]) Qp[rowi[r] + (int64_t)jo * n] = win[r][(s + 2) % W];
                    }
                }
            });
        }
    }
}

#pragma clang fp contract(on)

struct Config {
    int ML, KF, R;
};
constexpr Config kConfigs[] = {
    {32, 8, 1}, {32, 8, 2}, {64, 8, 1}, {64, 8, 2}, {16, 8, 1}, {32, 4, 1}, {16, 4, 1},
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);

struct Layout {
    std::vector<BProb> probs;
    std::vector<BWork> work;
    int nmax = 0;
    int64_t tot_lt = 0, tot_g = 0;
};

int check_batch(int64_t B, const int64_t* n, const int64_t* off_d, const int64_t* off_q)
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
        if (n[p] % 2) {
            set_error("batched: problem %lld has odd order %lld (the reference asserts iseven(n))",
                      (long long)p, (long long)n[p]);
            return SEP_ERR_ODD_N;
        }
    }
    return SEP_OK;
}

void build_layout(int64_t B, const int64_t* n, const int64_t* off_d, const int64_t* off_q,
                  const Config& cfg, Layout& L)
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
    std::vector<int> order(B);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    const int rows = kLanes * cfg.R;
    for (int p : order)
        for (int r0 = 0; r0 < (int)n[p]; r0 += rows) L.work.push_back(BWork{p, r0});
}

template <int KF, int R>
void launch_apply(int nwork, hipStream_t s, const BWork* w, const BProb* p, const int* lt,
                  const double* g, const int64_t* st, double* Q)
{
    hipLaunchKernelGGL((k_apply_all<KF, R>), dim3(nwork), dim3(kLanes), 0, s, w, p, lt, g, st, Q);
}

// The whole batched solve on the handle's stream; returns after completion.
int run_batched(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                const int64_t* off_q, double* dv, double* ev, double* Q, const Config& cfg,
                int32_t* converged, int64_t* iters)
{
    Layout L;
    build_layout(B, n, off_d, off_q, cfg, L);
    const size_t probs_b = sizeof(BProb) * (size_t)B;
    const size_t work_b = sizeof(BWork) * L.work.size();
    const size_t st_b = sizeof(int64_t) * chase::ST_WORDS * (size_t)B;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    if (int rc = h->batch_meta.reserve(al(probs_b) + al(work_b) + al(st_b) + 256)) return rc;
    if (int rc = h->batch_scratch.reserve(al(sizeof(int) * (size_t)L.tot_lt) +
                                          sizeof(double) * 2 * (size_t)L.tot_g + 256))
        return rc;
    char* m = static_cast<char*>(h->batch_meta.p);
    BProb* d_probs = reinterpret_cast<BProb*>(m);
    BWork* d_work = reinterpret_cast<BWork*>(m + al(probs_b));
    int64_t* d_st = reinterpret_cast<int64_t*>(m + al(probs_b) + al(work_b));
    int* d_count = reinterpret_cast<int*>(m + al(probs_b) + al(work_b) + al(st_b));
    char* sc = static_cast<char*>(h->batch_scratch.p);
    int* d_lt = reinterpret_cast<int*>(sc);
    double* d_grid = reinterpret_cast<double*>(sc + al(sizeof(int) * (size_t)L.tot_lt));

    // host staging lives in the handle; every copy below completes before it is reused
    h->batch_host_st.assign((size_t)chase::ST_WORDS * B, 0);
    for (int64_t p = 0; p < B; ++p) {
        int64_t* s = &h->batch_host_st[(size_t)p * chase::ST_WORDS];
        s[chase::ST_TO] = n[p];
        s[chase::ST_LBEA] = n[p];
        s[chase::ST_ITER] = 0;
    }
    int done = 0;
    SEP_HIP(hipMemcpyAsync(d_probs, L.probs.data(), probs_b, hipMemcpyHostToDevice, h->stream));
    SEP_HIP(hipMemcpyAsync(d_work, L.work.data(), work_b, hipMemcpyHostToDevice, h->stream));
    SEP_HIP(hipMemcpyAsync(d_st, h->batch_host_st.data(), st_b, hipMemcpyHostToDevice, h->stream));
    SEP_HIP(hipMemcpyAsync(d_count, &done, sizeof(int), hipMemcpyHostToDevice, h->stream));
    SEP_HIP(hipStreamSynchronize(h->stream));

    const int nwork = (int)L.work.size();
    const int64_t max_rounds = 100 * (int64_t)L.nmax + 3;
    int64_t round = 0;
    for (; round < max_rounds && done < B; ++round) {
        hipLaunchKernelGGL(k_chase_all, dim3((unsigned)B), dim3(kLanes), 0, h->stream, d_probs,
                           dv, ev, d_lt, d_grid, d_st, cfg.ML, d_count);
        SEP_CHECK_LAUNCH();
        if (cfg.KF == 8 && cfg.R == 2)
            launch_apply<8, 2>(nwork, h->stream, d_work, d_probs, d_lt, d_grid, d_st, Q);
        else if (cfg.KF == 8)
            launch_apply<8, 1>(nwork, h->stream, d_work, d_probs, d_lt, d_grid, d_st, Q);
        else
            launch_apply<4, 1>(nwork, h->stream, d_work, d_probs, d_lt, d_grid, d_st, Q);
        SEP_CHECK_LAUNCH();
        SEP_HIP(hipMemcpyAsync(&done, d_count, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        SEP_HIP(hipStreamSynchronize(h->stream));
    }
    if (done < B) {
        set_error("batched: %d of %lld problems unfinished after %lld rounds", (int)(B - done),
                  (long long)B, (long long)round);
        return SEP_ERR_HIP;
    }
    SEP_HIP(hipMemcpyAsync(h->batch_host_st.data(), d_st, st_b, hipMemcpyDeviceToHost, h->stream));
    SEP_HIP(hipStreamSynchronize(h->stream));
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
                               h->stream));
    if (iters)
        SEP_HIP(hipMemcpyAsync(iters, its.data(), sizeof(int64_t) * B, hipMemcpyHostToDevice,
                               h->stream));
    SEP_HIP(hipStreamSynchronize(h->stream));
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
    if (int rc = check_batch(B, n, off_d, off_q)) return rc;
    if (B == 0) return SEP_OK;
    if (!dv || !ev || !Q) {
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
                    (fused_depth == 4 && rows_per_lane == 1);
    if (!ok) {
        set_error("batched: unsupported configuration fused_depth=%d rows_per_lane=%d "
                  "(supported: 8/1, 8/2, 4/1; 0/0 restores the autotuned choice)",
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
    if (int rc = check_batch(B, n, off_d, off_q)) return rc;
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
    for (int c = 0; c < kNumConfigs && he == hipSuccess && rc == SEP_OK; ++c) {
        he = hipMemcpyAsync(cd.p, dv, sizeof(double) * nd, hipMemcpyDeviceToDevice, h->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(ce.p, ev, sizeof(double) * nd, hipMemcpyDeviceToDevice, h->stream);
        if (he != hipSuccess) break;
        hipLaunchKernelGGL(k_identity_batch, dim3(64, (unsigned)B), dim3(256), 0, h->stream,
                           cq.as<double>(), d_n, d_offq, (int)B);
        if ((he = hipGetLastError()) != hipSuccess) break;
        if ((he = hipEventRecord(e0, h->stream)) != hipSuccess) break;
        rc = run_batched(h, B, n, off_d, off_q, cd.as<double>(), ce.as<double>(), cq.as<double>(),
                         kConfigs[c], nullptr, nullptr);
        if (rc == SEP_NOT_CONVERGED) rc = SEP_OK;
        if (rc) break;
        if ((he = hipEventRecord(e1, h->stream)) != hipSuccess) break;
        if ((he = hipEventSynchronize(e1)) != hipSuccess) break;
        float ms = 0;
        if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
        if (ms < best) {
            best = ms;
            best_i = c;
        }
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

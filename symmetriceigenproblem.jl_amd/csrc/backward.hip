// backward.hip -- reverse-mode derivative of the eigendecomposition of a symmetric
// tridiagonal T(d, e) (no counterpart in the reference, which has no AD support).
//
// With T = Q diag(lam) Q^T and cotangents (lam_bar, Q_bar):
//   W = Q^T Q_bar,  M = diag(lam_bar) + F o W,  F_ij = 1/(lam_j - lam_i) (i != j), 0 (i == j)
//   T_bar = Q M Q^T,  d_bar_k = T_bar_kk,  e_bar_k = T_bar_{k,k+1} + T_bar_{k+1,k}.
// T_bar is never formed.  With P = Q M, d_bar_k = <P_k., Q_k.>, and e_bar_k =
// <P_k., Q_{k+1}.> + <P_{k+1}., Q_k.>, so every row k of P contributes three dot products
// with rows k-1, k, k+1 of Q; those are taken in the epilogue of the second product, and P
// is never written either.
//
// Kernels (Float64 MFMA, v_mfma_f64_16x16x4f64):
//   k_bwd_gemm<kModeM>: W tile = Q^T Q_bar tile; epilogue writes M = F o W + diag(lam_bar)
//   k_bwd_gemm<kModeP>: P tile = Q M tile; epilogue reduces the tile's three row dot-product
//                       partial sums and stores them per column tile (no atomics)
//   k_bwd_reduce:       sums the column-tile partials in a fixed order -> d_bar, e_bar
// Block tile 128 x 128, k step 16, 256 threads = 2 x 2 waves of 64 x 64 (4 x 4 MFMA tiles of
// 16 x 16, four f64 accumulators per lane each), double-buffered LDS, one barrier per k step;
// batches whose orders are all at most 512 use 64 x 64 block tiles (waves of 32 x 32).
// Problems are described by a table, so the single and the batched entry points share the
// kernels (grid z = problem; tiles beyond a problem's order exit at once).
#include <algorithm>
#include <numeric>
#include <vector>

#include "sep_internal.hpp"

namespace sep {
namespace {

using f64x4 = double __attribute__((ext_vector_type(4)));

constexpr int kBT = 128;      // block tile (rows = columns)
constexpr int kBK = 16;       // k step
constexpr int kThreads = 256;
constexpr int kSmallTileMaxN = 512;  // largest order for which 64-tiles are used

struct BwdProb {
    int n;
    int ld;         // leading dimension of Q and Q_bar
    int64_t off_q;  // offset of Q / Q_bar (elements)
    int64_t off_d;  // offset of lam / lam_bar / d_bar / e_bar, and partial-sum column
    int64_t off_m;  // offset of this problem's M (n x n, ld n) in the scratch
};

enum { kModeM = 0, kModeP = 1 };

// Tile of this workgroup.  Workgroups go to the XCDs in turn (workgroup b to XCD b mod X,
// in dispatch order b = x + y nt), and each XCD has its own L2.  For a square tile grid
// whose size divides evenly, each XCD takes a contiguous run of 8 x 8 tile groups, in
// dispatch order, so the workgroups resident on one XCD at a time share 8 row panels of A
// and 8 column panels of B in its L2; otherwise the grid position is the tile.
__device__ __forceinline__ void tile_of(int xcds, int& tx, int& ty)
{
    constexpr int G = 8;
    const int nt = gridDim.x;
    tx = blockIdx.x;
    ty = blockIdx.y;
    if (gridDim.z != 1 || (int)gridDim.y != nt || nt % G != 0 || xcds <= 0 ||
        (nt * nt) % xcds != 0)
        return;
    const int b = blockIdx.x + blockIdx.y * nt;
    const int per = nt * nt / xcds;
    const int t = (b % xcds) * per + b / xcds;  // a bijection of 0 .. nt^2 - 1
    const int g = t / (G * G), w = t % (G * G), gpr = nt / G;
    ty = (g / gpr) * G + w / G;
    tx = (g % gpr) * G + w % G;
}

// At most 256 registers per lane (accumulators included), so that two workgroups share a
// CU and each SIMD has a second wave to issue MFMAs while one waits on memory or a barrier.
// BT is the block tile: 128 (2 x 2 waves of 64 x 64) for large orders, 64 (2 x 2 waves of
// 32 x 32) for batches of small problems, where a 128 tile would mostly compute padding.
// One BT x BT tile (tx, ty) of problem pr: the product, then the epilogue of MODE.  As and
// Bs are the workgroup's double-buffered LDS tiles; the caller separates two tiles run by
// one workgroup with a barrier.
template <int MODE, int BT>
__device__ __forceinline__ void bwd_tile(const BwdProb pr, int tx, int ty,
                                         const double* __restrict__ lam,
                                         const double* __restrict__ lam_bar,
                                         const double* __restrict__ Q,
                                         const double* __restrict__ Qb, double* __restrict__ M,
                                         double* __restrict__ part, int64_t part_ld,
                                         double (&As)[2][kBK][BT + 4],
                                         double (&Bs)[2][kBK][BT + 4])
{
    constexpr int LDS_LD = BT + 4;
    constexpr int WT = BT / 32;          // 16 x 16 MFMA tiles per wave and dimension
    constexpr int TPI = kThreads / BT;   // threads per A row (first product) / B row
    constexpr int EPT = kBK / TPI;       // elements each of them loads
    constexpr int EPK = BT / 16;         // second product: elements per thread along i
    static_assert(EPT * TPI == kBK && EPK * 16 == BT, "tile geometry");

    const int n = pr.n, ld = pr.ld;
    const int i0 = tx * BT, j0 = ty * BT;
    if (i0 >= n || j0 >= n) return;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int wr = wid >> 1, wc = wid & 1;
    const double* __restrict__ Qp = Q + pr.off_q;
    // B operand: Q_bar (ld) for the first product, M (ld n) for the second
    const double* __restrict__ Bp = (MODE == kModeM) ? Qb + pr.off_q : M + pr.off_m;
    const int ldb = (MODE == kModeM) ? ld : n;

    // this thread's A and B elements: A[i][k] = Q[k + i ld] (first product, Q^T) or
    // Q[i + k ld] (second product), B[k][j] = Bp[k + j ldb]
    const int ai = (MODE == kModeM) ? t / TPI : (t & 15) * EPK;
    const int ak = (MODE == kModeM) ? (t % TPI) * EPT : t >> 4;
    const int bj = t / TPI, bk = (t % TPI) * EPT;
    constexpr int AE = (MODE == kModeM) ? EPT : EPK;
    double ra[AE], rb[EPT];
    // Tiles entirely inside the matrix (all of them when n is a multiple of BT) load
    // without per-element range tests.
    const bool inner_ij = i0 + BT <= n && j0 + BT <= n;
    auto load_tile = [&](int k0) {
        const double* srcb = Bp + (int64_t)(j0 + bj) * ldb + k0 + bk;
        const double* srca = (MODE == kModeM) ? Qp + (int64_t)(i0 + ai) * ld + k0 + ak
                                              : Qp + (int64_t)(k0 + ak) * ld + i0 + ai;
        if (inner_ij && k0 + kBK <= n) {
#pragma unroll
            for (int u = 0; u < AE; ++u) ra[u] = srca[u];
#pragma unroll
            for (int u = 0; u < EPT; ++u) rb[u] = srcb[u];
            return;
        }
#pragma unroll
        for (int u = 0; u < AE; ++u) {
            const bool in = (MODE == kModeM) ? (i0 + ai < n && k0 + ak + u < n)
                                             : (k0 + ak < n && i0 + ai + u < n);
            ra[u] = in ? srca[u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < EPT; ++u) rb[u] = (j0 + bj < n && k0 + bk + u < n) ? srcb[u] : 0.0;
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int u = 0; u < AE; ++u) {
            if constexpr (MODE == kModeM)
                As[buf][ak + u][ai] = ra[u];
            else
                As[buf][ak][ai + u] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < EPT; ++u) Bs[buf][bk + u][bj] = rb[u];
    };

    f64x4 acc[WT][WT];
#pragma unroll
    for (int a = 0; a < WT; ++a)
#pragma unroll
        for (int b = 0; b < WT; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};

    const int nk = (n + kBK - 1) / kBK;
    load_tile(0);
    store_tile(0);
    __syncthreads();
    const int fr = lane & 15, fk = lane >> 4;  // fragment row / k of this lane
    constexpr int WS = BT / 2;                 // wave tile
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load_tile((kt + 1) * kBK);
#pragma unroll
        for (int ks = 0; ks < kBK / 4; ++ks) {
            const int kr = ks * 4 + fk;
            double a[WT], b[WT];
#pragma unroll
            for (int r = 0; r < WT; ++r) a[r] = As[cur][kr][wr * WS + r * 16 + fr];
#pragma unroll
            for (int c = 0; c < WT; ++c) b[c] = Bs[cur][kr][wc * WS + c * 16 + fr];
#pragma unroll
            for (int r = 0; r < WT; ++r)
#pragma unroll
                for (int c = 0; c < WT; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        if (kt + 1 < nk) store_tile(cur ^ 1);
        __syncthreads();
    }

    // accumulator (r, c), element v of this lane: row fk + 4 v, column fr of the 16 x 16 tile
    // (the v_mfma_f64_16x16x4f64 result layout: lane j + 16 (i mod 4), element i / 4)
    if constexpr (MODE == kModeM) {
        const double* __restrict__ lp = lam + pr.off_d;
        const double* __restrict__ lbp = lam_bar + pr.off_d;
        double* __restrict__ Mp = M + pr.off_m;
#pragma unroll
        for (int c = 0; c < WT; ++c) {
            const int j = j0 + wc * WS + c * 16 + fr;
            if (j >= n) continue;
            const double lj = lp[j];
#pragma unroll
            for (int r = 0; r < WT; ++r) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = i0 + wr * WS + r * 16 + fk + 4 * v;
                    if (i >= n) continue;
                    Mp[i + (int64_t)j * n] = (i == j) ? lbp[i] : acc[r][c][v] / (lj - lp[i]);
                }
            }
        }
    } else {
        // the LDS tiles are free after the last barrier: reuse As for the wave partials
        static_assert(2 * kBK * LDS_LD >= 2 * BT * 3, "partials fit in the A tiles");
        double* red = &As[0][0][0];  // [2 wc][BT rows][3]
#pragma unroll
        for (int r = 0; r < WT; ++r) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int il = wr * WS + r * 16 + fk + 4 * v;
                const int i = i0 + il;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int c = 0; c < WT; ++c) {
                    const int j = j0 + wc * WS + c * 16 + fr;
                    if (i < n && j < n) {
                        const double p = acc[r][c][v];
                        const double* qc = Qp + (int64_t)j * ld;
                        s0 = __builtin_fma(p, qc[i], s0);
                        if (i + 1 < n) s1 = __builtin_fma(p, qc[i + 1], s1);
                        if (i > 0) s2 = __builtin_fma(p, qc[i - 1], s2);
                    }
                }
#pragma unroll
                for (int off = 8; off > 0; off >>= 1) {
                    s0 += __shfl_xor(s0, off, 64);
                    s1 += __shfl_xor(s1, off, 64);
                    s2 += __shfl_xor(s2, off, 64);
                }
                if (fr == 0) {
                    double* o = red + ((size_t)wc * BT + il) * 3;
                    o[0] = s0;
                    o[1] = s1;
                    o[2] = s2;
                }
            }
        }
        __syncthreads();
        if (t < BT) {
            const int i = i0 + t;
            if (i < n) {
                const double* a = red + (size_t)t * 3;
                const double* b = red + ((size_t)BT + t) * 3;
                double* dst = part + (int64_t)ty * 3 * part_ld + pr.off_d + i;
                dst[0] = a[0] + b[0];
                dst[part_ld] = a[1] + b[1];
                dst[2 * part_ld] = a[2] + b[2];
            }
        }
    }
}

// Grid mode (tiles == nullptr): one tile per workgroup, (x, y) the tile, z the problem.
// List mode: the resident workgroups run the list's tiles blockIdx.x, + gridDim.x, ...
// For batches of small problems most tiles are short (orders 64..256, a few k steps), and
// one workgroup per tile left the CUs waiting on workgroup dispatch (0.5 waves resident per
// SIMD on average, the MFMA pipe ~35% busy); resident workgroups that loop over the tiles
// keep it fed.  Entries with prob < 0 pad the list's XCD layout.
template <int MODE, int BT>
__global__ void __launch_bounds__(kThreads, 2)
    k_bwd_gemm(const BwdProb* __restrict__ probs, const double* __restrict__ lam,
               const double* __restrict__ lam_bar, const double* __restrict__ Q,
               const double* __restrict__ Qb, double* __restrict__ M,
               double* __restrict__ part, int64_t part_ld, int xcds,
               const int4* __restrict__ tiles, int ntiles)
{
    __shared__ double As[2][kBK][BT + 4];
    __shared__ double Bs[2][kBK][BT + 4];
    if (tiles) {
        for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
            const int4 tl = tiles[w];
            if (tl.x < 0) continue;
            bwd_tile<MODE, BT>(probs[tl.x], tl.y, tl.z, lam, lam_bar, Q, Qb, M, part, part_ld,
                               As, Bs);
            __syncthreads();  // this tile's last LDS reads before the next tile's stores
        }
    } else {
        int tx, ty;
        tile_of(xcds, tx, ty);
        bwd_tile<MODE, BT>(probs[blockIdx.z], tx, ty, lam, lam_bar, Q, Qb, M, part, part_ld, As,
                           Bs);
    }
}

// d_bar_k = sum_ct S0[ct][k];  e_bar_k = sum_ct (S1[ct][k] + S2[ct][k+1])
__global__ void __launch_bounds__(256)
    k_bwd_reduce(const BwdProb* __restrict__ probs, const double* __restrict__ part,
                 int64_t part_ld, int bt, double* __restrict__ d_bar, double* __restrict__ e_bar)
{
    const BwdProb pr = probs[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= pr.n) return;
    const int nct = (pr.n + bt - 1) / bt;
    const int64_t o = pr.off_d + k;
    double d = 0.0, e = 0.0;
    for (int ct = 0; ct < nct; ++ct) {
        const double* p = part + (int64_t)ct * 3 * part_ld;
        d += p[o];
        if (k + 1 < pr.n) e += p[part_ld + o] + p[2 * part_ld + o + 1];
    }
    d_bar[o] = d;
    if (k + 1 < pr.n) e_bar[o] = e;
}

int run_backward(sep_handle* h, const std::vector<BwdProb>& probs, int nmax, int64_t tot_d,
                 int64_t tot_m, const double* lam, const double* Q, const double* lam_bar,
                 const double* Q_bar, double* d_bar, double* e_bar)
{
    const int64_t B = (int64_t)probs.size();
    if (B == 0 || nmax == 0) return SEP_OK;
    if (B > 65535) {
        set_error("eig_backward: at most 65535 problems per call (got %lld)", (long long)B);
        return SEP_ERR_ARG;
    }
    // 64-tiles when every order is small (the 128 tile's padding would dominate)
    const int bt = nmax <= kSmallTileMaxN ? 64 : kBT;
    const int nct = (nmax + bt - 1) / bt;
    // Batches run a tile list (list mode of k_bwd_gemm): every problem's tiles, problems
    // largest first, each problem on one XCD (the XCD with the least work so far, in n^3),
    // so that its tiles share Q and Q_bar in that XCD's L2.  Workgroup g runs the list
    // entries g, g + G, ... on XCD g mod X (G a multiple of X), so XCD x's entries sit at
    // positions x, x + X, ...; shorter queues are padded with prob = -1.
    std::vector<int4> tiles;
    if (B > 1) {
        const int X = std::max(1, h->num_xcds);
        std::vector<int> order((size_t)B);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return probs[a].n > probs[b].n; });
        std::vector<std::vector<int4>> q((size_t)X);
        std::vector<double> load((size_t)X, 0.0);
        for (int p : order) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            const int nt = (probs[p].n + bt - 1) / bt;
            for (int ty = 0; ty < nt; ++ty)
                for (int tx = 0; tx < nt; ++tx) q[x].push_back(make_int4(p, tx, ty, 0));
            load[x] += (double)probs[p].n * probs[p].n * probs[p].n;
        }
        size_t len = 0;
        for (auto& v : q) len = std::max(len, v.size());
        tiles.assign(len * X, make_int4(-1, 0, 0, 0));
        for (int x = 0; x < X; ++x)
            for (size_t j = 0; j < q[x].size(); ++j) tiles[j * X + x] = q[x][j];
    }
    const size_t probs_bytes = (sizeof(BwdProb) * (size_t)B + 15) / 16 * 16;
    const size_t meta_bytes = probs_bytes + sizeof(int4) * tiles.size();
    if (int rc = h->bwd_meta.reserve(meta_bytes)) return rc;
    if (int rc = h->bwd_M.reserve(sizeof(double) * (size_t)tot_m)) return rc;
    if (int rc = h->bwd_part.reserve(sizeof(double) * 3 * (size_t)nct * (size_t)(tot_d + 1)))
        return rc;
    // the table and the tile list are uploaded from a handle-owned host copy, and only when
    // they change
    std::vector<char> raw(meta_bytes, 0);
    std::memcpy(raw.data(), probs.data(), sizeof(BwdProb) * (size_t)B);
    if (!tiles.empty())
        std::memcpy(raw.data() + probs_bytes, tiles.data(), sizeof(int4) * tiles.size());
    if (h->bwd_meta_host != raw || h->bwd_meta_dev != h->bwd_meta.p) {
        h->bwd_meta_host = std::move(raw);
        SEP_HIP(hipMemcpyAsync(h->bwd_meta.p, h->bwd_meta_host.data(), meta_bytes,
                               hipMemcpyHostToDevice, h->stream));
        SEP_HIP(hipStreamSynchronize(h->stream));
        h->bwd_meta_dev = h->bwd_meta.p;
    }
    const BwdProb* dp = h->bwd_meta.as<BwdProb>();
    const int4* dtiles =
        tiles.empty() ? nullptr
                      : reinterpret_cast<const int4*>(static_cast<char*>(h->bwd_meta.p) + probs_bytes);
    const int ntiles = (int)tiles.size();
    const int64_t part_ld = tot_d + 1;
    double* Mb = h->bwd_M.as<double>();
    double* part = h->bwd_part.as<double>();
    // grid mode: (x, y) tiles, z problems; list mode: as many workgroups as fit the CUs at
    // once (64-tiles: 35 KB of LDS and at most 128 registers per lane, four per CU; 128-tiles
    // two)
    const int64_t per_cu = bt == 64 ? 4 : 2;
    const dim3 grid = dtiles ? dim3((unsigned)std::min<int64_t>(ntiles, per_cu * h->num_cus))
                             : dim3(nct, nct, (unsigned)B);
    if (bt == 64) {
        hipLaunchKernelGGL((k_bwd_gemm<kModeM, 64>), grid, dim3(kThreads), 0, h->stream, dp,
                           lam, lam_bar, Q, Q_bar, Mb, part, part_ld, h->num_xcds, dtiles, ntiles);
        SEP_CHECK_LAUNCH();
        hipLaunchKernelGGL((k_bwd_gemm<kModeP, 64>), grid, dim3(kThreads), 0, h->stream, dp,
                           lam, lam_bar, Q, Q_bar, Mb, part, part_ld, h->num_xcds, dtiles, ntiles);
    } else {
        hipLaunchKernelGGL((k_bwd_gemm<kModeM, kBT>), grid, dim3(kThreads), 0, h->stream, dp,
                           lam, lam_bar, Q, Q_bar, Mb, part, part_ld, h->num_xcds, dtiles, ntiles);
        SEP_CHECK_LAUNCH();
        hipLaunchKernelGGL((k_bwd_gemm<kModeP, kBT>), grid, dim3(kThreads), 0, h->stream, dp,
                           lam, lam_bar, Q, Q_bar, Mb, part, part_ld, h->num_xcds, dtiles, ntiles);
    }
    SEP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bwd_reduce, dim3((nmax + 255) / 256, (unsigned)B), dim3(256), 0,
                       h->stream, dp, part, part_ld, bt, d_bar, e_bar);
    SEP_CHECK_LAUNCH();
    return SEP_OK;
}

}  // namespace

int eig_backward_f64(sep_handle* h, int64_t n, const double* lam, const double* Q, int64_t ldq,
                     const double* lam_bar, const double* Q_bar, double* d_bar, double* e_bar)
{
    if (n < 0 || n > (int64_t)1 << 20 || ldq < std::max<int64_t>(n, 1)) {
        set_error("sep_eig_backward: invalid n=%lld ldq=%lld", (long long)n, (long long)ldq);
        return SEP_ERR_ARG;
    }
    if (n == 0) return SEP_OK;
    if (!lam || !Q || !lam_bar || !Q_bar || !d_bar || (n > 1 && !e_bar)) {
        set_error("sep_eig_backward: null device pointer");
        return SEP_ERR_ARG;
    }
    std::vector<BwdProb> p(1);
    p[0] = BwdProb{(int)n, (int)ldq, 0, 0, 0};
    // e_bar has n - 1 entries: the reduction writes e_bar[k] only for k + 1 < n
    return run_backward(h, p, (int)n, n, n * n, lam, Q, lam_bar, Q_bar, d_bar, e_bar);
}

int eig_backward_batched_f64(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                             const int64_t* off_q, const double* lam, const double* Q,
                             const double* lam_bar, const double* Q_bar, double* d_bar,
                             double* e_bar)
{
    if (B < 0 || (B > 0 && (!n || !off_d || !off_q))) {
        set_error("sep_eig_backward_batched: invalid batch description");
        return SEP_ERR_ARG;
    }
    if (B == 0) return SEP_OK;
    if (!lam || !Q || !lam_bar || !Q_bar || !d_bar || !e_bar) {
        set_error("sep_eig_backward_batched: null device pointer");
        return SEP_ERR_ARG;
    }
    std::vector<BwdProb> p((size_t)B);
    int nmax = 0;
    int64_t tot_d = 0, tot_m = 0;
    for (int64_t i = 0; i < B; ++i) {
        if (n[i] < 1 || n[i] > 65536 || off_d[i] < 0 || off_q[i] < 0) {
            set_error("sep_eig_backward_batched: problem %lld has order %lld or bad offsets",
                      (long long)i, (long long)n[i]);
            return SEP_ERR_ARG;
        }
        p[i] = BwdProb{(int)n[i], (int)n[i], off_q[i], off_d[i], tot_m};
        tot_m += n[i] * n[i];
        tot_d = std::max(tot_d, off_d[i] + n[i]);
        nmax = std::max(nmax, (int)n[i]);
    }
    return run_backward(h, p, nmax, tot_d, tot_m, lam, Q, lam_bar, Q_bar, d_bar, e_bar);
}

}  // namespace sep

// tridiag.hip -- dense symmetric -> tridiagonal reduction and the explicit orthogonal
// factor, the step before qr_algorithm! when the input is a dense symmetric matrix.
//
// The reference starts from a SymTridiagonal (its README states the scope); this file adds
// the reduction in front of it with LAPACK's storage conventions, so the results can be
// compared with dsytrd / dorgtr (uplo = 'L'):
//   sytrd: A = Q T Q^T, Q = H(0) H(1) ... H(n-2), H(i) = I - tau_i v_i v_i^T with
//          v_i[i+1] = 1 and v_i[i+2 .. n-1] stored in A[i+2 .., i]; d = diag(T),
//          e = subdiag(T), also left in A[i+1, i].  Only the lower triangle is referenced.
//   orgtr: the n x n matrix Q from those reflectors.
// Started from Q, qr_algorithm!(T, Q) accumulates the tridiagonal eigenvectors onto it,
// so its output columns are eigenvectors of the dense A.
//
// Reduction, column i (dsytd2's unblocked order): reflector of the updated column i, then
// y = tau A22 v, w = y - (tau/2)(y.v) v and the rank-2 update A22 -= v w^T + w v^T.  The
// update is applied lazily: the pass that computes y for column i first applies column
// i-1's pending update to each element it reads, so every column costs one read and one
// write of the trailing lower triangle (the HBM bound: 8 n^3 / 3 bytes in all).  Per column:
//   k_trd_tiles   one 64 x 64 lower tile per workgroup: update, store, and the tile's
//                 contributions A_IJ v_J (to row block I) and A_IJ^T v_I (to row block J)
//                 written to a partial slot of their own (slot = the other tile index);
//   k_trd_reduce  sums the slots in a fixed order (deterministic, no atomics); up to 16
//                 slots the next column step sums them itself;
//   k_trd_vec_r   one workgroup, the column in registers (k_trd_vec: passes over memory
//                 for columns beyond 8192 rows): w of the previous column, its update applied
//                 to column i, d[i], and the reflector of column i (dlarfg: beta =
//                 -sign(alpha) hypot(alpha, |x|), tau = (beta - alpha) / beta,
//                 v = x / (alpha - beta)).
// The explicit factor: below 256 rows accumulated backwards one reflector at a time
// (Q = H(0) (H(1) (... H(n-2)))), each reflector's rank-1 update applied in the pass that
// forms the next reflector's column dots z = v^T Q; from 256 rows in blocks of 64
// reflectors (compact WY: Q -= V T V^T Q with library GEMMs, the products over the block's
// rows split along those rows).
#include <rocblas/rocblas.h>

#include "sep_internal.hpp"

namespace sep {
namespace {

constexpr int kTile = 64;
constexpr int kThreads = 256;  // 4 waves; wave q owns 16 columns of the tile
constexpr int kVecThreads = 1024;

// Tile (I, J), I >= J, of the lower triangle from its linear index t = I (I + 1) / 2 + J.
__device__ __forceinline__ void lower_tile(int t, int& I, int& J)
{
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    I = r;
    J = t - r * (r + 1) / 2;
}

// Column i's trailing block A[s.., s..], s = i + 1, m = n - s.  vp / wp: the pending update
// of column i - 1 (absolute row indices; has_pending = 0 for i = 0); vn: column i's
// reflector.  part[slot * m + row] receives the tile contributions to y = A22 vn.
__global__ void __launch_bounds__(kThreads)
    k_trd_tiles(double* __restrict__ A, int64_t lda, int s, int m, const double* __restrict__ vp,
                const double* __restrict__ wp, int has_pending, const double* __restrict__ vn,
                double* __restrict__ part)
{
    __shared__ double prod[kTile][kTile + 1];  // [column][row] products a * vn[row]
    __shared__ double rows[4][kTile];
    __shared__ double cvec[3][kTile];  // the tile's columns of vp, wp, vn (0 past m)
    int I, J;
    lower_tile(blockIdx.x, I, J);
    const int lane = threadIdx.x % kTile, q = threadIdx.x / kTile;
    // the column vectors once per workgroup (waves 0-2, one load per lane), read back from
    // LDS as broadcasts, before the tile's own loads
    if (q < 3) {
        const int colv = J * kTile + lane;
        const double* __restrict__ src = q == 0 ? vp : (q == 1 ? wp : vn);
        cvec[q][lane] = (colv < m && (q == 2 || has_pending)) ? src[s + colv] : 0.0;
    }
    const int row = I * kTile + lane;
    const bool rin = row < m;
    const double vr = rin ? vn[s + row] : 0.0;
    const double vpr = (rin && has_pending) ? vp[s + row] : 0.0;
    const double wpr = (rin && has_pending) ? wp[s + row] : 0.0;
    double* __restrict__ colbase = A + (s + row);
    double a[16];
    int col0 = J * kTile + q * 16;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int col = col0 + c;
        const bool ok = rin && col < m && (I > J || row >= col);
        a[c] = ok ? colbase[(int64_t)(s + col) * lda] : 0.0;
    }
    lds_barrier();
    double rs = 0.0;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int col = col0 + c;
        const bool ok = rin && col < m && (I > J || row >= col);
        if (ok && has_pending) {
            a[c] -= vpr * cvec[1][q * 16 + c] + wpr * cvec[0][q * 16 + c];
            colbase[(int64_t)(s + col) * lda] = a[c];
        }
        const double vc = cvec[2][q * 16 + c];
        rs += a[c] * vc;
        // transposed contribution: strictly-lower elements only (the diagonal is counted
        // once, in the row sum)
        prod[q * 16 + c][lane] = (I > J || row > col) ? a[c] * vr : 0.0;
    }
    rows[q][lane] = rs;
    lds_barrier();
    // column sums: thread (cc, q) adds rows q*16 .. q*16+15 of column cc
    const int cc = threadIdx.x % kTile;
    double cs = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) cs += prod[cc][q * 16 + r];
    lds_barrier();
    prod[cc][q] = cs;  // reuse: column cc's four partials in prod[cc][0..3]
    lds_barrier();
    if (threadIdx.x < kTile) {
        const int x = threadIdx.x;
        const double rsum = ((rows[0][x] + rows[1][x]) + rows[2][x]) + rows[3][x];
        const double csum = ((prod[x][0] + prod[x][1]) + prod[x][2]) + prod[x][3];
        if (I == J) {
            if (I * kTile + x < m) part[(int64_t)I * m + I * kTile + x] = rsum + csum;
        } else {
            if (I * kTile + x < m) part[(int64_t)J * m + I * kTile + x] = rsum;
            if (J * kTile + x < m) part[(int64_t)I * m + J * kTile + x] = csum;
        }
    }
}

// y[s + r] = sum over slots of part[slot * m + r].  Workgroup of 64 rows x 4 slot
// groups: thread (row, q) adds slots q, q + 4, q + 8, ... (16 loads in flight at a time,
// index-clamped so that none is predicated), then the four group sums are added in a fixed
// order (deterministic, no atomics).
constexpr int kRedRows = 64, kRedGroups = 4, kRedBatch = 16;
__global__ void __launch_bounds__(kRedRows * kRedGroups)
    k_trd_reduce(const double* __restrict__ part, int nslots, int s, int m, double* __restrict__ y)
{
    __shared__ double grp[kRedGroups][kRedRows];
    const int lane = threadIdx.x % kRedRows, q = threadIdx.x / kRedRows;
    const int r = blockIdx.x * kRedRows + lane;
    const int rc = r < m ? r : m - 1;
    double acc = 0.0;
    for (int t0 = 0; t0 < nslots; t0 += kRedGroups * kRedBatch) {
        double p[kRedBatch];
#pragma unroll
        for (int u = 0; u < kRedBatch; ++u) {
            const int t = t0 + q + kRedGroups * u;
            p[u] = part[(int64_t)(t < nslots ? t : nslots - 1) * m + rc];
        }
#pragma unroll
        for (int u = 0; u < kRedBatch; ++u)
            if (t0 + q + kRedGroups * u < nslots) acc += p[u];
    }
    grp[q][lane] = acc;
    lds_barrier();
    if (q == 0 && r < m) y[s + r] = ((grp[0][lane] + grp[1][lane]) + grp[2][lane]) + grp[3][lane];
}

__device__ double block_sum(double v, double* red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x / 64); ++k) t += red[k];
    return t;
}
__device__ double block_max(double v, double* red)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x / 64); ++k) t = fmax(t, red[k]);
    return t;
}

constexpr int kFuseSlots = 16;     // at most this many partial slots: summed in k_trd_vec

// y[r] of the previous column: the reduce kernel's output, or (part != nullptr) the sum of
// the tile partials in the same fixed slot order.
__device__ __forceinline__ double y_at(const double* __restrict__ y,
                                       const double* __restrict__ part, int nslots, int m,
                                       int s, int r)
{
    if (!part) return y[r];
    double acc = 0.0;
    for (int t = 0; t < nslots; ++t) acc += part[(int64_t)t * m + (r - s)];
    return acc;
}

// The reflector H = I - tau v v^T with H (alpha, x) = (beta, 0), v = (1, x / (alpha - beta)),
// as LAPACK's dlarfg: beta = -sign(alpha) hypot(alpha, |x|), tau = (beta - alpha) / beta.
// Like dlarfg, a |beta| below safmin = 2^-969 (x and alpha near the underflow threshold,
// where 1 / (alpha - beta) would lose precision or overflow) is recomputed from alpha and x
// scaled by 2^969, up to 20 times; v is formed from the scaled x and beta scaled back.
// |x| = unit sqrt(ssn) (unit a power of two, ssn = sum (x_j / unit)^2) scales exactly with x.
struct Refl {
    double beta, tau, sc;
    int knt;
    __device__ double scaled(double x) const
    {
        for (int t = 0; t < knt; ++t) x *= 0x1p969;
        return x * sc;
    }
};
__device__ __forceinline__ Refl reflector(double alpha, double xnorm, double unit, double ssn)
{
    Refl o;
    double beta = -copysign(hypot(alpha, xnorm), alpha);
    o.knt = 0;
    if (fabs(beta) < 0x1p-969) {
        double a = alpha, m = unit;
        do {
            ++o.knt;
            beta *= 0x1p969;
            a *= 0x1p969;
            m *= 0x1p969;
        } while (fabs(beta) < 0x1p-969 && o.knt < 20);
        beta = -copysign(hypot(a, m * sqrt(ssn)), a);
        alpha = a;
    }
    o.tau = (beta - alpha) / beta;
    o.sc = 1.0 / (alpha - beta);
    for (int t = 0; t < o.knt; ++t) beta *= 0x1p-969;
    o.beta = beta;
    return o;
}

// ---- one-workgroup column step, rows in registers ---------------------------------------
// 256 threads (one wave per SIMD).  Wave reductions run on DPP lane moves inside each
// 16-lane row (xor 1, xor 2, half-row mirror, row mirror: every lane of a row then holds
// the same sum, bit for bit, as IEEE addition commutes) and four lane reads across rows;
// the four waves' results meet in LDS behind one barrier.
constexpr int kVecR = 256;

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int L>
__device__ __forceinline__ double lane_value(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, L);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), L);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    v += dpp_move<0x140>(v);  // row_mirror
    return (lane_value<0>(v) + lane_value<16>(v)) + (lane_value<32>(v) + lane_value<48>(v));
}
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_move<0xB1>(v));
    v = fmax(v, dpp_move<0x4E>(v));
    v = fmax(v, dpp_move<0x141>(v));
    v = fmax(v, dpp_move<0x140>(v));
    return fmax(fmax(lane_value<0>(v), lane_value<16>(v)),
                fmax(lane_value<32>(v), lane_value<48>(v)));
}
// Workgroup sum over the four waves; `red` (4 doubles) must not be reused by a later
// reduction of the same launch (no second barrier guards it).
__device__ __forceinline__ double wg_sum4(double v, double* red)
{
    v = wave_sum(v);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
    lds_barrier();
    return (red[0] + red[1]) + (red[2] + red[3]);
}
__device__ __forceinline__ void wg_max_sum4(double& mx, double& sq, double* rmx, double* rsq)
{
    mx = wave_max(mx);
    sq = wave_sum(sq);
    if (threadIdx.x % 64 == 0) {
        rmx[threadIdx.x / 64] = mx;
        rsq[threadIdx.x / 64] = sq;
    }
    lds_barrier();
    mx = fmax(fmax(rmx[0], rmx[1]), fmax(rmx[2], rmx[3]));
    sq = (rsq[0] + rsq[1]) + (rsq[2] + rsq[3]);
}

// y[r] of the previous column from its tile partials (at most kFuseSlots slots, summed in
// slot order); every load is issued unpredicated (slot and row clamped into range) so that
// they are all in flight together.
__device__ __forceinline__ double y_from_slots(const double* __restrict__ part, int nslots,
                                               int m, int row)
{
    double p[kFuseSlots];
#pragma unroll
    for (int t = 0; t < kFuseSlots; ++t)
        p[t] = part[(int64_t)(t < nslots ? t : nslots - 1) * m + row];
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < kFuseSlots; ++t)
        if (t < nslots) acc += p[t];
    return acc;
}

// The previous column's update of one element: w = tp y + aw v, c - (v wi + w vi).
__device__ __forceinline__ double pending_w(double tp, double aw, double yr, double vr)
{
    return tp * yr + aw * vr;
}

// Column i's step (as k_trd_vec) with rows [i, n) in registers, KP per thread:
// m = n - i <= kVecR * KP.  The values every thread needs (y, v and A at rows i and i + 1,
// tau) are read by all threads, so the only barriers are those of the reductions: the dot
// product of w, then the largest magnitude and sum of squares of x.
template <int KP>
__global__ void __launch_bounds__(kVecR)
    k_trd_vec_r(double* __restrict__ A, int64_t lda, int n, int i, const double* __restrict__ vp,
                double* __restrict__ wp, const double* __restrict__ y,
                const double* __restrict__ part, int nslots, double* __restrict__ vn,
                double* __restrict__ d, double* __restrict__ e, double* __restrict__ tau)
{
    __shared__ double red[4][4];
    double* __restrict__ col = A + (int64_t)i * lda;
    const int m = n - i;  // rows of the previous column's y: [i, n), index r - i
    const int tid = threadIdx.x;
    const bool has_next = i + 1 < n;
    double cv[KP], yv[KP], vv[KP];
    double yi = 0.0, vi = 0.0, y1 = 0.0, v1 = 0.0, tp = 0.0;
    double c1 = has_next ? col[i + 1] : 0.0;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int r = i + tid + k * kVecR;
        cv[k] = r < n ? col[r] : 0.0;
    }
    if (i > 0) {
        tp = tau[i - 1];
        vi = vp[i];
        if (has_next) v1 = vp[i + 1];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int r = i + tid + k * kVecR;
            vv[k] = r < n ? vp[r] : 0.0;
        }
        if (part) {
            yi = y_from_slots(part, nslots, m, 0);
            if (has_next) y1 = y_from_slots(part, nslots, m, 1);
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const int r = i + tid + k * kVecR;
                yv[k] = r < n ? y_from_slots(part, nslots, m, r - i) : 0.0;
            }
        } else {
            yi = y[i];
            if (has_next) y1 = y[i + 1];
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const int r = i + tid + k * kVecR;
                yv[k] = r < n ? y[r] : 0.0;
            }
        }
        // x = tp y, w = x - (tp/2)(x.v) v over rows [i, n)
        double dot = 0.0;
#pragma unroll
        for (int k = 0; k < KP; ++k) dot += (tp * yv[k]) * vv[k];
        dot = wg_sum4(dot, red[0]);
        const double aw = -0.5 * tp * dot;
        const double wi = pending_w(tp, aw, yi, vi);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int r = i + tid + k * kVecR;
            const double w = pending_w(tp, aw, yv[k], vv[k]);
            if (r < n) wp[r] = w;
            cv[k] -= vv[k] * wi + w * vi;
        }
        if (has_next) c1 -= v1 * wi + pending_w(tp, aw, y1, v1) * vi;
    }
    if (tid == 0) {
        d[i] = cv[0];
        col[i] = cv[0];
    }
    if (!has_next) return;
    // dlarfg on alpha = A[i+1, i], x = rows [i+2, n)
    const double alpha = c1;
    double mx = 0.0, sq = 0.0;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int r = i + tid + k * kVecR;
        if (r >= i + 2 && r < n) {
            mx = fmax(mx, fabs(cv[k]));
            sq += cv[k] * cv[k];
        }
    }
    wg_max_sum4(mx, sq, red[1], red[2]);
    // |x| = 2^ex sqrt(ssn) with 2^ex <= mx < 2^(ex+1) when squares could leave the range
    double xnorm, ssn = 1.0, unit = 1.0;
    if (mx == 0.0 || (mx > 0x1p-400 && mx < 0x1p400)) {
        xnorm = sqrt(sq);
    } else {
        const int ex = ilogb(mx);
        double ss = 0.0;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int r = i + tid + k * kVecR;
            const double t = (r >= i + 2 && r < n) ? ldexp(cv[k], -ex) : 0.0;
            ss += t * t;
        }
        ssn = wg_sum4(ss, red[3]);
        unit = ldexp(1.0, ex);
        xnorm = ldexp(sqrt(ssn), ex);
    }
    if (xnorm == 0.0) {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int r = i + tid + k * kVecR;
            if (r >= i + 2 && r < n) {
                col[r] = cv[k];
                vn[r] = 0.0;
            }
        }
        if (tid == 0) {
            col[i + 1] = alpha;
            vn[i + 1] = 1.0;
            e[i] = alpha;
            tau[i] = 0.0;
        }
        return;
    }
    const Refl rf = reflector(alpha, xnorm, unit, ssn);
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int r = i + tid + k * kVecR;
        if (r >= i + 2 && r < n) {
            const double v = rf.scaled(cv[k]);
            col[r] = v;
            vn[r] = v;
        }
    }
    if (tid == 0) {
        vn[i + 1] = 1.0;
        col[i + 1] = rf.beta;
        e[i] = rf.beta;
        tau[i] = rf.tau;
    }
}

// Column i: finish the previous column's w (from y = A22 v of column i - 1), apply its
// update to column i, write d[i] and generate column i's reflector into A, vn, e, tau.
// Columns too long for k_trd_vec_r: each step is a pass over memory.
__global__ void __launch_bounds__(kVecThreads)
    k_trd_vec(double* __restrict__ A, int64_t lda, int n, int i, const double* __restrict__ vp,
              double* __restrict__ wp, const double* __restrict__ y,
              const double* __restrict__ part, int nslots, double* __restrict__ vn,
              double* __restrict__ d, double* __restrict__ e, double* __restrict__ tau)
{
    __shared__ double red[kVecThreads / 64];
    double* __restrict__ col = A + (int64_t)i * lda;
    const int s = i, m = n - i;  // rows of the previous column's y: [i, n)
    // long columns: the same steps, one pass over memory each
    if (i > 0) {
        const double tp = tau[i - 1];
        double dot = 0.0;
        for (int r = i + threadIdx.x; r < n; r += blockDim.x)
            dot += (tp * y_at(y, part, nslots, m, s, r)) * vp[r];
        dot = block_sum(dot, red);
        const double alpha = -0.5 * tp * dot;
        for (int r = i + threadIdx.x; r < n; r += blockDim.x)
            wp[r] = tp * y_at(y, part, nslots, m, s, r) + alpha * vp[r];
        __syncthreads();
        const double vi = vp[i], wi = wp[i];
        for (int r = i + threadIdx.x; r < n; r += blockDim.x)
            col[r] -= vp[r] * wi + wp[r] * vi;
        __syncthreads();
    }
    if (threadIdx.x == 0) d[i] = col[i];
    if (i >= n - 1) return;
    const double alpha = col[i + 1];
    double mx = 0.0;
    for (int r = i + 2 + threadIdx.x; r < n; r += blockDim.x) mx = fmax(mx, fabs(col[r]));
    mx = block_max(mx, red);
    // |x| = 2^ex sqrt(ss), ss the sum of squares of x / 2^ex, 2^ex <= mx < 2^(ex+1)
    double ss = 0.0;
    const int ex = mx > 0.0 ? ilogb(mx) : 0;
    if (mx > 0.0) {
        for (int r = i + 2 + threadIdx.x; r < n; r += blockDim.x) {
            const double t = ldexp(col[r], -ex);
            ss += t * t;
        }
    }
    ss = block_sum(ss, red);
    const double xnorm = mx > 0.0 ? ldexp(sqrt(ss), ex) : 0.0;
    if (xnorm == 0.0) {
        for (int r = i + 1 + threadIdx.x; r < n; r += blockDim.x) vn[r] = r == i + 1 ? 1.0 : 0.0;
        if (threadIdx.x == 0) {
            e[i] = alpha;
            tau[i] = 0.0;
        }
        return;
    }
    const Refl rf = reflector(alpha, xnorm, ldexp(1.0, ex), ss);
    __syncthreads();  // every thread has read col[i + 1] and x before they change
    for (int r = i + 2 + threadIdx.x; r < n; r += blockDim.x) {
        const double v = rf.scaled(col[r]);
        col[r] = v;
        vn[r] = v;
    }
    if (threadIdx.x == 0) {
        vn[i + 1] = 1.0;
        col[i + 1] = rf.beta;
        e[i] = rf.beta;
        tau[i] = rf.tau;
    }
}

// Q <- I (n x n, ld ldq).
__global__ void k_eye(double* __restrict__ Q, int64_t ldq, int n)
{
    const int64_t tot = (int64_t)n * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = q / n, r = q % n;
        Q[r + c * ldq] = r == c ? 1.0 : 0.0;
    }
}

// Reflector j's column of A as v_j[r] (r >= j + 1; v_j[j + 1] = 1).
__device__ __forceinline__ double refl(const double* __restrict__ A, int64_t lda, int j, int r)
{
    return r == j + 1 ? 1.0 : A[r + (int64_t)j * lda];
}

// Backward accumulation, step for reflector i (i = -1: the final pass).  Region
// [s, n)^2 with s = i + 2: applies reflector i+1's pending update
// Q[r, c] -= tau_{i+1} v_{i+1}[r] z[c] and, for i >= 0, writes the column partials of
// v_i^T Q over the region's rows (one slot per row tile) to part.
__global__ void __launch_bounds__(kThreads)
    k_org_tiles(double* __restrict__ Q, int64_t ldq, const double* __restrict__ A, int64_t lda,
                const double* __restrict__ tau, int i, int s, int m, int ntc,
                const double* __restrict__ zp, double* __restrict__ part)
{
    __shared__ double prod[kTile][kTile + 1];
    const int I = blockIdx.x / ntc, J = blockIdx.x % ntc;
    const int lane = threadIdx.x % kTile, q = threadIdx.x / kTile;
    const int row = I * kTile + lane;
    const bool rin = row < m;
    const int p = i + 1;  // pending reflector
    const double tp = tau[p];
    const double vpr = rin ? refl(A, lda, p, s + row) : 0.0;
    const double vnr = (rin && i >= 0) ? refl(A, lda, i, s + row) : 0.0;
    double* __restrict__ base = Q + (s + row);
    const int col0 = J * kTile + q * 16;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int col = col0 + c;
        double a = 0.0;
        if (rin && col < m) {
            double* pq = base + (int64_t)(s + col) * ldq;
            a = *pq - tp * vpr * zp[s + col];
            *pq = a;
        }
        prod[q * 16 + c][lane] = a * vnr;
    }
    if (i < 0) return;
    lds_barrier();
    const int cc = threadIdx.x % kTile;
    double cs = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) cs += prod[cc][q * 16 + r];
    lds_barrier();
    prod[cc][q] = cs;
    lds_barrier();
    if (threadIdx.x < kTile) {
        const int x = threadIdx.x, col = J * kTile + x;
        if (col < m)
            part[(int64_t)I * m + col] = ((prod[x][0] + prod[x][1]) + prod[x][2]) + prod[x][3];
    }
}

// z_i over columns [i+1, n): z[i+1] = 1 (row i+1 of Q is still e_{i+1}), z[c] = sum of the
// row-tile partials for c >= i+2.
__global__ void __launch_bounds__(256)
    k_org_reduce(const double* __restrict__ part, int nslots, int i, int s, int m,
                 double* __restrict__ z)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) z[i + 1] = 1.0;
    if (r >= m) return;
    double acc = 0.0;
#pragma unroll 8
    for (int t = 0; t < nslots; ++t) acc += part[(int64_t)t * m + r];
    z[s + r] = acc;
}

// Blocked accumulation of the explicit factor (compact WY, LAPACK dlarft / dlarfb):
// reflectors j0 .. j0+nb-1 form H(j0) ... H(j0+nb-1) = I - V T V^T, applied to the trailing
// block Q[j0+1.., j0+1..] with three library GEMMs (W = V^T Q, W = T W, Q -= V W).
constexpr int kNB = 64;           // reflectors per block
constexpr int kBlockedMin = 256;  // smaller orders take the unblocked path

// V (m x nb, ld m) for the block starting at reflector j0: rows [j0+1, n), column c holds
// v_{j0+c} (zero above its unit entry).
__global__ void __launch_bounds__(256)
    k_org_vblock(const double* __restrict__ A, int64_t lda, int n, int j0, int nb,
                 double* __restrict__ V)
{
    const int m = n - j0 - 1;
    const int64_t tot = (int64_t)m * nb;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(q % m), c = (int)(q / m);
        const int row = j0 + 1 + r, j = j0 + c;
        V[q] = row < j + 1 ? 0.0 : row == j + 1 ? 1.0 : A[row + (int64_t)j * lda];
    }
}

// T (nb x nb, ld nb, upper) from S = V^T V and tau, column by column as dlarft
// (forward, columnwise): T[j, j] = tau_j, T[0:j, j] = T[0:j, 0:j] (-tau_j S[0:j, j]).
__global__ void __launch_bounds__(kNB)
    k_org_tfactor(const double* __restrict__ S, const double* __restrict__ tau, int j0, int nb,
                  double* __restrict__ T)
{
    // Thread i owns row i of T in registers (T(i, k) = 0 for k < i, so the sums run over
    // every k < j); S and tau are read at uniform addresses (scalar loads, SGPR operands).
    // No thread reads another's row, so there is no barrier.
    const int i = threadIdx.x;
    double tr[kNB];
#pragma unroll
    for (int j = 0; j < kNB; ++j) {
        tr[j] = 0.0;
        if (j < nb) {
            const double* __restrict__ sj = S + (int64_t)j * nb;
            const double tj = tau[j0 + j];
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < j; ++k) acc[k % 4] += tr[k] * sj[k];
            const double dot = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            tr[j] = i < j ? -tj * dot : (i == j ? tj : 0.0);
        }
    }
#pragma unroll
    for (int j = 0; j < kNB; ++j)
        if (j < nb && i < nb) T[i + j * nb] = tr[j];
}

int blas_status(rocblas_status st, const char* what)
{
    if (st == rocblas_status_success) return SEP_OK;
    set_error("sep_orgtr: %s -> rocBLAS status %d", what, (int)st);
    return SEP_ERR_HIP;
}

int check_args(const char* who, int64_t n, int64_t lda, const void* A)
{
    // leading dimensions reach the blocked path's library GEMMs as 32-bit integers
    if (n < 0 || n > 65536 || lda < std::max<int64_t>(n, 1) || lda > INT32_MAX) {
        set_error("%s: invalid n=%lld ld=%lld", who, (long long)n, (long long)lda);
        return SEP_ERR_ARG;
    }
    if (n > 0 && !A) {
        set_error("%s: null device pointer", who);
        return SEP_ERR_ARG;
    }
    return SEP_OK;
}

}  // namespace

namespace {

// Q <- H(0) ... H(n-2) (Q = I on entry), blocks of kNB reflectors from the last one.
// out = sum over p of part[p * len + e], e < len (fixed order).
__global__ void __launch_bounds__(256)
    k_sum_parts(const double* __restrict__ part, int np, int64_t len, double* __restrict__ out)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len;
         e += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int p = 0; p < np; ++p) acc += part[(int64_t)p * len + e];
        out[e] = acc;
    }
}

// Split-K factor for the products that reduce over the m rows of the block (V^T V, V^T Q):
// their outputs are only b x b and b x m, too few tiles to fill the chip, so the m rows are
// cut into np chunks of at least kSplitRows, multiplied as one strided batch and summed.
constexpr int kSplitMax = 16, kSplitRows = 256;

// out (b x ncols, ld b) = V^T B over m rows, V (m x b, ld m), B (m x ncols, ld ldb).
int gemm_tn_split(rocblas_handle bh, hipStream_t st, int b, int ncols, int m, const double* V,
                  const double* B, int64_t ldb, double* out, double* part)
{
    const double one = 1.0, zero = 0.0;
    const int np = std::max(1, std::min(kSplitMax, m / kSplitRows));
    if (np == 1)
        return blas_status(rocblas_dgemm(bh, rocblas_operation_transpose, rocblas_operation_none,
                                         b, ncols, m, &one, V, m, B, (rocblas_int)ldb, &zero, out,
                                         b),
                           "V^T B");
    const int rc = m / np, rem = m - np * rc;
    const rocblas_stride sc = (rocblas_stride)b * ncols;
    if (int rc_ = blas_status(rocblas_dgemm_strided_batched(
                                  bh, rocblas_operation_transpose, rocblas_operation_none, b, ncols,
                                  rc, &one, V, m, rc, B, (rocblas_int)ldb, rc, &zero, part, b, sc,
                                  np),
                              "V^T B (split)"))
        return rc_;
    if (rem > 0) {
        // the last rows' product joins chunk 0
        if (int rc_ = blas_status(rocblas_dgemm(bh, rocblas_operation_transpose,
                                                rocblas_operation_none, b, ncols, rem, &one,
                                                V + (int64_t)np * rc, m, B + (int64_t)np * rc,
                                                (rocblas_int)ldb, &one, part, b),
                                  "V^T B (rows past the chunks)"))
            return rc_;
    }
    const int64_t len = (int64_t)b * ncols;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)std::min<int64_t>(2048, (len + 255) / 256)),
                       dim3(256), 0, st, part, np, len, out);
    SEP_CHECK_LAUNCH();
    return SEP_OK;
}

int orgtr_blocked(sep_handle* h, int n, const double* A, int64_t lda, const double* tau,
                  double* Q, int64_t ldq)
{
    hipStream_t st = h->stream;
    if (!h->blas) {
        rocblas_handle b = nullptr;
        if (int rc = blas_status(rocblas_create_handle(&b), "create handle")) return rc;
        h->blas = b;
    }
    rocblas_handle bh = static_cast<rocblas_handle>(h->blas);
    if (int rc = blas_status(rocblas_set_stream(bh, st), "set stream")) return rc;
    const size_t nb = kNB;
    // V (n x nb), W and W2 (nb x n), S and T (nb x nb), split-K partials (kSplitMax x nb x n)
    if (int rc = h->trd_blk.reserve(sizeof(double) *
                                    ((3 + kSplitMax) * nb * (size_t)n + 2 * nb * nb)))
        return rc;
    double* V = h->trd_blk.as<double>();
    double* W = V + nb * (size_t)n;
    double* W2 = W + nb * (size_t)n;
    double* S = W2 + nb * (size_t)n;
    double* T = S + nb * nb;
    double* part = T + nb * nb;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    for (int jend = n - 1; jend > 0;) {
        const int j0 = std::max(0, jend - kNB), b = jend - j0, m = n - j0 - 1;
        hipLaunchKernelGGL(k_org_vblock, dim3((unsigned)std::min<int64_t>(2048, ((int64_t)m * b + 255) / 256)),
                           dim3(256), 0, st, A, lda, n, j0, b, V);
        SEP_CHECK_LAUNCH();
        if (int rc = gemm_tn_split(bh, st, b, b, m, V, V, m, S, part)) return rc;
        hipLaunchKernelGGL(k_org_tfactor, dim3(1), dim3(kNB), 0, st, S, tau, j0, b, T);
        SEP_CHECK_LAUNCH();
        double* Qr = Q + (j0 + 1) + (int64_t)(j0 + 1) * ldq;
        if (int rc = gemm_tn_split(bh, st, b, m, m, V, Qr, ldq, W, part)) return rc;
        if (int rc = blas_status(rocblas_dgemm(bh, rocblas_operation_none, rocblas_operation_none,
                                               b, m, b, &one, T, b, W, b, &zero, W2, b),
                                 "T W"))
            return rc;
        if (int rc = blas_status(rocblas_dgemm(bh, rocblas_operation_none, rocblas_operation_none,
                                               m, m, b, &mone, V, m, W2, b, &one, Qr,
                                               (rocblas_int)ldq),
                                 "Q - V W"))
            return rc;
        jend = j0;
    }
    return SEP_OK;
}

}  // namespace

void release_blas(sep_handle* h)
{
    if (h->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(h->blas));
    h->blas = nullptr;
}

// Column i's one-workgroup step: rows in registers (k_trd_vec_r, the smallest KP that
// holds the n - i rows) up to 8192 rows, passes over memory beyond.
int launch_vec(hipStream_t st, double* A, int64_t lda, int n, int i, const double* vp, double* wp,
               const double* y, const double* part, int nslots, double* vn, double* d, double* e,
               double* tau)
{
    const int m = n - i;
    if (m <= kVecR * 4)
        hipLaunchKernelGGL(k_trd_vec_r<4>, dim3(1), dim3(kVecR), 0, st, A, lda, n, i, vp, wp, y,
                           part, nslots, vn, d, e, tau);
    else if (m <= kVecR * 8)
        hipLaunchKernelGGL(k_trd_vec_r<8>, dim3(1), dim3(kVecR), 0, st, A, lda, n, i, vp, wp, y,
                           part, nslots, vn, d, e, tau);
    else if (m <= kVecR * 16)
        hipLaunchKernelGGL(k_trd_vec_r<16>, dim3(1), dim3(kVecR), 0, st, A, lda, n, i, vp, wp, y,
                           part, nslots, vn, d, e, tau);
    else if (m <= kVecR * 32)
        hipLaunchKernelGGL(k_trd_vec_r<32>, dim3(1), dim3(kVecR), 0, st, A, lda, n, i, vp, wp, y,
                           part, nslots, vn, d, e, tau);
    else
        hipLaunchKernelGGL(k_trd_vec, dim3(1), dim3(kVecThreads), 0, st, A, lda, n, i, vp, wp, y,
                           part, nslots, vn, d, e, tau);
    SEP_CHECK_LAUNCH();
    return SEP_OK;
}

int sytrd_f64(sep_handle* h, int64_t n64, double* A, int64_t lda, double* d, double* e,
              double* tau)
{
    if (int rc = check_args("sep_sytrd", n64, lda, A)) return rc;
    if (n64 == 0) return SEP_OK;
    if (!d || (n64 > 1 && (!e || !tau))) {
        set_error("sep_sytrd: null device pointer");
        return SEP_ERR_ARG;
    }
    const int n = (int)n64;
    hipStream_t st = h->stream;
    if (n == 1) {
        SEP_HIP(hipMemcpyAsync(d, A, sizeof(double), hipMemcpyDeviceToDevice, st));
        return SEP_OK;
    }
    const int ntmax = (n - 1 + kTile - 1) / kTile;
    // vectors: v[2][n], w[n], y[n]
    if (int rc = h->trd_vec.reserve(sizeof(double) * 4 * (size_t)n)) return rc;
    if (int rc = h->trd_part.reserve(sizeof(double) * (size_t)ntmax * (size_t)(n - 1))) return rc;
    double* v[2] = {h->trd_vec.as<double>(), h->trd_vec.as<double>() + n};
    double* w = h->trd_vec.as<double>() + 2 * (size_t)n;
    double* y = h->trd_vec.as<double>() + 3 * (size_t)n;
    double* part = h->trd_part.as<double>();
    // partials of the previous column's y when few enough to be summed in k_trd_vec
    const double* fpart = nullptr;
    int fslots = 0;
    for (int i = 0; i < n; ++i) {
        // column i: previous column's w and update, d[i], reflector i (into v[i & 1])
        if (int rc = launch_vec(st, A, lda, n, i, v[(i + 1) & 1], w, y, fpart, fslots, v[i & 1],
                                d, e, tau))
            return rc;
        if (i >= n - 1) break;
        const int s = i + 1, m = n - s, nt = (m + kTile - 1) / kTile;
        hipLaunchKernelGGL(k_trd_tiles, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(kThreads), 0,
                           st, A, lda, s, m, v[(i + 1) & 1], w, i > 0 ? 1 : 0, v[i & 1], part);
        SEP_CHECK_LAUNCH();
        if (nt <= kFuseSlots) {
            fpart = part;
            fslots = nt;
        } else {
            fpart = nullptr;
            fslots = 0;
            hipLaunchKernelGGL(k_trd_reduce, dim3((unsigned)((m + kRedRows - 1) / kRedRows)),
                               dim3(kRedRows * kRedGroups), 0, st, part, nt, s, m, y);
            SEP_CHECK_LAUNCH();
        }
    }
    return SEP_OK;
}

int orgtr_f64(sep_handle* h, int64_t n64, const double* A, int64_t lda, const double* tau,
              double* Q, int64_t ldq)
{
    if (int rc = check_args("sep_orgtr", n64, lda, A)) return rc;
    if (int rc = check_args("sep_orgtr", n64, ldq, Q)) return rc;
    if (n64 > 1 && !tau) {
        set_error("sep_orgtr: null device pointer");
        return SEP_ERR_ARG;
    }
    if (n64 == 0) return SEP_OK;
    const int n = (int)n64;
    hipStream_t st = h->stream;
    hipLaunchKernelGGL(k_eye, dim3((unsigned)std::min<int64_t>(4096, ((int64_t)n * n + 255) / 256)),
                       dim3(256), 0, st, Q, ldq, n);
    SEP_CHECK_LAUNCH();
    if (n == 1) return SEP_OK;
    if (n >= kBlockedMin) return orgtr_blocked(h, n, A, lda, tau, Q, ldq);
    const int ntmax = (n - 1 + kTile - 1) / kTile;
    if (int rc = h->trd_vec.reserve(sizeof(double) * 4 * (size_t)n)) return rc;
    if (int rc = h->trd_part.reserve(sizeof(double) * (size_t)ntmax * (size_t)(n - 1))) return rc;
    double* z[2] = {h->trd_vec.as<double>(), h->trd_vec.as<double>() + n};
    double* part = h->trd_part.as<double>();
    // reflector n-2 first: its column dots over rows [n-1, n) are just z[n-1] = 1
    hipLaunchKernelGGL(k_org_reduce, dim3(1), dim3(256), 0, st, part, 0, n - 2, n, 0,
                       z[(n - 2) & 1]);
    SEP_CHECK_LAUNCH();
    for (int i = n - 3; i >= -1; --i) {
        // region [i+2, n): pending reflector i+1 (dots in z[(i+1) & 1]); dots of reflector i
        const int s = i + 2, m = n - s, nt = (m + kTile - 1) / kTile;
        hipLaunchKernelGGL(k_org_tiles, dim3((unsigned)(nt * nt)), dim3(kThreads), 0, st, Q,
                           ldq, A, lda, tau, i, s, m, nt, z[(i + 1) & 1], part);
        SEP_CHECK_LAUNCH();
        if (i < 0) break;
        hipLaunchKernelGGL(k_org_reduce, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                           part, nt, i, s, m, z[i & 1]);
        SEP_CHECK_LAUNCH();
    }
    return SEP_OK;
}

}  // namespace sep

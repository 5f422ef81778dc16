// apply_dpp.hip -- bulk Givens application (Float64 and Float32) with DPP-broadcast
// coefficients.
//
// Same product as bulk_wave_order_2x2_rmul! (src/wave.jl:62-162): A <- A * prod_L prod_i
// G(i,L)^T, each rotation a' = c a + s b, b' = c b - s a on the column pair (i, i+1)
// (Rotation2 rmul!, src/simple_givens.jl:53-64), and the same systolic organisation as the
// systolic organisation: one row per lane, G waves of KG layers each, columns handed from
// wave to wave through a double-buffered LDS ring, wave order t = i + 2L.
//
// How the wave-uniform rotation data reach the FMA units:
//  * Coefficients are not scalar-loaded (scalar loads return out of order, so every use
//    waits for all of them, exposing L2 latency each step).  Instead one 64-bit VGPR holds
//    16 coefficients (2*KG per step, so one or two steps) in lanes 0..15 of each 16-lane
//    row, loaded with ordinary vector loads one whole interval ahead, and every FMA takes
//    its coefficient through the DPP64 `row_newbcast:j` source modifier, which broadcasts
//    lane j of each row.
//  * FP64 FMA issue rate depends on how many waves share a SIMD, hardly on the operand
//    source: register-only streams of the three-shear pattern (eight chains of three
//    dependent FMAs) issue an FMA in ~63% of the cycles with one wave per SIMD, ~83%
//    with two and ~93% with four (row_newbcast operand; a plain VGPR or SGPR coefficient
//    and the VGPR bank of the operands change this by 1-3%), at 2.24-2.42 GHz on all
//    CUs.  The 64-row geometry for large matrices is "wide": G = 8 waves of KG = 8
//    layers, S = 16 steps per barrier interval, two waves per SIMD ("small", KG = 4 and
//    S = 8, serves grids of at most 32 layers).  Four waves per SIMD would need 4 layers
//    per wave, which synchronises twice as often and moves twice the LDS traffic per FMA;
//    that geometry measured slower and is not built.  The first wave prefetches Q two or
//    three intervals ahead.  Below 64 * 256 rows these leave CUs idle; the 16- and 32-row
//    geometries further down (one layer group per DPP row, several workgroups per CU)
//    take over.
//  * Within an interval each wave lowers its priority as it progresses, so the younger of
//    two waves sharing a SIMD is not starved until the older one waits at the barrier,
//    and reads its LDS inputs and coefficient rows two steps ahead instead of all at the
//    interval start (where the whole workgroup's reads queued behind each other).
//  * Q is loaded and stored through one buffer resource per column (built with scalar
//    instructions): the hardware range check returns 0 for columns past n and drops
//    stores outside the matrix, so every wave runs every interval on one code path with
//    no per-lane address arithmetic.
//  * Only the middle waves load coefficients: each its own rows straight into registers
//    (one row per step, for the next interval), and the first and last waves' rows one
//    interval ahead, staged through LDS.  The first wave issues only Q loads and the last
//    only Q stores, so no wave ever has to wait for a memory operation it issued for a
//    later interval.  (Staging every wave's rows through LDS costs 2% at m = n = 16384,
//    k = 64: the extra LDS reads raise the power draw that sets the clock.)
//  * Rotations are applied as three shears, a += p b; b += q a; a += p b: three in-place
//    FMAs per rotation and row instead of two multiplies and two FMAs.  For a rotation
//    (c^2 + s^2 = 1) p = s~/(1 + |c|) = tan(phi/2), q = -s~, with the sign of c tracked
//    per column; in general a real scale per column is tracked (k_colscale), which takes
//    each rotation's sign(c) sqrt(c^2 + s^2), so pairs that are near normalised (all the
//    QR chase produces, the benchmark's cos/sin pairs, and pairs within 1/64 of the unit
//    circle) are applied as given, to rounding.  The systolic kernel multiplies each
//    column by its final scale when it enters the pass (first wave), so it leaves with
//    scale one; the 16- and 32-row kernels apply the scale when the column leaves.
//  * Grids with pairs farther from the unit circle (the reference benchmark's
//    (rand, rand)) take a five-instruction path, t = a; a += (c-1) a; a += s b;
//    b += (c-1) b; b -= s t, chosen on the device by a check of the grid, so the entry
//    point accepts any (c, s) as the reference's kernels do.
#include <mutex>
#include <type_traits>

#include "sep_internal.hpp"
#include "shear.hpp"

namespace sep {
namespace {

constexpr int kLanes = 64;

// Normalisation flag of a pass: 1 if the grid needs the five-instruction path (some
// |c^2 + s^2 - 1| above 1/64 among the pass's pairs, or a c or s that is not finite); every
// other grid runs the scaled three-shear form (see k_colscale), which applies the given
// (c, s) to rounding: normalised pairs, and pairs that are only near normalised
// (c^2 + s^2 = 1 + delta), alike.  k_colscale evaluates it on the pairs it loads (the residual
// as fma(c, c, fma(s, s, -1)), accurate to a fraction of eps) and publishes it from its last
// workgroup: each workgroup ORs its verdict into sync[0] and takes a ticket from sync[1]; the
// one with the last ticket moves sync[0] into the flag and sets both words back to zero, so
// every launch leaves them as it found them (zero) and no memset is needed between calls
// (graph capture records none).
template <typename T>
__device__ __forceinline__ bool off_band(T c, T s)
{
    constexpr T kBand = T(1) / T(64);
    return !(fabs(fma(c, c, fma(s, s, T(-1)))) <= kBand);
}

// Zeroes the two sync words of a freshly allocated workspace (a kernel rather than
// hipMemsetAsync, so that a captured launch sequence holds no memset node).
__global__ void __launch_bounds__(64) k_zero_sync(unsigned* __restrict__ sync)
{
    if (threadIdx.x < 2) sync[threadIdx.x] = 0u;
}

// Layers of one pass of every geometry (the scale table holds this many per column).
constexpr int kMaxPassLayers = 64;

// Device memory of one apply call: the coefficient table and the auxiliary words (column
// scales, input / output scale factors, normalisation flag).
template <typename T>
struct ApplyMem {
    T* table;
    T* escale;  // [n][kMaxPassLayers]: column scale before each layer of the pass
    T* sigm1;   // [n]: the column's scale factor at the pass boundary, minus one
    int* flag;  // the pass's normalisation flag
    unsigned* sync;  // k_colscale's two completion words (zero between launches)
};

// The auxiliary words of an n-column apply, laid out from `aux` (256-byte aligned).
inline size_t aux_bytes(int64_t n, size_t tsize)
{
    return tsize * (size_t)n * (kMaxPassLayers + 1) + 256;
}
template <typename T>
ApplyMem<T> apply_mem(void* table, void* aux, int64_t n, void* sync)
{
    T* e = static_cast<T*>(aux);
    T* sigm1 = e + (size_t)n * kMaxPassLayers;
    return ApplyMem<T>{static_cast<T*>(table), e, sigm1, reinterpret_cast<int*>(sigm1 + n),
                       static_cast<unsigned*>(sync)};
}

// The reference's identity skip: bulk_wave_order_2x2_rmul! skips a fused block of four
// (three at the edges) rotations only when all their s are zero (src/wave.jl:83,98,116,
// 137,152), so a rotation with s = 0 and c != 1 is applied unless its whole block is
// s-zero.  With 0-based (i, L), the block pairs layers Le = L & ~1 and Le + 1 and holds
// rows b, b + 1 of layer Le and b - 1, b of layer Le + 1, where b = i & ~1 for even L and
// (i + 1) & ~1 for odd L; rows outside 0 .. n-2 and layers past k are not part of it.
template <typename T>
__device__ __forceinline__ bool rot_skipped(const T* __restrict__ cs, int64_t ld, int n,
                                            int64_t i, int L, int k)
{
    const int Le = L & ~1;
    const int64_t b = (L & 1) ? ((i + 1) & ~(int64_t)1) : (i & ~(int64_t)1);
    const int64_t rows[4] = {b, b + 1, b - 1, b};
    const int lays[4] = {Le, Le, Le + 1, Le + 1};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (rows[q] < 0 || rows[q] > n - 2 || lays[q] >= k) continue;
        if (cs[2 * (rows[q] + ld * lays[q]) + 1] != T(0)) return false;
    }
    return true;
}

// The scaled three-shear form.  Rotation (i, L) maps its columns a = col i, b = col i+1 to
// a' = c a + s b, b' = c b - s a (Rotation2 rmul!).  With rho = sqrt(c^2 + s^2) and the
// signed rho_s = sign(c) rho, it is rho_s times a rotation with C = |c| / rho >= 0,
// S = s / rho_s.  Columns are stored as x^ = e x with a per-column scale e that absorbs
// every rho_s: each rotation divides the scales of both its columns by rho_s, so on the
// stored values it is  a^' = C a^ + S r b^,  b^' = -(S / r) a^ + C b^  with r = e_a / e_b,
// a matrix of determinant C^2 + S^2 = 1, which three shears realise exactly:
//   a^ += p b^;  b^ += q a^;  a^ += p b^   with  p = S r / (1 + C),  q = -S / r.
// |p| <= |r| and |q| <= 1/|r|, and r stays near 1 inside the band the normalisation check
// admits.
// A plain rotation has rho = 1 and e = +-1: that is sign tracking.  Grids whose pairs are
// only near normalised thus cost no accuracy: the determinant 1 + delta of each rotation
// goes into the scales instead of being dropped.  Skipped and out-of-grid slots have
// rho_s = 1.
template <typename T>
__device__ __forceinline__ T rho_signed(T c, T s)
{
    const T rho = sqrt(fma(c, c, s * s));
    return c < T(0) ? -rho : rho;
}
// The rho_s applied by slot (i, L), from its pair (c, s) already loaded (in: the slot lies in
// the grid): 1 outside the grid and for a skipped rotation.  Only a pair with s = 0 and
// c != 1 (whose rho_s is c, not 1, unless its fused block is skipped) reads the skip rule's
// other entries.
template <typename T>
__device__ __forceinline__ T applied_rho_of(const T* __restrict__ cs, int64_t ld, int n,
                                            int64_t i, int L, int k, bool in, T c, T s)
{
    if (!in) return T(1);
    if (s == T(0) && c != T(1) && rot_skipped<T>(cs, ld, n, i, L, k)) return T(1);
    return rho_signed<T>(c, s);
}

// Column scales of one pass (layers L0 .. L0 + Kp - 1).  d_j(lr) is the product of the
// rho_s of the rotations that touched column j before layer L0 + lr ((j-1, L) then (j, L)
// in each layer), D_j = d_j(Kp).  The stored column is x^ = e x with e = D_j / d_j when
// in_fix (the kernel multiplies the column by D_j as it enters the pass, and it leaves
// with e = 1), and e = 1 / d_j otherwise (it enters unscaled, and the kernel multiplies it
// by D_j as it leaves).  escale[j][lr] = e before layer L0 + lr, sigm1[j] = D_j - 1 (not
// read on the five-instruction path, which tracks nothing).
template <typename T>
__global__ void __launch_bounds__(1024) k_colscale(const T* __restrict__ cs, int64_t ld, int n,
                                                   int k, int L0, int Kp, int in_fix, int check,
                                                   unsigned* __restrict__ sync,
                                                   int* __restrict__ flag,
                                                   T* __restrict__ escale, T* __restrict__ sigm1)
{
    // 64 columns per block, sixteen threads per column, each over 4 of the (at most 64)
    // layers: the running product within its chunk, then the chunks' products in order.
    // The scales go through an LDS tile so that the block writes its 64 x 64 region of
    // escale with consecutive lanes on consecutive layers.
    constexpr int kChunks = 16, kPer = kMaxPassLayers / kChunks;
    __shared__ T cp[kChunks][64];
    __shared__ T tile[64][kMaxPassLayers + 1];
    const int c = threadIdx.x % 64, q = threadIdx.x / 64;
    const int j0 = blockIdx.x * 64, j = j0 + c;
    // every pair of the chunk is loaded first (clamped in-grid addresses), so that all the
    // loads are in flight together.  The scales are computed whatever the flag says (the
    // five-instruction path reads none of them).
    T ca[kPer], sa[kPer], cb[kPer], sb[kPer];
    bool ia[kPer], ib[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int lr = q * kPer + u, L = L0 + lr;
        const bool use = j < n && lr < Kp && L < k;
        ia[u] = use && j >= 1;
        ib[u] = use && j <= n - 2;
        const int64_t a = ia[u] ? 2 * ((j - 1) + ld * L) : 0;
        const int64_t b = ib[u] ? 2 * (j + ld * L) : 0;
        ca[u] = cs[a];
        sa[u] = cs[a + 1];
        cb[u] = cs[b];
        sb[u] = cs[b + 1];
    }
    T part[kPer];  // product of the chunk's layers before layer q kPer + u
    T run = T(1);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int L = L0 + q * kPer + u;
        part[u] = run;
        run *= applied_rho_of<T>(cs, ld, n, j - 1, L, k, ia[u], ca[u], sa[u]) *
               applied_rho_of<T>(cs, ld, n, j, L, k, ib[u], cb[u], sb[u]);
    }
    // the normalisation check, on the pairs (j, L) of columns j = 0 .. n-2: every pair once
    bool bad = false;
    if (check) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) bad |= ib[u] && off_band<T>(cb[u], sb[u]);
    }
    // The words live in L2 (device-coherent atomics).  A wave waits for its OR to return
    // before the barrier, so every verdict of the workgroup is in sync[0] before thread 0
    // takes the ticket, and the last ticket sees all of them.
    if (__any(bad) && (threadIdx.x % 64) == 0) {
        const unsigned old = atomicOr(&sync[0], 1u);
        asm volatile("" ::"v"(old));
    }
    cp[q][c] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (atomicAdd(&sync[1], 1u) == gridDim.x - 1) {
            *flag = (int)atomicExch(&sync[0], 0u);
            atomicExch(&sync[1], 0u);
        }
    }
    T pre = T(1), D = T(1);
    for (int t = 0; t < kChunks; ++t) {
        if (t < q) pre *= cp[t][c];
        D *= cp[t][c];
    }
    const T num = in_fix ? D : T(1);
#pragma unroll
    for (int u = 0; u < kPer; ++u) tile[c][q * kPer + u] = num / (pre * part[u]);
    if (q == 0 && j < n) sigm1[j] = D - T(1);
    __syncthreads();
    T* __restrict__ out = escale + (int64_t)j0 * kMaxPassLayers;
    for (int e = threadIdx.x; e < 64 * kMaxPassLayers; e += blockDim.x) {
        const int col = e / kMaxPassLayers;
        if (j0 + col < n) out[e] = tile[col][e % kMaxPassLayers];
    }
}

// Table entries of slot (i, L) = layer L0 + lr: (0, 0) outside the grid and for a skipped
// rotation (rot_skipped), (c - 1, s) on the five-instruction path, else the shear
// coefficients (p, q) of the scaled form from the column scales.  Every load is issued
// first, from clamped in-grid addresses, so that they are all in flight together; only a
// pair with s = 0 reads the skip rule's other entries.
template <typename T>
__device__ __forceinline__ void slot_coefs(const T* __restrict__ cs, int64_t ld, int n,
                                           int64_t i, int L, int k, int lr, int general,
                                           const T* __restrict__ escale, T& p, T& qq)
{
    const bool in = L < k && i >= 0 && i <= n - 2;
    const bool inm = in && i >= 1;
    const int64_t at = in ? 2 * (i + ld * L) : 0, atm = inm ? 2 * (i - 1 + ld * L) : 0;
    const int64_t ie = in ? i : 0;
    const T c = cs[at], s = cs[at + 1], cm = cs[atm], sm = cs[atm + 1];
    // scale of column i before layer L, of column i + 1 before layer L
    const T e0 = escale[ie * kMaxPassLayers + lr], eb = escale[(ie + 1) * kMaxPassLayers + lr];
    p = T(0);
    qq = T(0);
    if (!in || (s == T(0) && rot_skipped<T>(cs, ld, n, i, L, k))) return;
    if (general) {
        p = c - T(1);
        qq = s;
        return;
    }
    const T rs = rho_signed<T>(c, s);
    const T C = fabs(c) / fabs(rs), S = s / rs;
    // scale of column i after rotation (i-1, L) touched it
    const T ea = e0 / applied_rho_of<T>(cs, ld, n, i - 1, L, k, inm, cm, sm);
    const T r = ea / eb;
    p = S * r / (T(1) + C);
    qq = -S / r;
}

// Coefficient table of one pass: tab[g][u][2 KG] for g < G, u = t + S + PADU (time t,
// PADU rows of padding before and enough after that every interval of every wave reads
// inside the table).  Slot l of wave g at time t is rotation i = t - 2 (g KG + l) of layer
// L0 + g KG + l; [2l] = p, [2l+1] = q (shear form) or c - 1, s (general form);
// skipped (rot_skipped) and out-of-grid slots hold (0, 0), the identity in both forms.
// Section g = G holds, for every interval bt, the first wave's 16-lane input-scale row
// (sigm1 of the columns bt S + 1 .. bt S + S that enter the pass in that interval) at the
// position a coefficient row of a (virtual) wave G would have, so the coefficient loaders
// fetch it with the same address pattern.
template <typename T, int KG, int G, int S>
__global__ void __launch_bounds__(256) k_coef_table(T* __restrict__ tab,
                                                    const T* __restrict__ cs, int64_t ld,
                                                    int n, int k, int L0, int Up, int PADU,
                                                    int NI, const int* __restrict__ flag,
                                                    const T* __restrict__ escale,
                                                    const T* __restrict__ sigm1)
{
    constexpr int NC = 2 * KG;
    const int64_t rot_slots = (int64_t)G * Up * KG;
    const int64_t total = rot_slots + (int64_t)NI * 16;
    const int general = *flag;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        if (q >= rot_slots) {  // input-scale section
            const int qq = (int)(q - rot_slots);
            const int bt = qq / 16, c = qq % 16;
            const int col = bt * S + 1 + c;
            const T v = (c < S && col >= 0 && col < n && !general) ? sigm1[col] : 0.0;
            tab[(int64_t)G * Up * NC + (int64_t)(bt * S + G * (1 - S) + S + PADU) * NC + c] = v;
            continue;
        }
        const int l = (int)(q % KG);
        const int64_t r = q / KG;
        const int u = (int)(r % Up);
        const int g = (int)(r / Up);
        const int lr = g * KG + l;
        const int L = L0 + lr;
        const int64_t i = (int64_t)(u - S - PADU) - 2 * (int64_t)lr;
        T p, qq;
        slot_coefs<T>(cs, ld, n, i, L, k, lr, general, escale, p, qq);
        tab[2 * q] = p;
        tab[2 * q + 1] = qq;
    }
}

// General (c, s): t = a; a += (c-1) a; a += s b; b += (c-1) b; b += (-s) t, two
// rotations per asm block (lanes 2l, 2l+1 hold c-1, s of rotation l).  The outputs are
// early-clobber: the block writes a0/a1 before it reads the copies t0/t1, so those must
// not share registers with them.
#define SEP_GEN2(C0, S0, C1, S1)                                                             \
    asm("s_nop 1\n"                                                                          \
        SEP_DPP("%0", "%6", "%0", C0) SEP_DPP("%2", "%6", "%2", C1)                          \
        SEP_DPP("%0", "%6", "%1", S0) SEP_DPP("%2", "%6", "%3", S1)                          \
        SEP_DPP("%1", "%6", "%1", C0) SEP_DPP("%3", "%6", "%3", C1)                          \
        SEP_DPP("%1", "-%6", "%4", S0) SEP_DPP("%3", "-%6", "%5", S1)                        \
        : "+&v"(a0), "+&v"(b0), "+&v"(a1), "+&v"(b1)                                         \
        : "v"(t0), "v"(t1), "v"(cf))

template <typename T, int quad>
__device__ __forceinline__ void gen2(T& a0, T& b0, T& a1, T& b1, T cf)
{
    const T t0 = a0, t1 = a1;
#define SEP_GEN2_ANY                                                                        \
    if constexpr (quad == 0) SEP_GEN2(0, 1, 2, 3);                                          \
    else if constexpr (quad == 1) SEP_GEN2(4, 5, 6, 7);                                     \
    else if constexpr (quad == 2) SEP_GEN2(8, 9, 10, 11);                                   \
    else SEP_GEN2(12, 13, 14, 15);
    if constexpr (sizeof(T) == 8) {
#define SEP_FMAC "v_fmac_f64_dpp"
        SEP_GEN2_ANY
#undef SEP_FMAC
    } else {
#define SEP_FMAC "v_fmac_f32_dpp"
        SEP_GEN2_ANY
#undef SEP_FMAC
    }
#undef SEP_GEN2_ANY
}

template <typename T, int KG, int s, int base, int WR>
__device__ __forceinline__ void step_general(T (&w)[WR], T cf)
{
    static_assert(WR >= 2 * KG, "window ring too small");
#define SLOT(l, o) w[(((s) - 2 * (l) + (o)) % WR + WR) % WR]
    gen2<T, base / 4 + 0>(SLOT(0, 0), SLOT(0, 1), SLOT(1, 0), SLOT(1, 1), cf);
    gen2<T, base / 4 + 1>(SLOT(2, 0), SLOT(2, 1), SLOT(3, 0), SLOT(3, 1), cf);
    if constexpr (KG == 8) {
        gen2<T, 2>(SLOT(4, 0), SLOT(4, 1), SLOT(5, 0), SLOT(5, 1), cf);
        gen2<T, 3>(SLOT(6, 0), SLOT(6, 1), SLOT(7, 0), SLOT(7, 1), cf);
    }
#undef SLOT
}

// Scale of a column entering the pass (first wave): x += (D - 1) x, with D - 1 broadcast
// from lane s of the interval's scale row (exact when D = +-1, one rounding otherwise).
#define SEP_FIX(N) asm("s_nop 1\n" SEP_DPP("%0", "%1", "%0", N) : "+v"(x) : "v"(sg))

template <typename T, int s>
__device__ __forceinline__ void fix_sign(T& x, T sg)
{
    static_assert(s >= 0 && s < 16, "sign row has 16 lanes");
#define SEP_FIX_ANY                                                                         \
    if constexpr (s == 0) SEP_FIX(0);                                                       \
    else if constexpr (s == 1) SEP_FIX(1);                                                  \
    else if constexpr (s == 2) SEP_FIX(2);                                                  \
    else if constexpr (s == 3) SEP_FIX(3);                                                  \
    else if constexpr (s == 4) SEP_FIX(4);                                                  \
    else if constexpr (s == 5) SEP_FIX(5);                                                  \
    else if constexpr (s == 6) SEP_FIX(6);                                                  \
    else if constexpr (s == 7) SEP_FIX(7);                                                  \
    else if constexpr (s == 8) SEP_FIX(8);                                                  \
    else if constexpr (s == 9) SEP_FIX(9);                                                  \
    else if constexpr (s == 10) SEP_FIX(10);                                                \
    else if constexpr (s == 11) SEP_FIX(11);                                                \
    else if constexpr (s == 12) SEP_FIX(12);                                                \
    else if constexpr (s == 13) SEP_FIX(13);                                                \
    else if constexpr (s == 14) SEP_FIX(14);                                                \
    else SEP_FIX(15);
    if constexpr (sizeof(T) == 8) {
#define SEP_FMAC "v_fmac_f64_dpp"
        SEP_FIX_ANY
#undef SEP_FMAC
    } else {
#define SEP_FMAC "v_fmac_f32_dpp"
        SEP_FIX_ANY
#undef SEP_FMAC
    }
#undef SEP_FIX_ANY
}

template <int S, int s0>
struct Steps {
    template <class F>
    __device__ __forceinline__ static void run(F&& body)
    {
        if constexpr (s0 < S) {
            body(std::integral_constant<int, s0>{});
            Steps<S, s0 + 1>::run(body);
        }
    }
};

// Launch geometry of one systolic apply kernel.
template <int KG_, int G_, int S_, int PFD_>
struct Geom {
    static constexpr int KG = KG_;    // layers per wave
    static constexpr int G = G_;      // waves per workgroup (layer groups)
    static constexpr int S = S_;      // time steps per barrier interval
    static constexpr int PFD = PFD_;  // intervals the first wave prefetches Q ahead
    static constexpr int NC = 2 * KG;               // coefficients per step
    static constexpr int SPR = 16 / NC;             // steps per coefficient row
    static constexpr int NR = S / SPR;              // coefficient rows per interval
    static constexpr int W = 2 * KG;                // window columns
    static constexpr int Kp = KG * G;               // layers per pass
    static_assert(S % W == 0 && S % SPR == 0 && S <= 16, "bad interval length");
};

// One column of the 64-row tile at At (uniform) through a buffer resource covering exactly
// that column when it lies in [0, n) and nothing otherwise: the hardware range check makes
// loads of columns outside the matrix return 0 and drops stores to them, so neither needs
// a branch or a per-lane address (voff = the lane's row offset in bytes).
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t col_rsrc(T* p, int col, int n)
{
    const int bytes = (col >= 0 && col < n) ? kLanes * (int)sizeof(T) : 0;
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, bytes, 0x00020000);
}
// p = the column's first element in this tile (uniform), col = its index
template <typename T>
__device__ __forceinline__ T col_load(T* p, int col, int n, uint32_t voff)
{
    const __amdgpu_buffer_rsrc_t r = col_rsrc(p, col, n);
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
}
template <typename T>
__device__ __forceinline__ void col_store(T* p, int col, int n, uint32_t voff, T v)
{
    const __amdgpu_buffer_rsrc_t r = col_rsrc(p, col, n);
    if constexpr (sizeof(T) == 8) {
        using u2 = unsigned __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, 0, 0);
    }
}

// One barrier interval (S time steps t0 .. t0 + S - 1) of wave g.  Every wave runs the
// same code at every interval: the first wave's Q loads and the last wave's Q stores are
// range-checked by the hardware, and a middle wave's inputs outside [0, n) need no test
// (columns past n are zeros that only identity rotations touch; the LDS ring starts
// zeroed, so an input slot its producer has not written yet also reads 0).
template <typename T, class Gm, bool FIRST, bool LAST, bool GENERAL, int D>
__device__ __forceinline__ void interval(T* At, uint32_t voff, int64_t lda, int n,
                                         T (&win)[Gm::W], T (&pf)[FIRST ? Gm::PFD : 1][Gm::S],
                                         const T* cbw, T sg, const T (&cfo)[Gm::NR],
                                         T (&cfn)[Gm::NR], const T* __restrict__ tn,
                                         const T* lds_in, T* lds_out, int t0)
{
    constexpr int S = Gm::S, W = Gm::W, KG = Gm::KG, PFD = Gm::PFD, NR = Gm::NR;
    constexpr int SPR = Gm::SPR;
    // Reads of LDS (this wave's input columns, and the first and last waves' coefficient
    // rows) are issued LA steps before the step that uses them.  Issuing them all at the
    // start of the interval makes every wave of the workgroup queue its whole interval's
    // reads at the LDS right after the barrier, so the last wave in the queue waits for all
    // of them before its first FMA.  The ring holds the columns in pairs of steps (lane l's
    // pair p at 2 (64 p + l)): one 16-byte write per two steps from the producer and one
    // 16-byte read per two steps into the consumer (0.9% faster at m = n = 16384 than one
    // 8-byte access per step: fewer LDS instructions, less power).
    constexpr int LA = 2;
    constexpr bool OWN = !FIRST && !LAST;  // coefficient rows in registers
    using V2 = T __attribute__((ext_vector_type(2)));
    T in[FIRST ? 1 : S];
    T cfl[OWN ? 1 : NR];
    T hold = T(0);  // the even step's output, written with the odd step's
    auto read_in = [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        if constexpr (!FIRST && s < S && s % 2 == 0) {
            const V2 v = *reinterpret_cast<const V2*>(lds_in + (s / 2) * kLanes * 2);
            in[s] = v.x;
            in[s + 1] = v.y;
        }
        if constexpr (!OWN && s < S && s % SPR == 0) cfl[s / SPR] = cbw[(s / SPR) * 16];
    };
    Steps<LA, 0>::run(read_in);
    // first wave: prefetch slot D holds this interval's columns (loaded PFD intervals
    // ago); each is refilled, right after its step consumed it, with the column PFD
    // intervals ahead (t0 + PFD S + 1 + s)
    const int c0f = t0 + PFD * S + 1;
    const int o0 = t0 - 2 * Gm::Kp + 2;  // last wave: column leaving at step 0
    // running column pointers (uniform; one scalar add per step)
    T* pq = At + (int64_t)(FIRST ? c0f : o0) * lda;
    Steps<S, 0>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int row = s / SPR, base = (s % SPR) * Gm::NC;
        // Progress-based priority: 3 in the first quarter of the interval down to 0 in
        // the last, so that of two waves sharing a SIMD the one that is behind gets the
        // issue slots (the arbiter otherwise favours the older wave, which then reaches
        // the barrier early and leaves the younger one to run alone).
        if constexpr (s % (S / 4) == 0) __builtin_amdgcn_s_setprio(3 - s / (S / 4));
        read_in(std::integral_constant<int, s + LA>{});
        // middle waves: row s of the next interval's coefficients (cfo holds this one's)
        if constexpr (OWN && s < NR) cfn[s] = tn[s * 16];
        T cfr;
        if constexpr (OWN)
            cfr = cfo[row];
        else
            cfr = cfl[row];
        if constexpr (FIRST) {
            // the column enters the pass with its final scale (see k_colscale)
            win[(s + 1) % W] = pf[D][s];
            if constexpr (!GENERAL) fix_sign<T, s>(win[(s + 1) % W], sg);
            pf[D][s] = col_load<T>(pq, c0f + s, n, voff);
            pq += lda;
        } else {
            win[(s + 1) % W] = in[s];
        }
        if constexpr (GENERAL)
            step_general<T, KG, s, base>(win, cfr);
        else
            step_shear<T, KG, s, base>(win, cfr);
        T out = win[(s + 2) % W];
        if constexpr (LAST) {
            col_store<T>(pq, o0 + s, n, voff, out);
            pq += lda;
        } else if constexpr (s % 2 == 0) {
            hold = out;
        } else {
            *reinterpret_cast<V2*>(lds_out + (s / 2) * kLanes * 2) = V2{hold, out};
        }
        __builtin_amdgcn_sched_barrier(0);
    });
}

// LDS layout of one interval's coefficient block: rows [2][r][16] of the first and the
// last wave, then the first wave's 16-lane input-scale row.
template <class Gm>
struct CoefBlock {
    static constexpr int ROWS = 2 * Gm::NR * 16;
    static constexpr int SIZE = ROWS + 16;
    static constexpr int LOADERS = Gm::G > 2 ? Gm::G - 2 : 1;          // middle waves
    static constexpr int PER = (SIZE + LOADERS * kLanes - 1) / (LOADERS * kLanes);
};

// Element offsets (interval 0) of the block entries a loader lane fetches; interval bt
// adds bt S 2KG.  Entries past the block fetch element 0 and are not stored.
template <class Gm>
__device__ __forceinline__ void block_offsets(int (&off)[CoefBlock<Gm>::PER], int Up, int PADU,
                                              int loader, int lane)
{
    using CB = CoefBlock<Gm>;
    constexpr int S = Gm::S, NC = Gm::NC, NR = Gm::NR, G = Gm::G;
#pragma unroll
    for (int j = 0; j < CB::PER; ++j) {
        const int q = loader * kLanes + lane + j * CB::LOADERS * kLanes;
        int gq = 0, r = 0, c = 0;
        if (q < CB::ROWS) {
            gq = q < NR * 16 ? 0 : G - 1;
            r = (q / 16) % NR;
            c = q % 16;
        } else if (q < CB::SIZE) {
            gq = G;
            c = q - CB::ROWS;
        }
        off[j] = (q < CB::SIZE) ? (gq * Up + gq * (1 - S) + S + PADU) * NC + r * 16 + c : 0;
    }
}

// The middle waves fetch the next interval's coefficient block (the first and last
// waves' rows and the input scales) from the table into registers at the start of an
// interval and store it to LDS at the end; their other global memory operations are
// loads of their own coefficient rows, so waiting for these loads never drains the first
// wave's Q prefetch or the last wave's stores.
template <typename T, class Gm>
__device__ __forceinline__ void fetch_block(T (&v)[CoefBlock<Gm>::PER],
                                            const int (&off)[CoefBlock<Gm>::PER],
                                            const T* __restrict__ tab, int bt)
{
    const T* __restrict__ base = tab + (int64_t)bt * (Gm::S * Gm::NC);
#pragma unroll
    for (int j = 0; j < CoefBlock<Gm>::PER; ++j) v[j] = base[off[j]];
}

template <typename T, class Gm>
__device__ __forceinline__ void store_block(T* cb, const T (&v)[CoefBlock<Gm>::PER],
                                            int loader, int lane)
{
    using CB = CoefBlock<Gm>;
#pragma unroll
    for (int j = 0; j < CB::PER; ++j) {
        const int q = loader * kLanes + lane + j * CB::LOADERS * kLanes;
        if (q < CB::SIZE) cb[q] = v[j];
    }
}

template <typename T, class Gm, bool FIRST, bool LAST, bool GENERAL>
__device__ __forceinline__ void wave(T* __restrict__ A, int64_t lda, int m, int n,
                                     const T* __restrict__ tab, const T* __restrict__ sigm1,
                                     int Up, int PADU, int NI,
                                     T (*ring)[2][Gm::S][kLanes],
                                     T (*cbuf)[CoefBlock<Gm>::SIZE], int g, int lane)
{
    using CB = CoefBlock<Gm>;
    constexpr int S = Gm::S, W = Gm::W, PFD = Gm::PFD, NR = Gm::NR;
    constexpr bool LOADER = !FIRST && !LAST;
    const int loader = g - 1;
    // Lanes past the last row alias row m-1: they load, compute and store exactly the
    // values of that row's own lane, so no per-lane predication is needed.
    const int64_t row0 = (int64_t)blockIdx.x * kLanes;
    T* At = A + row0;  // the tile's first row (uniform)
    const uint32_t voff = (uint32_t)(min(row0 + lane, (int64_t)m - 1) - row0) * sizeof(T);
    const int cl = lane & 15;  // lane within the 16-lane DPP row

    T win[W];
#pragma unroll
    for (int w = 0; w < W; ++w) win[w] = 0.0;
    T pf[FIRST ? PFD : 1][S];
    if constexpr (FIRST) {
        // column 0 enters before step 0 (slot 0); interval d takes columns dS+1 .. dS+S
        win[0] = col_load<T>(At, 0, n, voff);
        if constexpr (!GENERAL) win[0] = fma(sigm1[0], win[0], win[0]);  // its final scale
#pragma unroll
        for (int d = 0; d < PFD; ++d)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int col = d * S + s + 1;
                pf[d][s] = col_load<T>(At + (int64_t)col * lda, col, n, voff);
            }
    }
    if constexpr (!LAST) {
        // zero both buffers of this wave's output ring (read before first written, see
        // interval())
#pragma unroll
        for (int s = 0; s < S; ++s) {
            ring[g][0][s][lane] = 0.0;
            ring[g][1][s][lane] = 0.0;
        }
    }
    // a middle wave's own coefficient rows: interval b in cfb[b & 1], interval b + 1
    // loaded into the other buffer one row per step while b runs
    constexpr int NC = Gm::NC;
    const T* __restrict__ tw = tab + (int64_t)(g * Up + g * (1 - S) + S + PADU) * NC + cl;
    T cfb[2][NR];
    if constexpr (LOADER) {
#pragma unroll
        for (int r = 0; r < NR; ++r) cfb[0][r] = tw[r * 16];
    }
    T blk[CB::PER];
    int off[CB::PER];
    if constexpr (LOADER) {  // block of interval 0
        block_offsets<Gm>(off, Up, PADU, loader, lane);
        fetch_block<T, Gm>(blk, off, tab, 0);
        store_block<T, Gm>(cbuf[0], blk, loader, lane);
    }
    lds_barrier();
    // The interval loop is unrolled by the prefetch depth (rounded up to even) so that
    // prefetch slot D % PFD and coefficient buffer D & 1 are compile-time registers,
    // consumed and refilled in place (no register shuffling of in-flight loads).
    constexpr int U = PFD % 2 == 0 ? PFD : 2 * PFD;
    for (int b0 = 0; b0 < NI; b0 += U) {
        Steps<U, 0>::run([&](auto dc) {
            constexpr int D = decltype(dc)::value;
            const int b = b0 + D;
            if (b >= NI) return;
            if constexpr (LOADER) {
                if (b + 1 < NI) fetch_block<T, Gm>(blk, off, tab, b + 1);
            }
            // Every wave runs every interval, also before its first column arrives and
            // after its last one left: those intervals move zeros through identity
            // rotations (their table rows are in range and hold zeros), store nothing
            // (range check) and keep the window in the same registers on every path.
            const int t0 = (b - g) * S + g;
            const T* cb = cbuf[b & 1];
            const T* cbw = cb + (LAST ? NR * 16 : 0) + cl;  // first / last wave's rows
            const T sg = FIRST ? cb[CB::ROWS + cl] : 0.0;  // input scales
            // next interval's rows (the last interval re-reads its own: in range, unused)
            const int64_t o = (int64_t)min(b + 1, NI - 1) * (S * NC);
            const T* lds_in =
                FIRST ? nullptr : &ring[g > 0 ? g - 1 : 0][(b - 1) & 1][0][0] + 2 * lane;
            T* lds_out = LAST ? nullptr : &ring[g][b & 1][0][0] + 2 * lane;
            interval<T, Gm, FIRST, LAST, GENERAL, D % PFD>(At, voff, lda, n, win, pf, cbw, sg,
                                                           cfb[D & 1], cfb[(D + 1) & 1], tw + o,
                                                           lds_in, lds_out, t0);
            if constexpr (LOADER) {
                if (b + 1 < NI) store_block<T, Gm>(cbuf[(b + 1) & 1], blk, loader, lane);
            }
            lds_barrier();
        });
    }
}

template <typename T, class Gm>
__global__ void __launch_bounds__(kLanes * Gm::G)
    k_apply_dpp(T* __restrict__ A, int64_t lda, int m, int n,
                const T* __restrict__ tab, const T* __restrict__ sigm1,
                const int* __restrict__ flag, int Up, int PADU, int NI)
{
    constexpr int G = Gm::G;
    static_assert(G >= 3, "the coefficient loaders are the middle waves");
    __shared__ __align__(16) T ring[G - 1][2][Gm::S][kLanes];
    __shared__ T cbuf[2][CoefBlock<Gm>::SIZE];
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int lane = threadIdx.x % kLanes;
    const bool general = *flag != 0;
#define SEP_WAVE(GEN)                                                                         \
    if (g == 0)                                                                               \
        wave<T, Gm, true, false, GEN>(A, lda, m, n, tab, sigm1, Up, PADU, NI, ring, cbuf, g,     \
                                   lane);                                                     \
    else if (g == G - 1)                                                                      \
        wave<T, Gm, false, true, GEN>(A, lda, m, n, tab, sigm1, Up, PADU, NI, ring, cbuf, g,     \
                                   lane);                                                     \
    else                                                                                      \
        wave<T, Gm, false, false, GEN>(A, lda, m, n, tab, sigm1, Up, PADU, NI, ring, cbuf, g,    \
                                    lane);
    if (general) {
        SEP_WAVE(true)
    } else {
        SEP_WAVE(false)
    }
#undef SEP_WAVE
}

// Interval counts of the systolic geometry for n columns.
template <class Gm>
struct DppDims {
    int NI, PADU, Up;
    explicit DppDims(int64_t n)
    {
        const int T_out = (int)(n + 2 * Gm::Kp - 2);
        NI = Gm::G + (T_out + Gm::S - 1) / Gm::S + 1;
        PADU = (Gm::G - 1) * (Gm::S - 1);
        Up = NI * Gm::S + 2 * Gm::S + PADU;
    }
    size_t table_elems() const { return (size_t)(Gm::G + 1) * Up * Gm::NC + 64; }
};

// th: the handle whose timing counters record the launches (may be null).
template <typename T, class Gm>
int run_dpp(sep_handle* th, hipStream_t st, const ApplyMem<T>& mem, int64_t m, int64_t n, T* A,
            int64_t lda, const T* cs, int64_t ld_cs, int64_t k, bool rotations)
{
    constexpr int G = Gm::G, S = Gm::S, KG = Gm::KG, Kp = Gm::Kp;
    sep_handle* h = th;
    const DppDims<Gm> dd(n);
    const int NI = dd.NI, PADU = dd.PADU, Up = dd.Up;
    T* escale = mem.escale;
    T* sigm1 = mem.sigm1;
    int* flag = mem.flag;
    const int tp0 = timing_mark(h, st);
    const int64_t rb = (m + kLanes - 1) / kLanes;
    int tp_prev = tp0;
    for (int64_t L0 = 0; L0 < k; L0 += Kp) {
        hipLaunchKernelGGL(k_colscale<T>, dim3((unsigned)((n + 63) / 64)), dim3(1024), 0, st, cs,
                           ld_cs, (int)n, (int)k, (int)L0, Kp, 1, rotations ? 0 : 1, mem.sync,
                           flag, escale, sigm1);
        SEP_CHECK_LAUNCH();
        const int64_t total = (int64_t)G * Up * KG + (int64_t)NI * 16;
        const int tblocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL((k_coef_table<T, KG, G, S>), dim3(tblocks), dim3(256), 0, st,
                           mem.table, cs, ld_cs, (int)n, (int)k, (int)L0, Up, PADU, NI, flag,
                           escale, sigm1);
        SEP_CHECK_LAUNCH();
        const int ta = timing_mark(h, st);
        timing_pair(h, false, tp_prev, ta);
        hipLaunchKernelGGL((k_apply_dpp<T, Gm>), dim3((unsigned)rb), dim3(kLanes * G), 0, st, A,
                           lda, (int)m, (int)n, mem.table, sigm1, flag, Up, PADU, NI);
        SEP_CHECK_LAUNCH();
        tp_prev = timing_mark(h, st);
        timing_pair(h, true, ta, tp_prev);
    }
    return SEP_OK;
}

using GeomWide = Geom<8, 8, 16, 2>;   // 64 layers per pass, 2 waves per SIMD
using GeomSmall = Geom<4, 8, 8, 3>;   // 32 layers per pass (short grids)

// ---- 16-row geometry (R16_BEGIN) ---------------------------------------------------
// One workgroup owns 16 rows of A and all Kp = 4 NW KG layers of a pass.  Lane 16 r + j
// of wave w holds row j of the block in layer group q = 4 w + r (layers q KG .. q KG +
// KG - 1): every 16-lane DPP row is a different layer group, and the row_newbcast
// coefficient broadcast gives each group its own rotations.  Groups hand columns on
// through LDS: in interval b (S steps) every lane writes the column leaving its window to
// out[w][b & 1][s][lane]; in interval b + 1 lane 16 r + j (r > 0) reads the slot of lane
// 16 (r - 1) + j of its own wave, and lane j (r = 0) reads lane 48 + j of the previous
// wave's buffer (wave 0: a staging buffer that wave 0 fills with columns of A).  Each
// group therefore runs S - 1 time steps behind the previous one: group q at local step u
// applies rotation i = u - q (S - 1) - 2 (q KG + l) of its layer l, and the last group
// hands out column u - DL.  Every wave runs the same FMA stream with no bounds tests
// (columns outside the matrix are zeros under identity rotations); wave 0 additionally
// loads A four columns per instruction (lane 16 c + j: column c, row j) and the last wave
// stores, scales and reads back its outputs the same way.  With 16 rows per
// workgroup, four independent workgroups share a CU, so the waves sharing a SIMD belong
// to different workgroups and do not stall at the same barriers.
constexpr int gcd_(int a, int b) { return b == 0 ? a : gcd_(b, a % b); }
constexpr int lcm_(int a, int b) { return a / gcd_(a, b) * b; }

template <int KG_, int NW_, int S_, int RPG_ = 16>
struct G16 {
    static constexpr int KG = KG_, NW = NW_, S = S_;
    static constexpr int RPG = RPG_;           // rows of A per block = lanes per layer group
    static constexpr int GPW = kLanes / RPG;   // layer groups per wave
    static constexpr int OL = kLanes - RPG;    // first lane of a wave's last group
    static constexpr int NG = GPW * NW;        // layer groups
    static constexpr int Kp = NG * KG;  // layers per pass
    static constexpr int W = 2 * KG;    // window columns
    static constexpr int NC = 2 * KG;   // coefficients per step and group
    static constexpr int SPR = 16 / NC; // steps per coefficient register
    static constexpr int NR = S / SPR;  // coefficient registers per interval
    static constexpr int NQ = S / GPW;  // GPW-column loads / stores per interval
    static constexpr int DL = (NG - 1) * (S - 1) + 2 * Kp - 2;  // output column = u - DL
    static constexpr int WR = W + S;    // window ring: an interval's inputs load into free slots
    static constexpr int U = lcm_(WR, 2 * S) / S;  // intervals per unrolled loop body (even)
    static_assert(S % GPW == 0 && S % SPR == 0 && U % 2 == 0, "bad interval length");
    static_assert(RPG == 16 || RPG == 32, "a layer group is one or two DPP rows");
};

// Per-wave constants of the 16-row kernel.
template <typename T>
struct R16Args {
    T* A0;            // A + first row of the block (uniform)
    int64_t voff;          // lane offset: (lane / RPG) lda + row within the block
    int64_t lda;
    const T* tab;
    const T* sigm1;
    int n, NBt, w, lane;
    bool first, last;
};

// tab[((w NBt + b) NR + rr) 64 + lane]: the coefficient register rr of interval b of wave
// w.  Lane 16 r + j of it belongs to group q = 4 w + r, step u = b S + rr SPR + j / NC,
// layer q KG + (j % NC) / 2; even j hold p (shear) or c - 1 (general), odd j hold q or s.
template <typename T, class Gm>
__global__ void __launch_bounds__(256) k_coef_r16(T* __restrict__ tab,
                                                  const T* __restrict__ cs, int64_t ld,
                                                  int n, int k, int L0, int NBt,
                                                  const int* __restrict__ flag,
                                                  const T* __restrict__ escale)
{
    constexpr int S = Gm::S, KG = Gm::KG, NC = Gm::NC, SPR = Gm::SPR, NR = Gm::NR;
    const int64_t total = (int64_t)Gm::NW * NBt * NR * kLanes;
    const int general = *flag;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int lane = (int)(e % kLanes);
        const int64_t rest = e / kLanes;
        const int rr = (int)(rest % NR);
        const int64_t wb = rest / NR;
        const int b = (int)(wb % NBt);
        const int w = (int)(wb / NBt);
        const int j = lane % 16;  // coefficient slot within the DPP row
        const int q = Gm::GPW * w + lane / Gm::RPG;
        const int64_t u = (int64_t)b * S + rr * SPR + j / NC;
        const int lr = q * KG + (j % NC) / 2;
        const int L = L0 + lr;
        const int64_t i = u - (int64_t)q * (S - 1) - 2 * lr;
        T p, qq;
        slot_coefs<T>(cs, ld, n, i, L, k, lr, general, escale, p, qq);
        tab[e] = (j & 1) ? qq : p;
    }
}

// One interval (S steps) of every wave.  IU is the interval's index in the unrolled loop
// body: its parity P selects the LDS buffers and its first local step fixes the window
// slots, which are compile-time registers (the same for every lane).
template <typename T, class Gm, bool GENERAL, int IU>
__device__ __forceinline__ void r16_interval(const R16Args<T>& a, T (*buf)[2][Gm::S][kLanes],
                                             T (&win)[Gm::WR], T (&cf)[2][Gm::NR],
                                             T (&qv)[2][Gm::NQ], const T* rd,
                                             T* wr, int b)
{
    constexpr int S = Gm::S, WR = Gm::WR, KG = Gm::KG, NW = Gm::NW, NR = Gm::NR;
    constexpr int NQ = Gm::NQ, SPR = Gm::SPR, NC = Gm::NC, W = Gm::W;
    constexpr int P = IU & 1, s0 = IU * S;
    constexpr int RPG = Gm::RPG, GPW = Gm::GPW, OL = Gm::OL;
    const int lane = a.lane, r = lane / RPG, j = lane % RPG;
    // inputs of this interval (written in interval b - 1) go straight into their window
    // slots: slot (u + 1) mod WR was last used by the column that left S + 1 steps ago
#pragma unroll
    for (int s = 0; s < S; ++s) win[(s0 + s + 1) % WR] = rd[((1 - P) * S + s) * kLanes];
    // last wave: its outputs of interval b - 1, columns (b - 1) S + 4 k + r - DL
    T ob[NQ], sg[NQ];
    const int oc0 = (b - 1) * S - Gm::DL;  // output column of lane row 0, k = 0
    const bool oful = oc0 >= 0 && oc0 + S <= a.n;
    if (a.last) {
#pragma unroll
        for (int k4 = 0; k4 < NQ; ++k4) {
            ob[k4] = buf[NW][1 - P][GPW * k4 + r][OL + j];
            const int oc = oc0 + GPW * k4 + r;
            // the five-instruction path tracks no scales (its sigm1 is not computed for it)
            sg[k4] = (!GENERAL && (oful || (oc >= 0 && oc < a.n))) ? a.sigm1[oc] : 0.0;
        }
    }
    Steps<S, 0>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int sw = (s0 + s) % WR;
        constexpr int rr = s / SPR, base = (s % SPR) * NC;
        if constexpr (GENERAL)
            step_general<T, KG, sw, base>(win, cf[P][rr]);
        else
            step_shear<T, KG, sw, base>(win, cf[P][rr]);
        // the column leaving the window entered W - 1 steps ago
        wr[(P * S + s) * kLanes] = win[(sw + 2 - W + WR) % WR];
    });
    // coefficients of interval b + 2 (the table has NB + 2 intervals)
    const T* __restrict__ tb = a.tab + ((int64_t)(a.w * a.NBt + b + 2) * NR) * kLanes;
#pragma unroll
    for (int rr = 0; rr < NR; ++rr) cf[P][rr] = tb[rr * kLanes + lane];
    if (a.first) {
        // staging for interval b + 1 (columns (b + 1) S + 1 + s), then the columns of the
        // staging written in interval b + 2
#pragma unroll
        for (int k4 = 0; k4 < NQ; ++k4) buf[0][P][GPW * k4 + r][OL + j] = qv[P][k4];
        const int c0 = (b + 3) * S + 1;
        const T* __restrict__ src = a.A0 + (int64_t)c0 * a.lda;
        if (c0 + S <= a.n) {
#pragma unroll
            for (int k4 = 0; k4 < NQ; ++k4) qv[P][k4] = src[(int64_t)GPW * k4 * a.lda + a.voff];
        } else {
#pragma unroll
            for (int k4 = 0; k4 < NQ; ++k4)
                qv[P][k4] = c0 + GPW * k4 + r < a.n ? src[(int64_t)GPW * k4 * a.lda + a.voff] : 0.0;
        }
    }
    if (a.last) {
        T* __restrict__ dst = a.A0 + (int64_t)oc0 * a.lda;
        if (oful) {
#pragma unroll
            for (int k4 = 0; k4 < NQ; ++k4)
                dst[(int64_t)GPW * k4 * a.lda + a.voff] = fma(sg[k4], ob[k4], ob[k4]);
        } else {
#pragma unroll
            for (int k4 = 0; k4 < NQ; ++k4) {
                const int oc = oc0 + GPW * k4 + r;
                if (oc >= 0 && oc < a.n)
                    dst[(int64_t)GPW * k4 * a.lda + a.voff] = fma(sg[k4], ob[k4], ob[k4]);
            }
        }
    }
    lds_barrier();
}

template <typename T, class Gm, bool GENERAL>
__device__ __forceinline__ void r16_body(T* __restrict__ A, int64_t lda, int m, int n,
                                         const T* __restrict__ tab,
                                         const T* __restrict__ sigm1, int NB, int NBt,
                                         T (*buf)[2][Gm::S][kLanes])
{
    constexpr int S = Gm::S, WR = Gm::WR, NW = Gm::NW, NR = Gm::NR, NQ = Gm::NQ, U = Gm::U;
    R16Args<T> a;
    a.w = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    a.lane = threadIdx.x % kLanes;
    a.first = a.w == 0;
    a.last = a.w == NW - 1;
    a.n = n;
    a.lda = lda;
    a.NBt = NBt;
    a.tab = tab;
    a.sigm1 = sigm1;
    constexpr int RPG = Gm::RPG, GPW = Gm::GPW, OL = Gm::OL;
    const int r = a.lane / RPG, j = a.lane % RPG;
    const int64_t row0 = (int64_t)blockIdx.x * RPG;
    // Rows past m alias row m - 1: those lanes compute and store that row's own values.
    const int jr = (int)(min(row0 + j, (int64_t)m - 1) - row0);
    a.A0 = A + row0;                          // uniform
    a.voff = (int64_t)r * lda + jr;           // lane RPG c + j: column c, row j
    const T* rd = r == 0 ? &buf[a.w][0][0][OL + j] : &buf[a.w + 1][0][0][a.lane - RPG];
    T* wr = &buf[a.w + 1][0][0][a.lane];

    T win[WR];
#pragma unroll
    for (int s = 0; s < WR; ++s) win[s] = 0.0;
    // interval 0 reads parity 1: zero every wave's buffer, stage columns 1 .. S
#pragma unroll
    for (int s = 0; s < S; ++s) wr[(S + s) * kLanes] = 0.0;
    T qv[2][NQ];
    if (a.first) {
        if (r == 0) win[0] = A[row0 + jr];  // column 0 sits in slot 0 before step 0
#pragma unroll
        for (int k4 = 0; k4 < NQ; ++k4) {
            const int c0 = 1 + GPW * k4 + r, c1 = S + c0, c2 = 2 * S + c0;
            const T* __restrict__ p = a.A0 + (int64_t)(GPW * k4) * lda + a.voff;
            buf[0][1][GPW * k4 + r][OL + j] = c0 < n ? p[(int64_t)1 * lda] : 0.0;
            qv[0][k4] = c1 < n ? p[(int64_t)(S + 1) * lda] : 0.0;
            qv[1][k4] = c2 < n ? p[(int64_t)(2 * S + 1) * lda] : 0.0;
        }
    }
    T cf[2][NR];
#pragma unroll
    for (int rr = 0; rr < NR; ++rr) {
        cf[0][rr] = tab[((int64_t)(a.w * NBt + 0) * NR + rr) * kLanes + a.lane];
        cf[1][rr] = tab[((int64_t)(a.w * NBt + 1) * NR + rr) * kLanes + a.lane];
    }
    lds_barrier();
    for (int b = 0; b < NB; b += U) {
        Steps<U, 0>::run([&](auto ic) {
            constexpr int IU = decltype(ic)::value;
            r16_interval<T, Gm, GENERAL, IU>(a, buf, win, cf, qv, rd, wr, b + IU);
        });
    }
    // outputs of the last interval (NB - 1, parity 1)
    if (a.last) {
#pragma unroll
        for (int k4 = 0; k4 < NQ; ++k4) {
            const int oc = (NB - 1) * S + GPW * k4 + r - Gm::DL;
            if (oc >= 0 && oc < n) {
                const T x = buf[NW][1][GPW * k4 + r][OL + j];
                a.A0[(int64_t)(oc - r) * lda + a.voff] = GENERAL ? x : fma(sigm1[oc], x, x);
            }
        }
    }
}

template <typename T, class Gm>
__global__ void __launch_bounds__(kLanes * Gm::NW)
    k_apply_r16(T* __restrict__ A, int64_t lda, int m, int n,
                const T* __restrict__ tab, const int* __restrict__ flag,
                const T* __restrict__ sigm1, int NB, int NBt)
{
    __shared__ T buf[Gm::NW + 1][2][Gm::S][kLanes];
    if (*flag != 0)
        r16_body<T, Gm, true>(A, lda, m, n, tab, sigm1, NB, NBt, buf);
    else
        r16_body<T, Gm, false>(A, lda, m, n, tab, sigm1, NB, NBt, buf);
}
// (R16_END)

// Interval counts of a 16- or 32-row geometry for n columns.
template <class Gm>
struct R16Dims {
    int NB, NBt;
    explicit R16Dims(int64_t n)
    {
        NB = (int)((n + Gm::DL + Gm::S - 1) / Gm::S);
        NB = (NB + Gm::U - 1) / Gm::U * Gm::U;  // whole unrolled loop bodies
        NBt = NB + 2;
    }
    size_t table_elems() const { return (size_t)Gm::NW * NBt * Gm::NR * kLanes; }
};

template <typename T, class Gm>
int run_r16(sep_handle* th, hipStream_t st, const ApplyMem<T>& mem, int64_t m, int64_t n, T* A,
            int64_t lda, const T* cs, int64_t ld_cs, int64_t k, bool rotations)
{
    constexpr int Kp = Gm::Kp;
    sep_handle* h = th;
    const R16Dims<Gm> rd(n);
    const int NB = rd.NB, NBt = rd.NBt;
    const size_t tab_elems = rd.table_elems();
    T* escale = mem.escale;
    T* sigm1 = mem.sigm1;
    int* flag = mem.flag;
    const int tp0 = timing_mark(h, st);
    const int64_t wgs = (m + Gm::RPG - 1) / Gm::RPG;
    int tp_prev = tp0;
    for (int64_t L0 = 0; L0 < k; L0 += Kp) {
        hipLaunchKernelGGL(k_colscale<T>, dim3((unsigned)((n + 63) / 64)), dim3(1024), 0, st, cs,
                           ld_cs, (int)n, (int)k, (int)L0, Kp, 0, rotations ? 0 : 1, mem.sync,
                           flag, escale, sigm1);
        SEP_CHECK_LAUNCH();
        const int tblocks = (int)std::min<int64_t>(((int64_t)tab_elems + 255) / 256, 8192);
        hipLaunchKernelGGL((k_coef_r16<T, Gm>), dim3(tblocks), dim3(256), 0, st, mem.table, cs,
                           ld_cs, (int)n, (int)k, (int)L0, NBt, flag, escale);
        SEP_CHECK_LAUNCH();
        const int ta = timing_mark(h, st);
        timing_pair(h, false, tp_prev, ta);
        hipLaunchKernelGGL((k_apply_r16<T, Gm>), dim3((unsigned)wgs), dim3(kLanes * Gm::NW), 0, st,
                           A, lda, (int)m, (int)n, mem.table, flag, sigm1, NB, NBt);
        SEP_CHECK_LAUNCH();
        tp_prev = timing_mark(h, st);
        timing_pair(h, true, ta, tp_prev);
    }
    return SEP_OK;
}

// 64 layers per pass: 4 waves of 4 groups of 4 layers, barrier intervals of 8, 16 or 24
// steps.  Longer intervals synchronise less often and take more LDS per workgroup (40, 80,
// 120 KB in Float64): they pay while the workgroups still fit the CUs at that size.
using G16Deep = G16<4, 4, 8>;
using G16Deep16 = G16<4, 4, 16>;
using G16Deep24 = G16<4, 4, 24>;

// The five geometries (SEP_APPLY names them wide, small, r16deep, r16s16, r16s24).
enum ApplyGeomId {
    kGeomWide = 2,
    kGeomSmall = 5,
    kGeomR16Deep = 7,
    kGeomR16S16 = 8,
    kGeomR16S24 = 10
};

// Interval length of the 16-row geometry, kernel time per apply, m = n, k = 64 (ms; one
// box, outputs bit for bit the same; r32 is the 32-row geometry G16<8, 4, 16, 32>):
//   Float64  S = 8   S = 16  S = 24  r32      Float32  S = 8   S = 16  S = 24  r32
//   2000     0.191   0.168   0.167            2000     0.169   0.150   0.152   0.224
//   3000     0.274   0.239   0.238
//   4096     0.381   0.326   0.317            4096     0.321   0.279   0.273   0.420
//   5000     0.530   0.497   0.690
//   6144     0.708   0.633   0.832   0.698    6144     0.528   0.496   0.509   0.621
//   8192     1.012   0.875   1.114   1.029    8192     0.726   0.680   0.688   0.863
//   9000     1.252   1.379           1.484
//   10240    1.408   1.630           1.674
//   12000    1.709   1.870   2.297            12000    1.364   1.239   1.487   1.572
//   14336                                     14336    1.863   1.826   2.150   1.878
// (wide, Float32: 1.612 at 14336, 1.838 at 16384.)  A 16-row workgroup takes 40 / 80 /
// 120 KB of LDS in Float64 (S = 8 / 16 / 24), half in Float32: S = 24 pays while the
// workgroups fit one per CU, S = 16 while they fit two (Float64) or four (Float32) per CU.
// r32 is no longer built.
// Earlier measurements, kernel time per apply, m = n, k = 64 (ms):
//            wide   deep   r16    r16deep r32    r32deep
//   2000     0.304  0.317  0.200  0.192   0.227  0.259
//   4096     0.589  0.611  0.383  0.373   0.429  0.494
//   8192     1.181  1.203  1.076  0.987   0.887  1.199
//   12000    1.824  1.766  2.169  1.715   1.916  2.133
//   16384    2.737  2.825  3.445  3.026   2.913  3.332
// The 64-row geometries leave CUs idle below m = 64 * 256 rows; the 16- and 32-row ones
// fill them but cost more per row (per-workgroup coefficient traffic, and workgroups that
// drift apart: at m = 16384 r16's workgroups take 1.9 to 3.4 ms).  Also measured at 16384:
// wide with Q prefetched three intervals ahead 2.57 ms; two passes of 32 layers with two
// workgroups per CU 3.11 ms (KG 8, 4 waves) and 3.55 ms (KG 4, 8 waves); neighbour-to-
// neighbour interval counters in LDS instead of the workgroup barrier 2.66 ms.  (deep,
// r16 and r32deep are no longer built.)  Where wide's time goes at 16384, from timing-only
// copies with one cost removed each (2.44-2.50 ms as built): the middle waves' coefficient
// rows not loaded 2.19-2.28 ms, or all coefficient rows read from LDS with no global loads
// 2.44 ms (the cost is the 512 bytes per wave and step written into VGPRs, whichever way
// they come); no Q loads or stores 2.08 ms; no barrier 2.32 ms, though its clock then
// drops (1.88 GHz): the apply runs at the power limit, so cycles saved partly return as
// a lower clock; no priority ramp 2.51 ms; no column-scale FMA in the first wave 2.44 ms.
// Float32,
// m = n, k = 64 (ms): 2000: wide 0.276, r16deep 0.164; 4096: 0.531, 0.309; 8192: 1.030,
// 0.700; 16384: 2.038, 2.191 (r32 0.22 / 0.42 / 0.88 / 2.15).
template <typename T>
int choose_geom(const sep_handle* h, int64_t m, int64_t k)
{
    if (h && h->apply_geom) return h->apply_geom;  // SEP_APPLY (tests force each in turn)
    if (k <= 32) return kGeomSmall;
    const int64_t cus = h && h->num_cus > 0 ? h->num_cus : 256;
    const int64_t blocks = (m + 15) / 16;  // 16-row workgroups
    if (sizeof(T) == 4) return m <= 12288 ? kGeomR16S16 : kGeomWide;
    if (blocks <= cus) return kGeomR16S24;
    if (blocks <= 2 * cus) return kGeomR16S16;
    if (m <= 14336) return kGeomR16Deep;
    return kGeomWide;
}

template <typename T>
size_t table_bytes(int geom, int64_t n)
{
    switch (geom) {
        case kGeomWide: return sizeof(T) * DppDims<GeomWide>(n).table_elems();
        case kGeomSmall: return sizeof(T) * DppDims<GeomSmall>(n).table_elems();
        case kGeomR16Deep: return sizeof(T) * R16Dims<G16Deep>(n).table_elems();
        case kGeomR16S16: return sizeof(T) * R16Dims<G16Deep16>(n).table_elems();
        default: return sizeof(T) * R16Dims<G16Deep24>(n).table_elems();
    }
}

template <typename T>
int launch_geom(int geom, sep_handle* th, hipStream_t st, const ApplyMem<T>& mem, int64_t m,
                int64_t n, T* A, int64_t lda, const T* cs, int64_t ld_cs, int64_t k,
                bool rotations)
{
    switch (geom) {
        case kGeomWide:
            return run_dpp<T, GeomWide>(th, st, mem, m, n, A, lda, cs, ld_cs, k, rotations);
        case kGeomSmall:
            return run_dpp<T, GeomSmall>(th, st, mem, m, n, A, lda, cs, ld_cs, k, rotations);
        case kGeomR16Deep:
            return run_r16<T, G16Deep>(th, st, mem, m, n, A, lda, cs, ld_cs, k, rotations);
        case kGeomR16S16:
            return run_r16<T, G16Deep16>(th, st, mem, m, n, A, lda, cs, ld_cs, k, rotations);
        default:
            return run_r16<T, G16Deep24>(th, st, mem, m, n, A, lda, cs, ld_cs, k, rotations);
    }
}

int check_apply_range(int64_t m, int64_t n)
{
    if (m > INT32_MAX || n > (INT32_MAX / 4)) {
        set_error("bulk apply: m=%lld n=%lld beyond the 32-bit index range", (long long)m,
                  (long long)n);
        return SEP_ERR_ARG;
    }
    return SEP_OK;
}

constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kSyncBytes = 256;  // k_colscale's sync words at the head of the scratch

}  // namespace

template <typename T>
int bulk_apply_dev(sep_handle* h, hipStream_t st, ApplyWork& w, int64_t m, int64_t n, T* A,
                   int64_t lda, const T* cs, int64_t ld_cs, int64_t k, bool rotations)
{
    if (m <= 0 || n <= 1 || k <= 0) return SEP_OK;
    if (int rc = check_apply_range(m, n)) return rc;
    const int geom = choose_geom<T>(h, m, k);
    if (int rc = w.table.reserve(table_bytes<T>(geom, n))) return rc;
    if (int rc = w.aux.reserve(aux_bytes(n, sizeof(T)))) return rc;
    if (!w.sync.p) {
        if (int rc = w.sync.reserve(256)) return rc;
        hipLaunchKernelGGL(k_zero_sync, dim3(1), dim3(64), 0, st, w.sync.as<unsigned>());
        SEP_CHECK_LAUNCH();
    }
    return launch_geom<T>(geom, h, st, apply_mem<T>(w.table.p, w.aux.p, n, w.sync.p), m, n, A,
                          lda, cs, ld_cs, k, rotations);
}

template int bulk_apply_dev<double>(sep_handle*, hipStream_t, ApplyWork&, int64_t, int64_t,
                                    double*, int64_t, const double*, int64_t, int64_t, bool);
template int bulk_apply_dev<float>(sep_handle*, hipStream_t, ApplyWork&, int64_t, int64_t,
                                   float*, int64_t, const float*, int64_t, int64_t, bool);

// ---- the allocation-free apply --------------------------------------------------------
// Scratch of sep_bulk_rmul_f64_dev_noalloc: its coefficient table and auxiliary words live
// in this file-scope device array (one copy per GPU, held until the process ends) instead
// of handle buffers.  In host code `scratch` only names the symbol: kernels get the
// per-device address that hipGetSymbolAddress returns, recorded by scratch_init.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error("sep_create: device %d beyond the %d devices of the scratch table", device,
                  kMaxDevices);
        return SEP_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return SEP_OK;
    void* p = nullptr;
    SEP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the current device's copy
    if (!p) {
        set_error("hipGetSymbolAddress(scratch) returned a null address on device %d", device);
        return SEP_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
    return SEP_OK;
}

int bulk_apply_noalloc_f64(sep_handle* h, int64_t m, int64_t n, double* A, int64_t lda,
                           const double* cs, int64_t k)
{
    if (m <= 0 || n <= 1 || k <= 0) return SEP_OK;
    if (int rc = check_apply_range(m, n)) return rc;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = (h->device >= 0 && h->device < kMaxDevices) ? g_scratch_ptr[h->device] : nullptr;
    }
    if (!base) {
        set_error("sep_bulk_rmul_f64_dev_noalloc: no scratch address for device %d", h->device);
        return SEP_ERR_HIP;
    }
    const int geom = choose_geom<double>(h, m, k);
    // layout: k_colscale's sync words (zero since the array was loaded, and left zero by
    // every launch), the coefficient table, the auxiliary words
    const size_t tb = round256(table_bytes<double>(geom, n));
    const size_t need = kSyncBytes + tb + aux_bytes(n, sizeof(double));
    if (need > kNoAllocScratchBytes) {
        set_error("sep_bulk_rmul_f64_dev_noalloc: n=%lld needs %zu bytes of scratch, above the "
                  "fixed %zu (use sep_bulk_rmul_f64_dev)",
                  (long long)n, need, (size_t)kNoAllocScratchBytes);
        return SEP_ERR_ARG;
    }
    unsigned char* b = static_cast<unsigned char*>(base) + kSyncBytes;
    // no timing events: the instrumentation may create them
    return launch_geom<double>(geom, nullptr, h->stream, apply_mem<double>(b, b + tb, n, base), m,
                               n, A, lda, cs, n - 1, k, false);
}


}  // namespace sep

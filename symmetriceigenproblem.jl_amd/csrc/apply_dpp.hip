// apply_dpp.hip -- Float64 bulk Givens application with DPP-broadcast coefficients.
//
// Same product as bulk_wave_order_2x2_rmul! (src/wave.jl:62-162): A <- A * prod_L prod_i
// G(i,L)^T, each rotation a' = c a + s b, b' = c b - s a on the column pair (i, i+1)
// (Rotation2 rmul!, src/simple_givens.jl:53-64), and the same systolic organisation as the
// SGPR-fed kernel in apply.hip (one row per lane, G waves of KG layers each, columns
// handed from wave to wave through a double-buffered LDS ring, wave order t = i + 2L).
//
// What differs is how the wave-uniform rotation data reach the FMA units:
//  * Coefficients are not scalar-loaded (scalar loads return out of order, so every use
//    waits for all of them, exposing L2 latency each step).  Instead one 64-bit VGPR holds
//    16 coefficients (2*KG per step, so one or two steps) in lanes 0..15 of each 16-lane
//    row, loaded with ordinary vector loads one whole interval ahead, and every FMA takes
//    its coefficient through the DPP64 `row_newbcast:j` source modifier, which broadcasts
//    lane j of each row.
//  * A single wave issues FP64 FMAs at well under the SIMD's rate (measured: ~43% of peak
//    with one wave per SIMD, ~70% with two, ~76% with four, ~95% with sixteen).  Two
//    geometries are built: "wide" (G = 8 waves of KG = 8 layers, S = 16 steps per barrier
//    interval, two waves per SIMD) and "deep" (G = 16 waves of KG = 4 layers, S = 8, four
//    waves per SIMD); wide is the default because deep's doubled barrier count costs more
//    than its occupancy gains.  The first wave prefetches Q two or three intervals ahead.
//  * Only the middle waves load coefficients (for every wave, one interval ahead, staged
//    through LDS); the first wave issues only Q loads and the last only Q stores, so no
//    wave ever has to wait for a memory operation it issued for a later interval.
//  * Rotations with c^2 + s^2 = 1 (all rotations the QR chase produces, and the benchmark's
//    cos/sin pairs) are applied as three shears, a += p b; b += q a; a += p b with
//    p = s~/(1 + c~) = tan(phi/2), q = -s~: three in-place FMAs per rotation and row instead
//    of two multiplies and two FMAs.  To keep |p| <= 1 the rotation is taken with c~ = |c|;
//    the sign this drops is tracked per column (each rotation with c < 0 negates both of
//    its columns in the stored representation), folded into s~, and removed when a column
//    leaves the pass.  The column signs are prefix products over the layers, computed by a
//    small per-column kernel.
//  * Grids that are not rotations (the reference benchmark's (rand, rand) pairs) take a
//    five-instruction path, t = a; a += (c-1) a; a += s b; b += (c-1) b; b -= s t, chosen
//    on the device by a normalisation check, so the entry point accepts any (c, s) as the
//    reference's kernels do.
#include <type_traits>

#include "sep_internal.hpp"

namespace sep {
namespace {

constexpr int kLanes = 64;

// flag = 1 if any rotation of the grid is not normalised.
__global__ void __launch_bounds__(256) k_check_norm(const double* __restrict__ cs, int64_t ld,
                                                    int n, int k, int* __restrict__ flag)
{
    const int64_t total = (int64_t)(n - 1) * k;
    int bad = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = q % (n - 1), L = q / (n - 1);
        const double c = cs[2 * (i + ld * L)], s = cs[2 * (i + ld * L) + 1];
        bad |= !(fabs(c * c + s * s - 1.0) <= 1e-14);
    }
    if (__any(bad) && (threadIdx.x % kLanes) == 0) atomicOr(flag, 1);
}

__device__ __forceinline__ bool neg_c(const double* __restrict__ cs, int64_t ld, int n, int64_t i,
                                      int L, int k)
{
    if (i < 0 || i > n - 2 || L >= k) return false;
    return cs[2 * (i + ld * L)] < 0.0;
}

// Per column j: bit lr of mask[j] = the stored sign of column j is negative before the
// rotations of layer L0 + lr; sigm1[j] = (final sign) - 1, i.e. 0 or -2.
__global__ void __launch_bounds__(256) k_colsigns(const double* __restrict__ cs, int64_t ld,
                                                  int n, int k, int L0, int Kp,
                                                  const int* __restrict__ flag,
                                                  uint64_t* __restrict__ mask,
                                                  double* __restrict__ sigm1)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (*flag) {
        mask[j] = 0;
        sigm1[j] = 0.0;
        return;
    }
    bool neg = false;
    uint64_t m = 0;
    for (int lr = 0; lr < Kp; ++lr) {
        const int L = L0 + lr;
        if (L >= k) break;
        if (neg) m |= 1ull << lr;
        // rotations (j-1, L) then (j, L) touch column j in layer L
        neg ^= neg_c(cs, ld, n, j - 1, L, k) ^ neg_c(cs, ld, n, j, L, k);
    }
    mask[j] = m;
    sigm1[j] = neg ? -2.0 : 0.0;
}

// Coefficient table of one pass: tab[g][u][2 KG] for g < G, u = t + S + PADU (time t,
// PADU rows of padding before and enough after that every interval of every wave reads
// inside the table).  Slot l of wave g at time t is rotation i = t - 2 (g KG + l) of layer
// L0 + g KG + l; [2l] = p, [2l+1] = q (shear form) or c - 1, s (general form);
// identity / out-of-grid slots hold (0, 0), which is the identity in both forms.
// Section g = G holds, for every interval bt, the last wave's 16-lane output-sign row at
// the position a coefficient row of a (virtual) wave G would have, so the coefficient
// loaders fetch it with the same address pattern.
template <int KG, int G, int S>
__global__ void __launch_bounds__(256) k_coef_table(double* __restrict__ tab,
                                                    const double* __restrict__ cs, int64_t ld,
                                                    int n, int k, int L0, int Up, int PADU,
                                                    int NI, const int* __restrict__ flag,
                                                    const uint64_t* __restrict__ mask,
                                                    const double* __restrict__ sigm1)
{
    constexpr int NC = 2 * KG, Kp = KG * G;
    const int64_t rot_slots = (int64_t)G * Up * KG;
    const int64_t total = rot_slots + (int64_t)NI * 16;
    const int general = *flag;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        if (q >= rot_slots) {  // output-sign section
            const int qq = (int)(q - rot_slots);
            const int bt = qq / 16, c = qq % 16;
            const int tl = (bt - (G - 1)) * S + (G - 1);
            const int col = tl - 2 * Kp + 2 + c;
            const double v = (c < S && col >= 0 && col < n && !general) ? sigm1[col] : 0.0;
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
        double p = 0.0, qq = 0.0;
        if (L < k && i >= 0 && i <= n - 2) {
            const double c = cs[2 * (i + ld * L)], s = cs[2 * (i + ld * L) + 1];
            if (general) {
                p = c - 1.0;
                qq = s;
            } else {
                bool na = (mask[i] >> lr) & 1;
                na ^= neg_c(cs, ld, n, i - 1, L, k);
                const bool nb = (mask[i + 1] >> lr) & 1;
                const double ct = fabs(c);
                const double st = ((c < 0.0) ^ na ^ nb) ? -s : s;
                p = st / (1.0 + ct);
                qq = -st;
            }
        }
        tab[2 * q] = p;
        tab[2 * q + 1] = qq;
    }
}

// ---- the per-step FMA blocks ------------------------------------------------------
// Operand 2l / 2l+1 of each asm is window slot (s - 2l) / (s - 2l + 1) mod 2KG (the column
// pair of rotation l); the last operand is the coefficient row.  Lane numbers are the
// DPP row_newbcast sources.  The leading s_nop covers the VALU-write -> DPP-read hazard
// in case the compiler copied the coefficient row with a VALU move right before the block.
#define SEP_DPP(a, c, b, lane) \
    "v_fmac_f64_dpp " a ", " c ", " b " row_newbcast:" #lane " row_mask:0xf bank_mask:0xf\n"
#define SEP_ROT3(a, b, p, q) /* a += p b; b += q a; a += p b (three shears) */

#define SEP_SHEAR8                                                                             \
    asm("s_nop 1\n"                                                                            \
        SEP_DPP("%0", "%16", "%1", 0) SEP_DPP("%2", "%16", "%3", 2) SEP_DPP("%4", "%16", "%5", 4)  \
        SEP_DPP("%6", "%16", "%7", 6) SEP_DPP("%8", "%16", "%9", 8)                                \
        SEP_DPP("%10", "%16", "%11", 10) SEP_DPP("%12", "%16", "%13", 12)                        \
        SEP_DPP("%14", "%16", "%15", 14)                                                       \
        SEP_DPP("%1", "%16", "%0", 1) SEP_DPP("%3", "%16", "%2", 3) SEP_DPP("%5", "%16", "%4", 5)  \
        SEP_DPP("%7", "%16", "%6", 7) SEP_DPP("%9", "%16", "%8", 9)                                \
        SEP_DPP("%11", "%16", "%10", 11) SEP_DPP("%13", "%16", "%12", 13)                        \
        SEP_DPP("%15", "%16", "%14", 15)                                                       \
        SEP_DPP("%0", "%16", "%1", 0) SEP_DPP("%2", "%16", "%3", 2) SEP_DPP("%4", "%16", "%5", 4)  \
        SEP_DPP("%6", "%16", "%7", 6) SEP_DPP("%8", "%16", "%9", 8)                                \
        SEP_DPP("%10", "%16", "%11", 10) SEP_DPP("%12", "%16", "%13", 12)                        \
        SEP_DPP("%14", "%16", "%15", 14)                                                       \
        : "+v"(SLOT(0, 0)), "+v"(SLOT(0, 1)), "+v"(SLOT(1, 0)), "+v"(SLOT(1, 1)),              \
          "+v"(SLOT(2, 0)), "+v"(SLOT(2, 1)), "+v"(SLOT(3, 0)), "+v"(SLOT(3, 1)),              \
          "+v"(SLOT(4, 0)), "+v"(SLOT(4, 1)), "+v"(SLOT(5, 0)), "+v"(SLOT(5, 1)),              \
          "+v"(SLOT(6, 0)), "+v"(SLOT(6, 1)), "+v"(SLOT(7, 0)), "+v"(SLOT(7, 1))               \
        : "v"(cf))

#define SEP_SHEAR4(P0, Q0, P1, Q1, P2, Q2, P3, Q3)                                             \
    asm("s_nop 1\n"                                                                            \
        SEP_DPP("%0", "%8", "%1", P0) SEP_DPP("%2", "%8", "%3", P1) SEP_DPP("%4", "%8", "%5", P2)  \
        SEP_DPP("%6", "%8", "%7", P3)                                                          \
        SEP_DPP("%1", "%8", "%0", Q0) SEP_DPP("%3", "%8", "%2", Q1) SEP_DPP("%5", "%8", "%4", Q2)  \
        SEP_DPP("%7", "%8", "%6", Q3)                                                          \
        SEP_DPP("%0", "%8", "%1", P0) SEP_DPP("%2", "%8", "%3", P1) SEP_DPP("%4", "%8", "%5", P2)  \
        SEP_DPP("%6", "%8", "%7", P3)                                                          \
        : "+v"(SLOT(0, 0)), "+v"(SLOT(0, 1)), "+v"(SLOT(1, 0)), "+v"(SLOT(1, 1)),              \
          "+v"(SLOT(2, 0)), "+v"(SLOT(2, 1)), "+v"(SLOT(3, 0)), "+v"(SLOT(3, 1))               \
        : "v"(cf))

// Shear step: KG rotations, coefficients in lanes base .. base + 2KG - 1.
template <int KG, int s, int base>
__device__ __forceinline__ void step_shear(double (&w)[2 * KG], double cf)
{
    constexpr int W = 2 * KG;
#define SLOT(l, o) w[(((s) - 2 * (l) + (o)) % W + W) % W]
    if constexpr (KG == 8) {
        static_assert(base == 0, "KG = 8 fills a whole coefficient row");
        SEP_SHEAR8;
    } else {
        static_assert(KG == 4, "KG is 4 or 8");
        if constexpr (base == 0)
            SEP_SHEAR4(0, 1, 2, 3, 4, 5, 6, 7);
        else
            SEP_SHEAR4(8, 9, 10, 11, 12, 13, 14, 15);
    }
#undef SLOT
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

template <int quad>
__device__ __forceinline__ void gen2(double& a0, double& b0, double& a1, double& b1, double cf)
{
    const double t0 = a0, t1 = a1;
    if constexpr (quad == 0) SEP_GEN2(0, 1, 2, 3);
    else if constexpr (quad == 1) SEP_GEN2(4, 5, 6, 7);
    else if constexpr (quad == 2) SEP_GEN2(8, 9, 10, 11);
    else SEP_GEN2(12, 13, 14, 15);
}

template <int KG, int s, int base>
__device__ __forceinline__ void step_general(double (&w)[2 * KG], double cf)
{
    constexpr int W = 2 * KG;
#define SLOT(l, o) w[(((s) - 2 * (l) + (o)) % W + W) % W]
    gen2<base / 4 + 0>(SLOT(0, 0), SLOT(0, 1), SLOT(1, 0), SLOT(1, 1), cf);
    gen2<base / 4 + 1>(SLOT(2, 0), SLOT(2, 1), SLOT(3, 0), SLOT(3, 1), cf);
    if constexpr (KG == 8) {
        gen2<2>(SLOT(4, 0), SLOT(4, 1), SLOT(5, 0), SLOT(5, 1), cf);
        gen2<3>(SLOT(6, 0), SLOT(6, 1), SLOT(7, 0), SLOT(7, 1), cf);
    }
#undef SLOT
}

// Output sign fix of the last wave: x += (sigma - 1) x, with (sigma - 1) in {0, -2}
// broadcast from lane s of the interval's sign row (exact: the result is x or -x).
#define SEP_FIX(N) asm("s_nop 1\n" SEP_DPP("%0", "%1", "%0", N) : "+v"(x) : "v"(sg))

template <int s>
__device__ __forceinline__ void fix_sign(double& x, double sg)
{
    static_assert(s >= 0 && s < 16, "sign row has 16 lanes");
    if constexpr (s == 0) SEP_FIX(0);
    else if constexpr (s == 1) SEP_FIX(1);
    else if constexpr (s == 2) SEP_FIX(2);
    else if constexpr (s == 3) SEP_FIX(3);
    else if constexpr (s == 4) SEP_FIX(4);
    else if constexpr (s == 5) SEP_FIX(5);
    else if constexpr (s == 6) SEP_FIX(6);
    else if constexpr (s == 7) SEP_FIX(7);
    else if constexpr (s == 8) SEP_FIX(8);
    else if constexpr (s == 9) SEP_FIX(9);
    else if constexpr (s == 10) SEP_FIX(10);
    else if constexpr (s == 11) SEP_FIX(11);
    else if constexpr (s == 12) SEP_FIX(12);
    else if constexpr (s == 13) SEP_FIX(13);
    else if constexpr (s == 14) SEP_FIX(14);
    else SEP_FIX(15);
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

template <class Gm, bool FIRST, bool LAST, bool CHECKED, bool GENERAL, int D>
__device__ __forceinline__ void interval(double* __restrict__ Ar, int64_t lda, int n,
                                         double (&win)[Gm::W], double (&pf)[FIRST ? Gm::PFD : 1][Gm::S],
                                         const double (&cf)[Gm::NR], double sg,
                                         const double* lds_in, double* lds_out, int t0, int g)
{
    constexpr int S = Gm::S, W = Gm::W, KG = Gm::KG, PFD = Gm::PFD;
    double in[FIRST ? 1 : S];
    // first wave: prefetch slot D holds this interval's columns (loaded PFD intervals
    // ago); each is refilled, right after its step consumed it, with the column PFD
    // intervals ahead (t0 + PFD S + 1 + s)
    const int c0f = t0 + PFD * S + 1;
    const double* __restrict__ pfp = Ar + (int64_t)c0f * lda;
    if constexpr (!FIRST) {
        const int c0 = t0 - 2 * g * KG + 1;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if constexpr (CHECKED)
                in[s] = (c0 + s >= 0 && c0 + s < n) ? lds_in[s * kLanes] : 0.0;
            else
                in[s] = lds_in[s * kLanes];
        }
    }
    const int o0 = t0 - 2 * Gm::Kp + 2;
    double* __restrict__ po = Ar + (int64_t)o0 * lda;
    Steps<S, 0>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int row = s / Gm::SPR, base = (s % Gm::SPR) * Gm::NC;
        if constexpr (FIRST) {
            win[(s + 1) % W] = pf[D][s];
            if constexpr (CHECKED)
                pf[D][s] = (c0f + s < n) ? pfp[(int64_t)s * lda] : 0.0;
            else
                pf[D][s] = pfp[(int64_t)s * lda];
        } else {
            win[(s + 1) % W] = in[s];
        }
        if constexpr (GENERAL)
            step_general<KG, s, base>(win, cf[row]);
        else
            step_shear<KG, s, base>(win, cf[row]);
        double out = win[(s + 2) % W];
        if constexpr (LAST) {
            if constexpr (!GENERAL) fix_sign<s>(out, sg);
            if constexpr (CHECKED) {
                if (o0 + s >= 0 && o0 + s < n) po[(int64_t)s * lda] = out;
            } else {
                po[(int64_t)s * lda] = out;
            }
        } else {
            lds_out[s * kLanes] = out;
        }
    });
}

// LDS layout of one interval's coefficient block: rows [g][r][16] for every wave, then
// the last wave's 16-lane output-sign row.
template <class Gm>
struct CoefBlock {
    static constexpr int ROWS = Gm::G * Gm::NR * 16;
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
            gq = q / (NR * 16);
            r = (q / 16) % NR;
            c = q % 16;
        } else if (q < CB::SIZE) {
            gq = G;
            c = q - CB::ROWS;
        }
        off[j] = (q < CB::SIZE) ? (gq * Up + gq * (1 - S) + S + PADU) * NC + r * 16 + c : 0;
    }
}

// The middle waves fetch the next interval's coefficient block (every wave's rows and
// the output signs) from the table into registers at the start of an interval and store
// it to LDS at the end; they issue no other global memory operations, so waiting for
// these loads never drains the first wave's Q prefetch or the last wave's stores.
template <class Gm>
__device__ __forceinline__ void fetch_block(double (&v)[CoefBlock<Gm>::PER],
                                            const int (&off)[CoefBlock<Gm>::PER],
                                            const double* __restrict__ tab, int bt)
{
    const double* __restrict__ base = tab + (int64_t)bt * (Gm::S * Gm::NC);
#pragma unroll
    for (int j = 0; j < CoefBlock<Gm>::PER; ++j) v[j] = base[off[j]];
}

template <class Gm>
__device__ __forceinline__ void store_block(double* cb, const double (&v)[CoefBlock<Gm>::PER],
                                            int loader, int lane)
{
    using CB = CoefBlock<Gm>;
#pragma unroll
    for (int j = 0; j < CB::PER; ++j) {
        const int q = loader * kLanes + lane + j * CB::LOADERS * kLanes;
        if (q < CB::SIZE) cb[q] = v[j];
    }
}

template <class Gm, bool FIRST, bool LAST, bool GENERAL>
__device__ __forceinline__ void wave(double* __restrict__ A, int64_t lda, int m, int n,
                                     const double* __restrict__ tab, int Up, int PADU,
                                     int T_out, int NI,
                                     double (*ring)[2][Gm::S][kLanes],
                                     double (*cbuf)[CoefBlock<Gm>::SIZE], int g, int lane)
{
    using CB = CoefBlock<Gm>;
    constexpr int S = Gm::S, W = Gm::W, KG = Gm::KG, PFD = Gm::PFD, NR = Gm::NR;
    constexpr bool LOADER = !FIRST && !LAST;
    const int loader = g - 1;
    // Lanes past the last row alias row m-1: they load, compute and store exactly the
    // values of that row's own lane, so no per-lane predication is needed.
    const int64_t row = min((int64_t)blockIdx.x * kLanes + lane, (int64_t)m - 1);
    double* __restrict__ Ar = A + row;
    const int cl = lane & 15;  // lane within the 16-lane DPP row

    double win[W];
#pragma unroll
    for (int w = 0; w < W; ++w) win[w] = 0.0;
    double pf[FIRST ? PFD : 1][S];
    if constexpr (FIRST) {
        // column 0 enters before step 0 (slot 0); interval d takes columns dS+1 .. dS+S
        win[0] = Ar[0];
#pragma unroll
        for (int d = 0; d < PFD; ++d)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int col = d * S + s + 1;
                pf[d][s] = (col < n) ? Ar[(int64_t)col * lda] : 0.0;
            }
    }
    double blk[CB::PER];
    int off[CB::PER];
    if constexpr (LOADER) {  // block of interval 0
        block_offsets<Gm>(off, Up, PADU, loader, lane);
        fetch_block<Gm>(blk, off, tab, 0);
        store_block<Gm>(cbuf[0], blk, loader, lane);
    }
    lds_barrier();
    // the interval loop is unrolled by the prefetch depth so that prefetch slot D is
    // consumed and refilled in place (no register shuffling of in-flight loads)
    for (int b0 = 0; b0 < NI; b0 += PFD) {
        Steps<PFD, 0>::run([&](auto dc) {
            constexpr int D = decltype(dc)::value;
            const int b = b0 + D;
            if (b >= NI) return;
            if constexpr (LOADER) {
                if (b + 1 < NI) fetch_block<Gm>(blk, off, tab, b + 1);
            }
            const int t0 = (b - g) * S + g;
            const bool active = (t0 + S > 0) && (t0 < T_out);
            if (active) {
                const double* cb = cbuf[b & 1];
                double cf[NR];
#pragma unroll
                for (int r = 0; r < NR; ++r) cf[r] = cb[(g * NR + r) * 16 + cl];
                const double sg = LAST ? cb[CB::ROWS + cl] : 0.0;
                const double* lds_in =
                    FIRST ? nullptr : &ring[g > 0 ? g - 1 : 0][(b - 1) & 1][0][lane];
                double* lds_out = LAST ? nullptr : &ring[g][b & 1][0][lane];
                bool interior;
                if constexpr (FIRST) {
                    interior = t0 + (PFD + 1) * S < n;
                } else {
                    const int c0 = t0 - 2 * g * KG + 1;
                    interior = c0 >= 0 && c0 + S <= n;
                }
                if constexpr (LAST) {
                    const int o0 = t0 - 2 * Gm::Kp + 2;
                    interior = interior && o0 >= 0 && o0 + S <= n;
                }
                if (interior)
                    interval<Gm, FIRST, LAST, false, GENERAL, D>(Ar, lda, n, win, pf, cf, sg,
                                                                 lds_in, lds_out, t0, g);
                else
                    interval<Gm, FIRST, LAST, true, GENERAL, D>(Ar, lda, n, win, pf, cf, sg,
                                                                lds_in, lds_out, t0, g);
            }
            if constexpr (LOADER) {
                if (b + 1 < NI) store_block<Gm>(cbuf[(b + 1) & 1], blk, loader, lane);
            }
            lds_barrier();
        });
    }
}

template <class Gm>
__global__ void __launch_bounds__(kLanes * Gm::G)
    k_apply_dpp(double* __restrict__ A, int64_t lda, int m, int n,
                const double* __restrict__ tab, const int* __restrict__ flag, int Up, int PADU,
                int T_out, int NI)
{
    constexpr int G = Gm::G;
    static_assert(G >= 3, "the coefficient loaders are the middle waves");
    __shared__ double ring[G - 1][2][Gm::S][kLanes];
    __shared__ double cbuf[2][CoefBlock<Gm>::SIZE];
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int lane = threadIdx.x % kLanes;
    const bool general = *flag != 0;
#define SEP_WAVE(GEN)                                                                         \
    if (g == 0)                                                                               \
        wave<Gm, true, false, GEN>(A, lda, m, n, tab, Up, PADU, T_out, NI, ring, cbuf, g,     \
                                   lane);                                                     \
    else if (g == G - 1)                                                                      \
        wave<Gm, false, true, GEN>(A, lda, m, n, tab, Up, PADU, T_out, NI, ring, cbuf, g,     \
                                   lane);                                                     \
    else                                                                                      \
        wave<Gm, false, false, GEN>(A, lda, m, n, tab, Up, PADU, T_out, NI, ring, cbuf, g,    \
                                    lane);
    if (general) {
        SEP_WAVE(true)
    } else {
        SEP_WAVE(false)
    }
#undef SEP_WAVE
}

template <class Gm>
int run_dpp(sep_handle* h, hipStream_t st, ApplyWork& w, int64_t m, int64_t n, double* A,
            int64_t lda, const double* cs, int64_t ld_cs, int64_t k, bool rotations)
{
    constexpr int G = Gm::G, S = Gm::S, KG = Gm::KG, Kp = Gm::Kp;
    const int T_out = (int)(n + 2 * Kp - 2);
    const int NI = G + (T_out + S - 1) / S + 1;
    const int PADU = (G - 1) * (S - 1);
    const int Up = NI * S + 2 * S + PADU;
    if (int rc = w.table.reserve(sizeof(double) * ((size_t)(G + 1) * Up * Gm::NC + 64))) return rc;
    if (int rc = w.aux.reserve(sizeof(uint64_t) * (size_t)n + sizeof(double) * (size_t)n + 256))
        return rc;
    uint64_t* mask = w.aux.as<uint64_t>();
    double* sigm1 = reinterpret_cast<double*>(mask + n);
    int* flag = reinterpret_cast<int*>(sigm1 + n);
    const int tp0 = timing_mark(h, st);
    SEP_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
    if (!rotations) {
        const int64_t tot = (n - 1) * k;
        const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 2048);
        hipLaunchKernelGGL(k_check_norm, dim3(blocks), dim3(256), 0, st, cs, ld_cs, (int)n,
                           (int)k, flag);
        SEP_CHECK_LAUNCH();
    }
    const int64_t rb = (m + kLanes - 1) / kLanes;
    int tp_prev = tp0;
    for (int64_t L0 = 0; L0 < k; L0 += Kp) {
        hipLaunchKernelGGL(k_colsigns, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cs,
                           ld_cs, (int)n, (int)k, (int)L0, Kp, flag, mask, sigm1);
        SEP_CHECK_LAUNCH();
        const int64_t total = (int64_t)G * Up * KG + (int64_t)NI * 16;
        const int tblocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL((k_coef_table<KG, G, S>), dim3(tblocks), dim3(256), 0, st,
                           w.table.as<double>(), cs, ld_cs, (int)n, (int)k, (int)L0, Up, PADU,
                           NI, flag, mask, sigm1);
        SEP_CHECK_LAUNCH();
        const int ta = timing_mark(h, st);
        timing_pair(h, false, tp_prev, ta);
        hipLaunchKernelGGL((k_apply_dpp<Gm>), dim3((unsigned)rb), dim3(kLanes * G), 0, st, A,
                           lda, (int)m, (int)n, w.table.as<double>(), flag, Up, PADU, T_out, NI);
        SEP_CHECK_LAUNCH();
        tp_prev = timing_mark(h, st);
        timing_pair(h, true, ta, tp_prev);
    }
    return SEP_OK;
}

using GeomDeep = Geom<4, 16, 8, 3>;   // 64 layers per pass, 4 waves per SIMD
using GeomWide = Geom<8, 8, 16, 2>;   // 64 layers per pass, 2 waves per SIMD
using GeomSmall = Geom<4, 8, 8, 3>;   // 32 layers per pass (short grids)

}  // namespace

int bulk_apply_dpp_f64(sep_handle* h, hipStream_t st, ApplyWork& w, int64_t m, int64_t n,
                       double* A, int64_t lda, const double* cs, int64_t ld_cs, int64_t k,
                       bool rotations)
{
    // measured at m = n = 16384, k = 64 (kernel time per apply): wide 2.57 ms; deep 2.78 ms
    // (more barriers among 16 waves outweigh the higher occupancy); wide with Q prefetched
    // three intervals ahead 2.57 ms; two passes of 32 layers with two workgroups per CU
    // 3.11 ms (KG 8, 4 waves) and 3.55 ms (KG 4, 8 waves)
    const int geom = h ? h->apply_geom : 0;
    if (geom == 1) return run_dpp<GeomDeep>(h, st, w, m, n, A, lda, cs, ld_cs, k, rotations);
    if (geom == 5) return run_dpp<GeomSmall>(h, st, w, m, n, A, lda, cs, ld_cs, k, rotations);
    if (k <= 32) return run_dpp<GeomSmall>(h, st, w, m, n, A, lda, cs, ld_cs, k, rotations);
    return run_dpp<GeomWide>(h, st, w, m, n, A, lda, cs, ld_cs, k, rotations);
}

}  // namespace sep

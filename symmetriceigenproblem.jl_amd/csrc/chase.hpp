// chase.hpp -- device-side Wilkinson-shift bulge chasing (one wavefront per problem).
//
// Restates, as wave-level device code, the inner loop of qr_algorithm!
// (src/bulge_chasing.jl:178-276): deflation scan, block-capacity check, scaled
// Wilkinson shift and the single_shift! sweep (src/bulge_chasing.jl:11-144) that
// records each rotation at layer ++layer_thickness[i] of the (c,s) grid.
//
// The sweep is a serial recurrence; the wavefront's 64 lanes are used around it:
//  * the deflation scan and the capacity reduction test 64 indices per step (ballot /
//    lane max);
//  * a sweep is processed in chunks of 128 steps (64 in the batched chase): lane j preloads
//    dv[i+1], ev[i+1] and layer_thickness[i] of steps i = i0 + j, i0 + 64 + j (values no
//    earlier step of the sweep changes) for the next chunk while the current one runs and
//    stages them in LDS; the recurrence reads them from there two steps ahead and writes
//    its finished outputs (ev[i-1], dv[i], c, s) to LDS, both on the LDS pipe instead of
//    the vector ALU, and the chunk is stored with one coalesced store per array.  The
//    recurrence itself therefore never waits on memory;
//  * in Float64 the rotation's range test (extreme magnitudes take the complete DLARTG)
//    can be deferred to the end of a chunk: each step folds its operands' exponent words
//    into a running maximum, and a chunk that left the common range is run again from its
//    entry state with the per-step test (single solves defer; the batched chase, where
//    vector issue is shared by several waves, keeps the per-step scalar test);
//  * single solves up to a few thousand indices run with dv, ev and layer_thickness in
//    LDS (the kernel passes LDS pointers in the Problem view).
// Arithmetic follows the reference operation by operation without contraction, and
// uses IEEE division and square root, so the grid, dv, ev and iteration count equal
// the C restatement's bit for bit -- except under the opt-in fast rotation (kFast, see
// givens_fast), whose results agree with the reference to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace sep {
namespace chase {

constexpr int kWave = 64;

template <typename T>
struct Consts;
template <>
struct Consts<double> {
    static constexpr double safmn2 = 0x1p-485;
    static constexpr double safmx2 = 0x1p485;
};
template <>
struct Consts<float> {
    static constexpr float safmn2 = 0x1p-51f;
    static constexpr float safmx2 = 0x1p51f;
};

__device__ __forceinline__ double tabs(double x) { return __builtin_fabs(x); }

// One (c, s) pair as a single 16-byte (8-byte) store.
__device__ __forceinline__ void store_pair(double* p, double c, double s)
{
    *reinterpret_cast<double2*>(p) = make_double2(c, s);
}
__device__ __forceinline__ void store_pair(float* p, float c, float s)
{
    *reinterpret_cast<float2*>(p) = make_float2(c, s);
}
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }

#pragma clang fp contract(off)

template <typename T>
struct Rot {
    T c, s, r;
};

// LinearAlgebra.givensAlgorithm (classic DLARTG), via get_rotation
// (src/simple_givens.jl:20-23).  g == 0 -> (1,0,f); f == 0 -> (0,1,g); otherwise
// c = f/r, s = g/r, r = sqrt(f^2+g^2), negated when |f| > |g| and c < 0, with
// power-of-two rescaling when max(|f|,|g|) leaves (2^-485, 2^485) (Float64).
// This complete version serves the rare cases of givens() below.
template <typename T>
__device__ __attribute__((noinline)) Rot<T> givens_slow(T f, T g)
{
    const T safmn2 = Consts<T>::safmn2, safmx2 = Consts<T>::safmx2;
    Rot<T> o;
    if (g == T(0)) return {T(1), T(0), f};
    if (f == T(0)) return {T(0), T(1), g};
    T f1 = f, g1 = g;
    T scale = tabs(f1) > tabs(g1) ? tabs(f1) : tabs(g1);
    int count = 0;
    if (scale >= safmx2) {
        do {
            count++;
            f1 *= safmn2;
            g1 *= safmn2;
            scale = tabs(f1) > tabs(g1) ? tabs(f1) : tabs(g1);
        } while (scale >= safmx2 && count < 20);
        o.r = sqrt(f1 * f1 + g1 * g1);
        o.c = f1 / o.r;
        o.s = g1 / o.r;
        for (int i = 0; i < count; i++) o.r *= safmx2;
    } else if (scale <= safmn2) {
        do {
            count++;
            f1 *= safmx2;
            g1 *= safmx2;
            scale = tabs(f1) > tabs(g1) ? tabs(f1) : tabs(g1);
        } while (scale <= safmn2);
        o.r = sqrt(f1 * f1 + g1 * g1);
        o.c = f1 / o.r;
        o.s = g1 / o.r;
        for (int i = 0; i < count; i++) o.r *= safmn2;
    } else {
        o.r = sqrt(f1 * f1 + g1 * g1);
        o.c = f1 / o.r;
        o.s = g1 / o.r;
    }
    if (tabs(f) > tabs(g) && o.c < T(0)) {
        o.c = -o.c;
        o.s = -o.s;
        o.r = -o.r;
    }
    return o;
}

// Correctly rounded square root for arguments the common path of givens() can produce
// (t = f^2 + g^2 with 2^-380 < max(|f|, |g|) < 2^485): the same reciprocal-square-root
// refinement the compiler emits for sqrt(double), without its rescaling of tiny
// arguments, which never applies in that range -- so the result is the same bits.
__device__ __forceinline__ double sqrt_common(double t)
{
    const double y = __builtin_amdgcn_rsq(t);
    double gq = t * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, gq, 0.5);
    gq = __builtin_fma(gq, r, gq);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-gq, gq, t);
    gq = __builtin_fma(d, h, gq);
    d = __builtin_fma(-gq, gq, t);
    gq = __builtin_fma(d, h, gq);
    return gq;
}
__device__ __forceinline__ float sqrt_common(float t) { return sqrtf(t); }

// fma in T (Float32: one rounding in single precision, as the reference's muladd and the
// oracle's fmaf; __builtin_fma would widen to double and round twice)
__device__ __forceinline__ double tfma(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float tfma(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}

template <typename T>
struct CommonRange;
template <>
struct CommonRange<double> {
    static constexpr double lo = 0x1p-380;  // smallest |f|, |g| of the common path
};
template <>
struct CommonRange<float> {
    static constexpr float lo = 0x1p-51f;
};

// Two IEEE quotients over the same denominator r (2^-380 <= |num| <= r < 2^486).  This is
// the compiler's double-division sequence (reciprocal, two Newton steps, quotient,
// remainder, final fma) with its operand scaling (v_div_scale / v_div_fmas /
// v_div_fixup) dropped: for operands and quotients this far from the exponent limits the
// scaling steps are identities, so the results are the same correctly rounded bits, and
// the reciprocal refinement is shared by both quotients.
__device__ __forceinline__ void div2_common(double a, double b, double r, double& qa, double& qb)
{
    double y = __builtin_amdgcn_rcp(r);
    double e = __builtin_fma(-r, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-r, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double q1 = a * y, q2 = b * y;
    const double r1 = __builtin_fma(-r, q1, a), r2 = __builtin_fma(-r, q2, b);
    qa = __builtin_fma(r1, y, q1);
    qb = __builtin_fma(r2, y, q2);
}
__device__ __forceinline__ void div2_common(float a, float b, float r, float& qa, float& qb)
{
    qa = a / r;
    qb = b / r;
}

// All lanes of the wave evaluate the chase with identical values.  The common path
// (2^-380 < |f|, |g| < 2^485) is computed unconditionally: the sign rule is applied to
// the numerators (-(f/r) == (-f)/r exactly), and the wave-uniform rare-case test only
// afterwards redirects to givens_slow, so its branch never stalls the path.
template <typename T>
__device__ __forceinline__ Rot<T> givens_common_path(T f, T g, bool& rare)
{
    const T af = tabs(f), ag = tabs(g);
    // the larger and the smaller magnitude (in T: for Float32 the double fmax would widen
    // both operands and round the result back)
    T big, small;
    if constexpr (std::is_same<T, float>::value) {
        big = __builtin_fmaxf(af, ag);
        small = __builtin_fminf(af, ag);
    } else {
        big = __builtin_fmax(af, ag);
        small = __builtin_fmin(af, ag);
    }
    rare = !(big < Consts<T>::safmx2) | !(small > CommonRange<T>::lo);
    const bool flip = (af > ag) & (f < T(0));
    const T fs = flip ? -f : f, gs = flip ? -g : g;
    const T r = sqrt_common(f * f + g * g);
    Rot<T> o;
    div2_common(fs, gs, r, o.c, o.s);
    o.r = flip ? -r : r;
    return o;
}
template <typename T>
__device__ __forceinline__ Rot<T> givens(T f, T g)
{
    bool rare;
    Rot<T> o = givens_common_path(f, g, rare);
    asm volatile("" ::"v"(o.c), "v"(o.s), "v"(o.r));
    if (__builtin_expect(__builtin_amdgcn_readfirstlane((int)rare), 0)) o = givens_slow(f, g);
    return o;
}

__device__ __forceinline__ uint32_t hi_word(double x)
{
    return (uint32_t)((uint64_t)__double_as_longlong(x) >> 32);
}
// x with its sign bit xor-ed with the top bit of m (m is 0 or 0x80000000)
__device__ __forceinline__ double xor_sign(double x, uint32_t m)
{
    return __longlong_as_double(__double_as_longlong(x) ^ (long long)((uint64_t)m << 32));
}

// Float64 specialisation of the same rotation with the bookkeeping moved off the vector
// unit.  The values are wave-uniform, so:
//  * the range test runs on the scalar unit from the high words: |x| < 2^485 exactly when
//    hi(|x|) < hi(2^485), and hi(|x|) > hi(2^-380) implies |x| > 2^-380 (the few values
//    just above 2^-380 that fail it take the complete path, which is exact for them too);
//  * |f| > |g| is one vector compare read as a wave ballot, f < 0 is the sign bit;
//  * the sign rule (c, s, r) -> -(c, s, r) is applied to r before dividing, since
//    f / (-r) == -(f / r) exactly;
//  * the reciprocal of r is refined from twice the square root's half-reciprocal estimate
//    h ~ 1/(2 sqrt t) instead of a separate v_rcp_f64: two Newton steps from that seed
//    reach RN(1/r) as the compiler's division sequence does, so the quotients are the
//    same correctly rounded values, and the reciprocal no longer waits for r's final
//    correction before it starts.
// The sign rule's negation of r and of its reciprocal: xor with a mask sm (0, or 0x80000000
// when |f| > |g| and f < 0) ...
struct XorSign {
    uint32_t sm;
    __device__ __forceinline__ double operator()(double x) const { return xor_sign(x, sm); }
};
// ... or a select by the lane mask m of f < -|g| (all lanes or none: the values are
// wave-uniform), the same condition (f < -|g| <= 0 exactly when f < 0 and |f| > |g|; false
// for any NaN, as the compare is).  One v_cndmask_b32 on the high word takes x or -x, the
// negation as its source modifier: the compare and the two selects replace a compare, an
// and, a select and two xors.
struct SelSign {
    uint64_t m;
    __device__ __forceinline__ double operator()(double x) const
    {
        const uint64_t u = (uint64_t)__double_as_longlong(x);
        uint32_t hi;
        asm("v_cndmask_b32_e64 %0, %1, -%1, %2" : "=v"(hi) : "v"((uint32_t)(u >> 32)), "s"(m));
        return __longlong_as_double((long long)(((uint64_t)hi << 32) | (uint32_t)u));
    }
};
// The common-path arithmetic of givens<double> with the sign rule sg.
template <class Sign>
__device__ __forceinline__ Rot<double> rot_common(double f, double g, Sign sg)
{
    // square root of t = f^2 + g^2 (as sqrt_common), keeping the half-reciprocal estimate
    const double t = f * f + g * g;
    const double y = __builtin_amdgcn_rsq(t);
    double gq = t * y, h = 0.5 * y;
    const double e = __builtin_fma(-h, gq, 0.5);
    gq = __builtin_fma(gq, e, gq);
    h = __builtin_fma(h, e, h);
    const double y0 = sg(h + h);  // ~ 1/(+-r), available before r
    double d = __builtin_fma(-gq, gq, t);
    gq = __builtin_fma(d, h, gq);
    d = __builtin_fma(-gq, gq, t);
    gq = __builtin_fma(d, h, gq);
    const double rr = sg(gq);  // +-RN(sqrt(t))
    // two quotients over rr with the shared, Newton-refined reciprocal: y0 is already
    // accurate to ~eps_rsq^2 (h went through a Newton step above), as the compiler's
    // division sequence has it after its first step from v_rcp_f64, so one step brings it
    // to rounding level, as the compiler's second step does
    double yy = y0;
    const double ee = __builtin_fma(-rr, yy, 1.0);
    yy = __builtin_fma(yy, ee, yy);
    const double q1 = f * yy, q2 = g * yy;
    const double r1 = __builtin_fma(-rr, q1, f), r2 = __builtin_fma(-rr, q2, g);
    Rot<double> o;
    o.c = __builtin_fma(r1, yy, q1);
    o.s = __builtin_fma(r2, yy, q2);
    o.r = rr;
    return o;
}

// Bounds of the common path: hi(|x|) in (hi(2^-380), hi(2^485)) for both operands.
constexpr uint32_t kRangeLo = 0x28300000u, kRangeHi = 0x5E400000u;

// Float64 specialisation of the same rotation with the bookkeeping moved off the vector
// unit.  The values are wave-uniform, so:
//  * the range test runs on the scalar unit from the high words: |x| < 2^485 exactly when
//    hi(|x|) < hi(2^485), and hi(|x|) > hi(2^-380) implies |x| > 2^-380 (the few values
//    just above 2^-380 that fail it take the complete path, which is exact for them too);
//  * |f| > |g| is one vector compare read as a wave ballot, f < 0 is the sign bit;
//  * the sign rule (c, s, r) -> -(c, s, r) is applied to r before dividing, since
//    f / (-r) == -(f / r) exactly;
//  * the reciprocal of r is refined from twice the square root's half-reciprocal estimate
//    h ~ 1/(2 sqrt t) instead of a separate v_rcp_f64: two Newton steps from that seed
//    reach RN(1/r) as the compiler's division sequence does, so the quotients are the
//    same correctly rounded values, and the reciprocal no longer waits for r's final
//    correction before it starts.
template <>
__device__ __forceinline__ Rot<double> givens<double>(double f, double g)
{
    const uint32_t fh = __builtin_amdgcn_readfirstlane((int)hi_word(f));
    const uint32_t gh = __builtin_amdgcn_readfirstlane((int)hi_word(g));
    // one unsigned compare of the larger offset: (h - lo - 1) wraps around for h <= lo
    const uint32_t fo = (fh & 0x7fffffffu) - (kRangeLo + 1),
                   go = (gh & 0x7fffffffu) - (kRangeLo + 1);
    const bool rare = (fo > go ? fo : go) >= kRangeHi - (kRangeLo + 1);
    // sign rule: |f| > |g| (one compare, read as a wave ballot) and f < 0 (fh's sign bit)
    const bool gt = __ballot(tabs(f) > tabs(g)) != 0;
    const uint32_t sm = (gt ? fh : 0u) & 0x80000000u;
    Rot<double> o = rot_common(f, g, XorSign{sm});
    asm volatile("" ::"v"(o.c), "v"(o.s), "v"(o.r));
    if (__builtin_expect(rare, 0)) o = givens_slow(f, g);
    return o;
}

// The same rotation with its range test deferred: no scalar round trip and no branch per
// step (a single chase wave waits on the vector-to-scalar transfer of each step's operands
// and on the branch: with the deferred test the step takes ~141 ns instead of ~182 ns).
// acc collects 2 (hi(|x|) - hi(2^-380) - 1) mod 2^32 of both operands (the offset of the
// test above, doubled so that the shift drops the sign bit); a chunk of steps stayed
// in the common range iff acc < kRareAcc at its end, and is otherwise run again with
// givens().  The sign rule is one compare into a lane mask and two selects (SelSign).
// Both forms cut the exact step from 51 to 49 instructions: the n = 2000 solve takes
// 0.251-0.255 s against 0.258-0.262 s with the five-instruction sign rule and the
// compiler's max tree, bit for bit the same.
constexpr uint32_t kRareAcc = 2u * (kRangeHi - (kRangeLo + 1));
__device__ __forceinline__ uint32_t range_word(double x)
{
    // (hi << 1) + (-2 (lo + 1)) as one shift-and-add (the compiler would otherwise take
    // the shifted word from the 64-bit value with a funnel shift and a mask)
    uint32_t r;
    asm("v_lshl_add_u32 %0, %1, 1, %2"
        : "=v"(r)
        : "v"(hi_word(x)), "s"(0u - 2u * (kRangeLo + 1)));
    return r;
}
__device__ __forceinline__ Rot<double> givens_deferred(double f, double g, uint32_t& acc)
{
    // one v_max3_u32 per step (the compiler otherwise builds a tree of max and max3)
    asm("v_max3_u32 %0, %0, %1, %2" : "+v"(acc) : "v"(range_word(f)), "v"(range_word(g)));
    return rot_common(f, g, SelSign{__builtin_amdgcn_ballot_w64(f < -tabs(g))});
}

// Opt-in fast rotation (Float64 single solves, sep_set_chase_rotations): c = f y, s = g y,
// r = t y with t = f^2 + g^2 (one fma) and y = 1/sqrt(t) from v_rsq_f64 (relative error
// below 2^-24) refined by one third-order step, y (1 + e/2 + 3 e^2 / 8) with e = 1 - t y^2
// (truncation error ~ e^3 / 3 < 2^-70).  y takes the sign of f, so c = |f| y >= 0 and r
// carries f's sign: where the reference negates (|f| > |g|, f < 0) this agrees with it,
// elsewhere it may differ by the sign of G (G and -G give the same similarity), and c >= 0
// keeps the appliers' shear forms, with coefficients s / (1 + c), well conditioned.  c, s
// and r are within about an ulp of the reference's correctly rounded quotients, but not
// always equal to them, so the chase agrees with the reference to rounding instead of bit
// for bit, in 14 vector instructions instead of 31.  The range test is deferred as in
// givens_deferred: every step of the formula is accurate for t in (2^-1000, 2^1000), and a
// chunk with a step outside it is run again with givens().
constexpr uint32_t kFastLo = 0x01700000u;  // hi(2^-1000)
constexpr uint32_t kFastHi = 0x7E700000u;  // hi(2^1000)
constexpr uint32_t kRareFast = kFastHi - (kFastLo + 1);
struct FastRotation {};
__device__ __forceinline__ Rot<double> givens_fast(double f, double g, uint32_t& acc)
{
    const double t = __builtin_fma(f, f, g * g);
    // (hi(t) - hi(2^-1000) - 1) mod 2^32 (t >= 0, so no sign bit to drop)
    acc = max(acc, hi_word(t) - (kFastLo + 1u));
    const double y0 = __builtin_amdgcn_rsq(t);
    const double e = __builtin_fma(-(t * y0), y0, 1.0);
    const double y = __builtin_copysign(__builtin_fma(y0, __builtin_fma(e, 0.375, 0.5) * e, y0), f);
    Rot<double> o;
    o.c = f * y;
    o.s = g * y;
    o.r = t * y;
    return o;
}
__device__ __forceinline__ Rot<double> step_rotation(FastRotation, double f, double g,
                                                     uint32_t& acc)
{
    return givens_fast(f, g, acc);
}

// The chase step's rotation: with the range test deferred (single-solve chunks, and the
// batched chase's fast form) or complete.
__device__ __forceinline__ Rot<double> step_rotation(std::true_type, double f, double g,
                                                     uint32_t& acc)
{
    return givens_deferred(f, g, acc);
}
// Float32: givens<float>'s arithmetic, its rare-case flag or-ed into acc (a chunk with
// acc != 0 is run again with the complete test).  The per-step scalar round trip and
// branch go, as in Float64: with the single-precision fma and max this takes the n = 2000
// Float32 solve from 0.19 to 0.14 s (73 instructions per step instead of ~96).
__device__ __forceinline__ Rot<float> step_rotation(std::true_type, float f, float g,
                                                    uint32_t& acc)
{
    bool rare;
    const Rot<float> o = givens_common_path(f, g, rare);
    acc |= (uint32_t)rare;
    return o;
}
template <typename T>
__device__ __forceinline__ Rot<T> step_rotation(std::false_type, T f, T g, uint32_t&)
{
    return givens(f, g);
}

// is_offdiagonal_small, src/bulge_chasing.jl:7-8 (1-based i; dv, ev 0-based arrays)
template <typename T>
__device__ __forceinline__ bool is_small(T e_i, T d_i, T d_ip1, T tol)
{
    return tabs(e_i) <= tol * (tabs(d_i) + tabs(d_ip1));
}

// Scaled Wilkinson shift of the trailing 2x2 (src/bulge_chasing.jl:239-265).
template <typename T>
__device__ __forceinline__ T wilkinson(T H11, T H12, T H22)
{
    const T scale = tabs(H11) + T(2) * tabs(H12) + tabs(H22);
    H11 /= scale;
    H12 /= scale;
    H22 /= scale;
    const T t = (H11 + H22) / T(2);
    const T d = (H11 - t) * (H22 - t) - H12 * H12;
    const T sq = sqrt(tabs(d));
    const T l1 = t + sq, l2 = t - sq;
    T lam = tabs(H22 - l1) < tabs(H22 - l2) ? l1 : l2;
    lam *= scale;
    return lam;
}

// Problem view.  dv (n), ev (n-1), lt (n) and the (c,s) grid: the reference's
// (n-1) x max_layers matrix of interleaved pairs (leading dimension ldg pairs).  Indices
// in the member functions are 1-based.
template <typename T>
struct Problem {
    T* dv;
    T* ev;
    int* lt;
    T* grid;
    int n;
    int max_layers;
    int64_t ldg;
};

// Pair index of rotation (i, layer) (both 1-based) in the grid.
template <typename T>
__device__ __forceinline__ int64_t grid_pair(const Problem<T>& P, int i, int layer)
{
    return (int64_t)(i - 1) + P.ldg * (layer - 1);
}

// 64-entry groups whose loads are in flight together in the deflation scan and the
// thickness maximum.  The scan runs down from `to` until it finds a negligible ev, which
// in a single large problem is usually far away (often the whole active range), and the
// maximum always reads its whole range, so each is a chain of dependent memory round trips.
// A single solve whose problem sits in global memory (kLds false) takes wide rounds (a
// sweep of an n = 8192 problem spends a few round trips on them instead of dozens).  With
// the problem in LDS a round trip is short and a wide round mostly issues loads past the
// deflation point that the scan never needed: narrow rounds are faster there.  The batched
// chase, whose problems are at most a few groups long and whose waves share their SIMDs,
// keeps narrow ones too.
template <bool kSolo, bool kLds>
struct ScanBatch {
    static constexpr int scan = kSolo && !kLds ? 8 : 2;
    static constexpr int max = kSolo && !kLds ? 32 : 8;
};
// chase steps per iteration of the chunk loop: one compare and taken branch per this many
// steps.  A single solve runs 512-step chunks and unrolls 32 steps; the batched chase,
// whose chunks are 64 steps, keeps 8 (its code size stays small beside the apply's).
template <bool kSolo>
struct StepUnroll {
    static constexpr int n = kSolo ? 32 : 8;
};
// static LDS of the batched chase (single_shift's chunk staging, 64-step chunks), per wave
constexpr unsigned kChaseLdsBytes = ((kWave + 2) * 2 + kWave * 4) * sizeof(double);

// The highest 1-based q < to with is_small(q), scanning down to `start`, or start-1.
// Each round issues the loads of kScanBatch groups of 64 indices before testing any.
template <typename T, int kScanBatch>
__device__ int deflation_scan(const Problem<T>& P, int to, int start, T tol, int lane)
{
    for (int top = to - 1; top >= start; top -= kWave * kScanBatch) {
        bool sm[kScanBatch];
#pragma unroll
        for (int u = 0; u < kScanBatch; ++u) {
            const int q = top - kWave * u - lane;
            sm[u] = false;
            if (q >= start) sm[u] = is_small(P.ev[q - 1], P.dv[q - 1], P.dv[q], tol);
        }
#pragma unroll
        for (int u = 0; u < kScanBatch; ++u) {
            const uint64_t mask = __ballot(sm[u]);
            if (mask) return top - kWave * u - (int)__builtin_ctzll(mask);
        }
    }
    return start - 1;
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
    return v;
}

// max(layer_thickness[from .. to]) (1-based, inclusive), kMaxBatch groups per round trip.
template <typename T, int kMaxBatch>
__device__ int max_thickness(const Problem<T>& P, int from, int to, int lane)
{
    int v = 0;
    for (int q0 = from; q0 <= to; q0 += kWave * kMaxBatch) {
        int w[kMaxBatch];
#pragma unroll
        for (int u = 0; u < kMaxBatch; ++u) {
            const int q = q0 + kWave * u + lane;
            w[u] = q <= to ? P.lt[q - 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < kMaxBatch; ++u) v = max(v, w[u]);
    }
    return wave_max(v);
}

// 32-bit LDS address of a __shared__ object, and back.
__device__ __forceinline__ uint32_t lds_addr(const void* p)
{
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(3))) T* lds_ptr(uint32_t a)
{
    return (__attribute__((address_space(3))) T*)(uintptr_t)a;
}

// single_shift!(H, from, to, mu, givens, layer_thickness), src/bulge_chasing.jl:11-144,
// executed by the whole wave (all lanes compute the uniform recurrence; lane j owns the
// memory traffic of steps i0 + j, i0 + 64 + j, ... of each chunk).
// kSolo: the chase wave has its SIMD to itself (single solves), so the step is bound by
// the wave's own issue: the rotation's range test is deferred to the end of the chunk, and
// chunks are 512 steps long: fewer chunk ends, whose stores, range test and grid writes
// stall the wave.  With the 32-step loop unroll and the narrow scan rounds of an
// LDS-resident problem, an n = 2000 solve takes 259-263 ms against 266-269 ms with 128-step
// chunks, 8-step unrolls and wide rounds, bit for bit the same.  Otherwise (the batched
// chase, several waves and the apply per SIMD) the range test stays on the scalar unit per
// step (80-81.5 ms against 85 ms per 4096-problem batch with the deferred test) and chunks
// are 64 steps, which keeps the chunk staging at 3 KB of LDS per wave.  kFast (Float64):
// the fast rotation and the dsteqr-style 2x2 update, with the deferred range test in both
// forms.
template <typename T, bool kSolo = true, bool kFast = false>
__device__ void single_shift(const Problem<T>& P, int from, int to, T mu, int lane)
{
    // kSolo: chunks of R 64 steps (lane j handles the memory traffic of steps j, j + 64, ...);
    // step j's (dv[i+1], ev[i+1]) at solo_in[j] (two pad entries for the read-ahead) and its
    // outputs (nrm, B11'', c, s) at solo_out[j]: 24.6 KB of LDS for double, beside the
    // problem's own 20 bytes per row when it sits in LDS (120 KB at n = 6144).  The batched
    // chase keeps 64-step chunks (its outputs are deposited in registers, one step per lane).
    constexpr int R = kSolo ? 8 : 1;
    constexpr int kChunk = R * kWave;
    __shared__ __align__(16) T solo_in[kChunk + 2][2];
    __shared__ __align__(16) T solo_out[kChunk][4];
    // (x, y) pairs of the staging rows, as one 2 x sizeof(T) LDS access each
    using V2 = T __attribute__((ext_vector_type(2)));
    using LdsV2 = __attribute__((address_space(3))) V2;
    T* __restrict__ dv = P.dv;
    T* __restrict__ ev = P.ev;
    int* __restrict__ lt = P.lt;
    T* __restrict__ grid = P.grid;
    // ---- first rotation, :17-63.  x*y + z*w is evaluated as fma(x, y, z*w) (the
    // reference's single_shift! is @fastmath; the oracle contracts identically).
    const T H11 = dv[from - 1];
    const T H21 = ev[from - 1];
    const T H22 = dv[from];
    const Rot<T> G1 = givens(H11 - mu, H21);
    const T c1 = G1.c, s1 = G1.s;
    const T H11p = tfma(c1, H11, s1 * H21);
    const T H21p = tfma(-s1, H11, c1 * H21);
    const T H12p = tfma(c1, H21, s1 * H22);
    const T H22p = tfma(-s1, H21, c1 * H22);
    const T H11pp = tfma(H11p, c1, H12p * s1);
    const T H12pp = tfma(H11p, -s1, H12p * c1);
    const T H22pp = tfma(H21p, -s1, H22p * c1);
    T bulge = T(0);
    T e_cur = T(0);  // ev[from+1] after the first step
    if (from + 2 <= to) {
        const T H23 = ev[from];
        bulge = s1 * H23;
        e_cur = c1 * H23;
    }
    {
        // every lane stores the same value to the same address (no predication)
        const int layer = __builtin_amdgcn_readfirstlane(lt[from - 1]) + 1;
        dv[from - 1] = H11pp;
        lt[from - 1] = layer;
        store_pair(grid + 2 * grid_pair(P, from, layer), c1, s1);
    }
    // carried state entering step i: p1 = ev[i-1], d_cur = dv[i], e_cur = ev[i]
    T p1 = H12pp;
    T d_cur = H22pp;

    // ---- bulge chase, :91-140, in chunks of kChunk steps
    const int i_first = from + 1, i_last = to - 1;
    if (i_first <= i_last) {
        // A chunk's inputs are staged in LDS, from which the steps read them back with one
        // broadcast read per value.  They are staged at the end of the previous chunk,
        // before its stores: the wait for the prefetch loads then covers only loads issued
        // a chunk earlier, never the stores just issued (a wave's vector memory operations
        // complete in order, so a wait for a load issued after a store also waits for the
        // store).  A wave's LDS operations complete in order, so the chunk's own reads of
        // the staging area are done when it is overwritten.
        // lane j, slot r <-> step i = i0 + 64 r + j: dv[i+1], ev[i+1] and
        // layer_thickness[i] (1-based)
        T nD[R], nE[R];
        int nL[R];
        auto stage = [&]() {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                solo_in[kWave * r + lane][0] = nD[r];
                solo_in[kWave * r + lane][1] = nE[r];
            }
        };
        auto prefetch = [&](int i0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = i0 + kWave * r + lane;
                nD[r] = T(0);
                nE[r] = T(0);
                nL[r] = 0;
                if (i <= i_last) {
                    nD[r] = dv[i];
                    if (i + 2 <= to) nE[r] = ev[i];
                    nL[r] = lt[i - 1];
                }
            }
        };
        // the inputs of a chunk's steps 0 and 1, read from the staging area at the end of
        // the previous chunk, so that the chunk's first step does not wait for them
        T l1D = T(0), l1E = T(0), l2D = T(0), l2E = T(0);
        auto look = [&]() {
            l1D = solo_in[0][0];
            l1E = solo_in[0][1];
            l2D = solo_in[1][0];
            l2E = solo_in[1][1];
        };
        prefetch(i_first);
        stage();
        look();
        int Lt[R];
#pragma unroll
        for (int r = 0; r < R; ++r) Lt[r] = nL[r];
        for (int i0u = i_first; i0u <= i_last; i0u += kChunk) {
            const int i0 = __builtin_amdgcn_readfirstlane(i0u);
            prefetch(i0 + kChunk);  // next chunk
            const int cnt = __builtin_amdgcn_readfirstlane(min(kChunk, i_last - i0 + 1));
            // outputs of step i0 + 64 r + j are kept in lane j, slot r and stored once per
            // chunk
            T oR[R], oD[R], oC[R], oS[R];
            // the chunk's entry state, for the rerun of a chunk that left the common range
            const T p1_0 = p1, d_0 = d_cur, e_0 = e_cur, b_0 = bulge;
            uint32_t acc = 0;
            auto run = [&](auto deferred) {
                // the inputs of steps j + 1 and j + 2, read from LDS ahead of use
                T a1D = l1D, a1E = l1E, a2D = l2D, a2E = l2E;
                // in2 = solo_in[j + 2] and out = solo_out[j] for step j, as pairs
                auto step = [&](const LdsV2* __restrict__ in2, LdsV2* __restrict__ out) {
                    const T B22 = a1D, B23 = a1E;
                    a1D = a2D;
                    a1E = a2E;
                    const V2 nx = *in2;
                    a2D = nx.x;
                    a2E = nx.y;
                    const Rot<T> G = step_rotation(deferred, p1, bulge, acc);
                    const T c = G.c, s = G.s;
                    const T B11 = d_cur, B21 = e_cur;
                    T B11pp, B12pp, B22pp;
                    if constexpr (std::is_same<decltype(deferred), FastRotation>::value) {
                        // G B G^T for c^2 + s^2 = 1 (to rounding) in 7 operations instead of
                        // 14, as LAPACK's dsteqr sweeps: with X = s (B22 - B11) + 2 c B21,
                        // B11'' = B11 + s X, B22'' = B22 - s X, B12'' = c X - B21 (the
                        // differences are backward stable: errors of eps |B|)
                        const T X = tfma(s, B22 - B11, c * (B21 + B21));
                        B11pp = tfma(s, X, B11);
                        B22pp = tfma(-s, X, B22);
                        B12pp = tfma(c, X, -B21);
                    } else {
                        const T B11p = tfma(c, B11, s * B21);
                        const T B21p = tfma(-s, B11, c * B21);
                        const T B12p = tfma(c, B21, s * B22);
                        const T B22p = tfma(-s, B21, c * B22);
                        B11pp = tfma(B11p, c, B12p * s);
                        B12pp = tfma(B11p, -s, B12p * c);
                        B22pp = tfma(B21p, -s, B22p * c);
                    }
                    // every lane writes the same values (the LDS pipe, not the VALU)
                    out[0] = V2{G.r, B11pp};
                    out[1] = V2{c, s};
                    p1 = B12pp;
                    d_cur = B22pp;
                    // B23 = ev[i+1]; the last step of the sweep has none: its prefetch slot
                    // holds 0, and the bulge / e_cur it then produces are never read
                    bulge = s * B23;
                    e_cur = c * B23;
                };
                // kU steps per loop iteration: one compare and taken branch per kU steps (a
                // taken branch costs a single wave more than its issue slot: per step 141 ns
                // at 1, 129 at 4, 126 at 8, 125.6 at 16).  The iteration's two LDS base
                // addresses are pinned in VGPRs, each step reaching its slots by immediate
                // offsets: the compiler otherwise copies the scalar base into a VGPR before
                // every step's read and writes, two vector moves per step in the batched
                // chase (46 -> 44 vector instructions; the 4096-problem batch stays within
                // its 54-56 ms spread).  Not in the single solve: its fast chase's reads
                // then stay behind the previous steps' writes (a wait per step instead of
                // two per 32 steps), and its exact chase, 11 instructions shorter per 32
                // steps, showed no gain.
                constexpr int kU = StepUnroll<kSolo>::n;
                int j = 0;
                for (; j + kU <= cnt; j += kU) {
                    uint32_t ib = lds_addr(&solo_in[j + 2][0]), ob = lds_addr(&solo_out[j][0]);
                    if constexpr (!kSolo) asm volatile("" : "+v"(ib), "+v"(ob));
                    const LdsV2* in2 = lds_ptr<V2>(ib);
                    LdsV2* out = lds_ptr<V2>(ob);
#pragma unroll
                    for (int u = 0; u < kU; ++u) step(in2 + u, out + 2 * u);
                }
                for (; j < cnt; ++j)
                    step(lds_ptr<V2>(lds_addr(&solo_in[j + 2][0])),
                         lds_ptr<V2>(lds_addr(&solo_out[j][0])));
            };
            // (the fast rotation is a Float64 form; Float32 defers the complete test)
            constexpr bool kF64 = std::is_same<T, double>::value;
            if constexpr ((kF64 && (kSolo || kFast)) || (!kF64 && kSolo && !kFast)) {
                if constexpr (kFast)
                    run(FastRotation{});
                else
                    run(std::true_type{});
                const uint32_t rare = !kF64 ? 1u : kFast ? kRareFast : kRareAcc;
                if (__builtin_expect(__builtin_amdgcn_readfirstlane((int)(acc >= rare)), 0)) {
                    p1 = p1_0;
                    d_cur = d_0;
                    e_cur = e_0;
                    bulge = b_0;
                    run(std::false_type{});
                }
            } else {
                run(std::false_type{});
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                oR[r] = solo_out[kWave * r + lane][0];
                oD[r] = solo_out[kWave * r + lane][1];
                oC[r] = solo_out[kWave * r + lane][2];
                oS[r] = solo_out[kWave * r + lane][3];
            }
            int layer[R];
#pragma unroll
            for (int r = 0; r < R; ++r) layer[r] = Lt[r] + 1;
            stage();
            look();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                Lt[r] = nL[r];
                // take the next chunk's thickness into its register here, before the
                // stores, so that the wait for its load cannot land behind them
                asm volatile("" : "+v"(Lt[r]));
            }
            // coalesced stores of the chunk; the next chunk was prefetched from entries
            // these stores do not touch (dv/ev beyond i0 + kChunk - 2, lt beyond
            // i0 + kChunk - 2)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (kWave * r + lane < cnt) {
                    const int i = i0 + kWave * r + lane;
                    ev[i - 2] = oR[r];   // ev[i-1] (1-based) = nrm
                    dv[i - 1] = oD[r];   // dv[i]   (1-based) = B11''
                    lt[i - 1] = layer[r];
                    store_pair(grid + 2 * grid_pair(P, i, layer[r]), oC[r], oS[r]);
                }
            }
        }
    }
    // final carried values: dv[to] and ev[to-1] (1-based); ev[to] untouched
    dv[to - 1] = d_cur;
    ev[to - 2] = p1;
}

// Block state exchanged with the host (int64 words).
enum StateWord {
    ST_TO = 0,         // in: `to` of the next block; out: last_block_ends_at
    ST_LBEA = 1,       // last_block_ends_at
    ST_ITER = 2,       // iteration count (in/out)
    ST_DID = 3,        // out: a sweep ran in this block
    ST_EARLY = 4,      // out: iter > maxiter was hit (not converged)
    ST_NROT = 5,       // out: rotations created in this block
    ST_MAXLAYER = 6,   // out: largest layer index used (apply depth)
    ST_ILO = 7,        // out: smallest rotation index used (1-based)
    ST_IHI = 8,        // out: largest rotation index used (1-based)
    ST_WORDS = 16
};

// One outer iteration of qr_algorithm! without the apply (src/bulge_chasing.jl:166-276):
// layer_thickness is reset here; the grid must hold identity rotations on entry.  kLds:
// P's dv, ev and lt point into LDS.
template <typename T, bool kSolo = true, bool kFast = false, bool kLds = false>
__device__ void chase_block(const Problem<T>& P, T tol, int64_t maxiter, int64_t* st, int lane)
{
    const int start = 1;
    int to = (int)st[ST_TO];
    int lbea = (int)st[ST_LBEA];
    int64_t iter = st[ST_ITER];
    for (int q = lane; q < P.n; q += kWave) P.lt[q] = 0;
    __threadfence_block();
    int did = 0, early = 0;
    int64_t nrot = 0;
    int ilo = P.n, ihi = 0;
    for (;;) {
        if (iter > maxiter) {
            early = 1;
            break;
        }
        int from = to;
        if (from > start)
            from = __builtin_amdgcn_readfirstlane(
                deflation_scan<T, ScanBatch<kSolo, kLds>::scan>(P, to, start, tol, lane) + 1);
        if (to <= start) break;
        if (from == to) {
            if (lane == 0) P.ev[from - 2] = T(0);
            __threadfence_block();
            if (to == lbea) lbea -= 1;
            to -= 1;
        } else {
            const int thick = __builtin_amdgcn_readfirstlane(
                max_thickness<T, ScanBatch<kSolo, kLds>::max>(P, from, to, lane));
            if (thick == P.max_layers) {
                to = from - 1;
            } else {
                const T lam = wilkinson(P.dv[to - 2], P.ev[to - 2], P.dv[to - 1]);
                did = 1;
                single_shift<T, kSolo, kFast>(P, from, to, lam, lane);
                // the next scan and sweep read what this sweep's lanes stored
                __threadfence_block();
                nrot += to - from;
                iter += 1;
                ilo = min(ilo, from);
                ihi = max(ihi, to - 1);
            }
        }
        if (to <= start) break;
    }
    int maxlayer = 0;
    if (did) maxlayer = max_thickness<T, ScanBatch<kSolo, kLds>::max>(P, 1, P.n, lane);
    if (lane == 0) {
        st[ST_TO] = lbea;
        st[ST_LBEA] = lbea;
        st[ST_ITER] = iter;
        st[ST_DID] = did;
        st[ST_EARLY] = early;
        st[ST_NROT] = nrot;
        st[ST_MAXLAYER] = maxlayer;
        st[ST_ILO] = ilo;
        st[ST_IHI] = ihi;
    }
}

#pragma clang fp contract(on)

}  // namespace chase
}  // namespace sep

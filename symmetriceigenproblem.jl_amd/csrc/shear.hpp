// shear.hpp -- the three-shear rotation step with DPP-broadcast coefficients, shared by
// the bulk apply (apply_dpp.hip) and the batched solver's apply (batched.hip).
//
// A rotation on a column pair (a, b) is applied as three in-place FMAs, a += p b;
// b += q a; a += p b.  The coefficients of one time step sit in lanes 0..15 of each
// 16-lane DPP row of one 64-bit VGPR (p, q of rotation l in lanes base + 2l, base + 2l + 1,
// the same in every row), and each FMA takes its coefficient through the DPP64
// `row_newbcast:j` source modifier, which broadcasts lane j of each row.
#pragma once

#include <hip/hip_runtime.h>

namespace sep {

// ---- the per-step FMA blocks ------------------------------------------------------
// Operand 2l / 2l+1 of each asm is window slot (s - 2l) / (s - 2l + 1) mod 2KG (the column
// pair of rotation l); the last operand is the coefficient row.  Lane numbers are the
// DPP row_newbcast sources.  The leading s_nop covers the VALU-write -> DPP-read hazard
// in case the compiler copied the coefficient row with a VALU move right before the block.
#define SEP_DPP(a, c, b, lane) \
    SEP_FMAC " " a ", " c ", " b " row_newbcast:" #lane " row_mask:0xf bank_mask:0xf\n"
// SEP_FMAC names v_fmac_f64_dpp or v_fmac_f32_dpp; it is defined around each use below.
#define SEP_ROT3(a, b, p, q) /* a += p b; b += q a; a += p b (three shears) */

// Within each phase the layers run from 7 down to 0: only layer 0's rotation reads the
// column that entered the window at this step (from LDS, or from the Q prefetch), so the
// other seven rotations of the phase cover that read's latency, which at the first step of
// an interval follows the barrier.
#define SEP_SHEAR8                                                                             \
    asm("s_nop 1\n"                                                                            \
        SEP_DPP("%14", "%16", "%15", 14) SEP_DPP("%12", "%16", "%13", 12)                        \
        SEP_DPP("%10", "%16", "%11", 10) SEP_DPP("%8", "%16", "%9", 8)                           \
        SEP_DPP("%6", "%16", "%7", 6) SEP_DPP("%4", "%16", "%5", 4) SEP_DPP("%2", "%16", "%3", 2)  \
        SEP_DPP("%0", "%16", "%1", 0)                                                          \
        SEP_DPP("%15", "%16", "%14", 15) SEP_DPP("%13", "%16", "%12", 13)                        \
        SEP_DPP("%11", "%16", "%10", 11) SEP_DPP("%9", "%16", "%8", 9)                           \
        SEP_DPP("%7", "%16", "%6", 7) SEP_DPP("%5", "%16", "%4", 5) SEP_DPP("%3", "%16", "%2", 3)  \
        SEP_DPP("%1", "%16", "%0", 1)                                                          \
        SEP_DPP("%14", "%16", "%15", 14) SEP_DPP("%12", "%16", "%13", 12)                        \
        SEP_DPP("%10", "%16", "%11", 10) SEP_DPP("%8", "%16", "%9", 8)                           \
        SEP_DPP("%6", "%16", "%7", 6) SEP_DPP("%4", "%16", "%5", 4) SEP_DPP("%2", "%16", "%3", 2)  \
        SEP_DPP("%0", "%16", "%1", 0)                                                          \
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

// Shear step: KG rotations, coefficients in lanes base .. base + 2KG - 1.  The window is a
// ring of WR >= 2KG registers; the column pair of layer l at step s is slots s - 2l and
// s - 2l + 1 (mod WR).
template <typename T, int KG, int s, int base, int WR>
__device__ __forceinline__ void step_shear(T (&w)[WR], T cf)
{
    static_assert(WR >= 2 * KG, "window ring too small");
    static_assert(KG == 8 || KG == 4, "KG is 4 or 8");
    static_assert(KG == 4 || base == 0, "KG = 8 fills a whole coefficient row");
#define SLOT(l, o) w[(((s) - 2 * (l) + (o)) % WR + WR) % WR]
#define SEP_SHEAR_ANY                                                                       \
    if constexpr (KG == 8)                                                                  \
        SEP_SHEAR8;                                                                         \
    else if constexpr (base == 0)                                                           \
        SEP_SHEAR4(0, 1, 2, 3, 4, 5, 6, 7);                                                 \
    else                                                                                    \
        SEP_SHEAR4(8, 9, 10, 11, 12, 13, 14, 15);
    if constexpr (sizeof(T) == 8) {
#define SEP_FMAC "v_fmac_f64_dpp"
        SEP_SHEAR_ANY
#undef SEP_FMAC
    } else {
#define SEP_FMAC "v_fmac_f32_dpp"
        SEP_SHEAR_ANY
#undef SEP_FMAC
    }
#undef SEP_SHEAR_ANY
#undef SLOT
}

}  // namespace sep

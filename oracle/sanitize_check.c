/*
 * sanitize_check.c -- drives every oracle entry point under AddressSanitizer +
 * UndefinedBehaviorSanitizer (and the threaded wave apply under ThreadSanitizer).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Built with sep_oracle.c by oracle/sanitize.mk
 * into oracle/_san/ and run by tests/test_oracle_sanitizers.py; a sanitizer report or a
 * failed check ends the program with a non-zero status.  With -DORC_PREFIX_FAST it checks
 * the contracted (timed CPU baseline) variant of the oracle instead.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_PREFIX_FAST
#define PFX(x) x
#else
#define PFX(x) fast_##x
#endif
#define STR_(x) #x
#define STR(x) STR_(x)

/* splitmix64, uniform in [-1, 1) */
static double urand(uint64_t *st)
{
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-52 - 1.0;
}

/* the tridiagonal families of the eigensolver checks */
enum { SPEC_RANDOM, SPEC_SPLIT, SPEC_GRADED, SPEC_WILKINSON, SPEC_ZERO };

#define T double
#define SUF _f64
#define EPS 0x1p-52
#define BIG 0x1p1000
#define TINY 0x1p-1000
#include "sanitize_check_impl.h"
#undef T
#undef SUF
#undef EPS
#undef BIG
#undef TINY

#define T float
#define SUF _f32
#define EPS 0x1p-23
#define BIG 0x1p120f
#define TINY 0x1p-120f
#include "sanitize_check_impl.h"
#undef T
#undef SUF
#undef EPS
#undef BIG
#undef TINY

int main(void)
{
    int f64 = run_all_f64();
    int f32 = run_all_f32();
    printf("sanitize_check (%s): %d Float64 and %d Float32 failures\n",
           STR(PFX(oracle)), f64, f32);
    return (f64 || f32) ? 1 : 0;
}

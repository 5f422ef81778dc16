"""The chase's deferred range test (symmetriceigenproblem.jl_amd/csrc/chase.hpp,
givens_deferred) against its per-step form (givens<double>), on the CPU.

Per step, the common path of the rotation is exact when both operands' high words hi(|x|)
lie strictly between hi(2^-380) and hi(2^485); the per-step test is one unsigned compare
of (hi(|x|) - lo - 1) mod 2^32.  The deferred form folds 2 hi(x) - 2 (lo + 1) (mod 2^32,
the shift dropping the sign bit) of every operand of a chunk into a running maximum and
compares it once with 2 (hi - lo - 1).  These tests show the two flag exactly the same
operands, over random pairs and chunks, with
zeros, subnormals, infinities and NaNs among them, for every sign and exponent and at the
bounds."""
import numpy as np

K_LO, K_HI = 0x28300000, 0x5E400000  # hi(2^-380), hi(2^485)
M32 = 0xFFFFFFFF


def per_step_rare(fh, gh):
    fo = ((fh & 0x7FFFFFFF) - (K_LO + 1)) & M32
    go = ((gh & 0x7FFFFFFF) - (K_LO + 1)) & M32
    return np.maximum(fo, go) >= K_HI - (K_LO + 1)


def range_word(h):
    return ((h << 1) + ((-2 * (K_LO + 1)) & M32)) & M32


K_RARE_ACC = 2 * (K_HI - (K_LO + 1))


def test_every_exponent():
    # every sign and exponent (the top 12 bits) with 256 mantissa patterns each, and the
    # words at the bounds
    h = np.arange(0, 1 << 32, 1 << 12, dtype=np.uint64)
    h = np.concatenate([h, np.array([K_LO - 1, K_LO, K_LO + 1, K_HI - 1, K_HI, K_HI + 1,
                                     0x80000000 | K_LO, 0x80000000 | (K_HI - 1), 0, M32],
                                    dtype=np.uint64)])
    one = per_step_rare(h, np.full_like(h, 0x3FF00000))  # the other operand 1.0
    defer = np.maximum(range_word(h), range_word(np.full_like(h, 0x3FF00000))) >= K_RARE_ACC
    assert np.array_equal(one, defer)


def test_random_pairs_and_chunks():
    r = np.random.default_rng(7)
    special = np.array([0.0, -0.0, 5e-324, 2.0 ** -1022, 2.0 ** -380, 2.0 ** -379.9, 2.0 ** 485,
                        2.0 ** 484.99, np.inf, -np.inf, np.nan, 1.0, -1.0], dtype=np.float64)
    for _ in range(20):
        x = np.concatenate([r.standard_normal(4000) * 2.0 ** r.integers(-1100, 1000, 4000),
                            r.choice(special, 96)])
        y = np.concatenate([r.standard_normal(4000) * 2.0 ** r.integers(-1100, 1000, 4000),
                            r.choice(special, 96)])
        fh = (x.view(np.uint64) >> 32).astype(np.uint64)
        gh = (y.view(np.uint64) >> 32).astype(np.uint64)
        step = per_step_rare(fh, gh)
        acc = np.maximum(range_word(fh), range_word(gh))
        assert np.array_equal(step, acc >= K_RARE_ACC)
        # a chunk (running maximum) is flagged exactly when one of its steps is
        for c in range(0, x.size, 64):
            assert (acc[c:c + 64].max() >= K_RARE_ACC) == bool(step[c:c + 64].any())

"""Seeded synthetic inputs for the five workloads the library is measured and tested on.

All generators use numpy's PCG64 (``numpy.random.default_rng(seed)``), so the same seed
gives the same arrays on every machine with this numpy; Julia's RNG streams cannot be
reproduced here, so these replace the reference's unseeded ``rand`` calls.

1. Plain solve: SymTridiagonal of order n with d_i = i (1-based), e_i ~ U[0,1), Q = I
   (the README example, README.md "Example" section).
2. Bulk apply: m x n column-major Q ~ N(0,1) and a (n-1) x k grid of rotations with
   theta ~ U[0, 2 pi), c = cos(theta), s = sin(theta) (benchmark/benchmark.jl:4-16 with
   proper rotations); flops = 6 m (n-1) k.
3. Single large solve: as 1 with n = 16384 and max_layers = 64.
4. Batched: B problems of even order drawn uniformly from [nmin, nmax], d_i = i +
   N(0, 0.1^2), e_i ~ U[0,1), packed back to back.
5. Differentiated solve: d_i = i, e_i ~ U[0.1, 1) (well separated spectrum), loss
   weights w ~ N(0,1) and G ~ N(0,1).
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 20240601


def rng(seed: int = DEFAULT_SEED) -> np.random.Generator:
    return np.random.default_rng(seed)


def tridiag_linear(n: int, seed: int = DEFAULT_SEED, dtype=np.float64, e_low: float = 0.0):
    """d_i = i (i = 1..n), e_i ~ U[e_low, 1)."""
    r = rng(seed)
    d = np.arange(1, n + 1, dtype=np.float64)
    e = e_low + (1.0 - e_low) * r.random(max(n - 1, 0))
    return d.astype(dtype), e.astype(dtype)


def givens_grid(n: int, k: int, seed: int = DEFAULT_SEED, dtype=np.float64) -> np.ndarray:
    """(k, n-1, 2) grid: theta ~ U[0, 2 pi), (cos theta, sin theta) per slot."""
    r = rng(seed)
    th = r.random((k, max(n - 1, 0))) * (2.0 * np.pi)
    g = np.empty((k, max(n - 1, 0), 2), dtype=np.float64)
    g[:, :, 0] = np.cos(th)
    g[:, :, 1] = np.sin(th)
    return g.astype(dtype)


def gaussian_colmajor(m: int, n: int, seed: int = DEFAULT_SEED, dtype=np.float64) -> np.ndarray:
    r = rng(seed)
    return np.asfortranarray(r.standard_normal((m, n)).astype(dtype))


def batched_orders(B: int, nmin: int = 32, nmax: int = 512, seed: int = DEFAULT_SEED) -> np.ndarray:
    """B even orders drawn uniformly from the even numbers in [nmin, nmax]."""
    r = rng(seed)
    lo, hi = (nmin + 1) // 2, nmax // 2
    return (2 * r.integers(lo, hi + 1, size=B)).astype(np.int64)


def batched_problems(ns, seed: int = DEFAULT_SEED, sigma: float = 0.1):
    """Packed d, e for orders ns: d_i = i + N(0, sigma^2), e_i ~ U[0,1).

    Problem p occupies d[off_d[p] : off_d[p] + n_p] and e[off_d[p] : off_d[p] + n_p - 1]
    (e has one unused slot per problem so both share off_d); Q offsets are cumulative
    n_p^2.  Returns (d, e, off_d, off_q)."""
    ns = np.asarray(ns, dtype=np.int64)
    r = rng(seed + 1)
    off_d = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    off_q = np.concatenate([[0], np.cumsum(ns * ns)[:-1]]).astype(np.int64)
    tot = int(ns.sum())
    d = np.empty(tot)
    e = np.zeros(tot)
    for p, n in enumerate(ns):
        o = off_d[p]
        d[o:o + n] = np.arange(1, n + 1) + sigma * r.standard_normal(n)
        e[o:o + n - 1] = r.random(n - 1)
    return d, e, off_d, off_q


def backward_loss_weights(n: int, seed: int = DEFAULT_SEED):
    """w ~ N(0,1) (n) and G ~ N(0,1) (n x n, column-major) for L = w.lam + <G, Q>."""
    r = rng(seed + 7)
    w = r.standard_normal(n)
    G = np.asfortranarray(r.standard_normal((n, n)))
    return w, G

#!/usr/bin/env python3
"""Benchmark of the eigensolver hot path on one MI355X.

Headline (one JSON line on stdout): Givens-apply GFLOP/s at m = n = 16384, k = 64, Float64
-- the analogue of bench(Float64, 16384, 16384, 64) in benchmark/benchmark.jl:4-16, with
proper rotations (theta ~ U[0, 2 pi)) and Q ~ N(0, 1), inputs resident in HBM.  One step =
one bulk_wave_order_2x2_rmul! call (coefficient preparation + apply kernel); flops are
6 m (n-1) k per step (benchmark/benchmark.jl:10-11).

Also reported in the same line:
* roofline of the dominant kernel (the apply kernel), timed with HIP events the library
  records on the stream the kernel runs on;
* cpu_baseline: the oracle's wave-order apply (a C restatement of src/wave.jl and
  src/bulk_givens.jl) timed on this host's cores on a row sample of the same workload;
* extra: the other apply shapes (n = 2000 Float64/Float32, n = 16384 Float32), the plain
  solve of the README example (n = 2000) beside the oracle's CPU solve, the single large
  solve (n = 16384, max_layers = 64), the differentiated-solve workload (n = 4096: forward
  solve, backward, and their ratio; and the backward batched over 1024 problems of order
  64..256), the dense-input reduction (n = 4096: sytrd against its HBM floor, orgtr), and
  the batched workload (4096 problems: solves/s against the per-problem GPU loop).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
                      [--calibrate-copy-gib G]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "symmetriceigenproblem.jl_amd")

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak
HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "apply_dpp_n16384_k64_f64_pmc.md")
# Published bench_parallel(T, 2000, 2000, 64) figures of the reference README (README.md:43-47;
# 8 threads on the author's CPU), GFLOP/s -- context for the n = 2000 extras, not the headline
PUBLISHED_N2000_8T = {"apply_n2000_f64": 267.13, "apply_n2000_f32": 530.83}


def pmc_traffic_bytes():
    """HBM bytes per launch of the apply kernel, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain-copy calibration); None when absent."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def load_package():
    spec = importlib.util.spec_from_file_location(
        "sep_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sep_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n: int, k: int, rows: int, threads: int, reps: int):
    """Oracle wave-order apply (FMA-contracted build) on `rows` rows of the n x n workload."""
    import numpy as np
    sys.path.insert(0, ROOT)
    import oracle
    sep = sys.modules["sep_amd"]
    A = sep.inputs.gaussian_colmajor(rows, n, seed=3)
    g = sep.inputs.givens_grid(n, k, seed=4)
    oracle.bulk_wave(A, g, threads=threads, fast=True)  # warm-up (page in, thread start)
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        oracle.bulk_wave(A, g, threads=threads, fast=True)
        best = min(best, time.perf_counter() - t)
    flops = 6.0 * rows * (n - 1) * k
    return flops / best / 1e9, best


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_apply(sep, h, torch, n, k, dtype, steps=5):
    gen = torch.Generator(device="cuda").manual_seed(5)
    A = torch.randn((n, n), generator=gen, dtype=dtype, device="cuda").t()
    grid = torch.from_numpy(sep.inputs.givens_grid(n, k, seed=6)).to(dtype).cuda()
    sep.bulk_wave_order_2x2_rmul_(A, grid, handle=h)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        sep.bulk_wave_order_2x2_rmul_(A, grid, handle=h)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / steps
    return 6.0 * n * (n - 1) * k / (ms * 1e-3) / 1e9, ms


EXTRA_SECTIONS = ("apply", "solve", "large", "backward", "dense", "batched")


def extras(sep, h, torch, np, cpu=True, large=True, sections=EXTRA_SECTIONS):
    """Secondary figures: other apply shapes, config-1 solve, the differentiated solve, the
    dense-input reduction (sytrd / orgtr), the batched solve (each a section that
    `sections` selects)."""
    sys.path.insert(0, ROOT)
    ex = {}
    for n, dt, name in () if "apply" not in sections else ((2000, torch.float64, "apply_n2000_f64"),
                        (2000, torch.float32, "apply_n2000_f32"),
                        (16384, torch.float32, "apply_n16384_f32")):
        gfs, ms = time_apply(sep, h, torch, n, 64, dt)
        peak = FP64_PEAK_TFLOPS if dt == torch.float64 else FP32_PEAK_TFLOPS
        ex[name + "_gflops"] = round(gfs, 1)
        ex[name + "_frac_of_vector_peak"] = round(gfs / 1e3 / peak, 4)
        if name in PUBLISHED_N2000_8T:
            # the reference README's bench_parallel figure (8 threads, its own CPU)
            ex[name + "_over_published_8thread_cpu"] = round(gfs / PUBLISHED_N2000_8T[name], 1)
        log(f"[bench] {name}: {gfs:.1f} GFLOP/s ({ms:.3f} ms/step)")

    if "solve" in sections:
        solve_section(sep, h, torch, np, ex, cpu)
    if large and "large" in sections:
        large_section(sep, h, torch, ex)
    if "backward" in sections:
        backward_section(sep, h, torch, np, ex)
    if "dense" in sections:
        dense_section(sep, h, torch, np, ex)
    if "batched" in sections:
        batched_section(sep, h, torch, np, ex)
    return ex


def solve_section(sep, h, torch, np, ex, cpu):
    # config 1: README example solve, n = 2000, max_layers = 128
    d, e = sep.inputs.tridiag_linear(2000)
    H = sep.SymTridiagonal(torch.from_numpy(d).cuda(), torch.from_numpy(e).cuda())
    Q = torch.eye(2000, dtype=torch.float64, device="cuda").t()
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, conv, it, nrot = sep.qr_algorithm_(H, Q, handle=h, return_rotations=True)
    torch.cuda.synchronize()
    ex["solve_n2000_s"] = round(time.perf_counter() - t, 4)
    ex["solve_n2000_iter"] = it
    ex["solve_n2000_rotations"] = nrot
    ex["solve_n2000_apply_gflops"] = round(6.0 * 2000 * nrot / ex["solve_n2000_s"] / 1e9, 1)
    # the reference README's qr_algorithm! at n = 2000 (its own CPU, unseeded input): 0.258 s
    ex["solve_n2000_over_published_cpu"] = round(0.258 / ex["solve_n2000_s"], 2)
    log(f"[bench] solve n=2000: {ex['solve_n2000_s']} s, iter={it}, rotations={nrot}")
    # the same solve with the opt-in fast chase rotation (results equal to rounding)
    H = sep.SymTridiagonal(torch.from_numpy(d).cuda(), torch.from_numpy(e).cuda())
    Q = torch.eye(2000, dtype=torch.float64, device="cuda").t()
    h.set_chase_rotations("fast")
    try:
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, conv, it, nrot = sep.qr_algorithm_(H, Q, handle=h, return_rotations=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
    finally:
        h.set_chase_rotations("exact")
    ex["solve_n2000_fast_rotations_s"] = round(el, 4)
    ex["solve_n2000_fast_rotations_iter"] = it
    ex["solve_n2000_fast_rotations_over_published_cpu"] = round(0.258 / el, 2)
    log(f"[bench] solve n=2000, fast rotations: {el:.4f} s, iter={it}, rotations={nrot}")
    if cpu:
        import oracle
        dv, ev = d.copy(), e.copy()
        Qh = np.asfortranarray(np.eye(2000))
        t = time.perf_counter()
        oracle.qr_algorithm(dv, ev, Qh, 128, fast=True)
        ex["solve_n2000_cpu_port_s"] = round(time.perf_counter() - t, 3)
        ex["solve_n2000_cpu_port_note"] = ("oracle qr_algorithm (C restatement, FMA build), "
                                           "1 thread, chase + layer-major apply")
        log(f"[bench] solve n=2000 on the CPU port: {ex['solve_n2000_cpu_port_s']} s")



def large_section(sep, h, torch, ex):
    # config 3: single large solve, n = 16384, max_layers = 64 (chase of block b + 1 beside
    # the apply of block b)
    if True:
        n = 16384
        d, e = sep.inputs.tridiag_linear(n, seed=7)
        H = sep.SymTridiagonal(torch.from_numpy(d).cuda(), torch.from_numpy(e).cuda())
        Q = torch.eye(n, dtype=torch.float64, device="cuda").t()
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, conv, it, nrot = sep.qr_algorithm_(H, Q, 64, handle=h, return_rotations=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        ex["solve_n16384_s"] = round(el, 3)
        ex["solve_n16384_iter"] = it
        ex["solve_n16384_rotations"] = nrot
        ex["solve_n16384_chase_ns_per_rotation"] = round(el / nrot * 1e9, 1)
        ex["solve_n16384_apply_gflops"] = round(6.0 * n * nrot / el / 1e9, 1)
        log(f"[bench] solve n=16384: {el:.2f} s, iter={it}, rotations={nrot}")
        del H, Q



def backward_section(sep, h, torch, np, ex):
    # config 5: differentiated solve, n = 4096, d_i = i, e ~ U[0.1, 1)
    n = 4096
    d, e = sep.inputs.tridiag_linear(n, seed=1, e_low=0.1)
    w, G = sep.inputs.backward_loss_weights(n)
    H = sep.SymTridiagonal(torch.from_numpy(d).cuda(), torch.from_numpy(e).cuda())
    torch.cuda.synchronize()
    t = time.perf_counter()
    lam, Q = sep.eigen_decomp_(H, handle=h)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t
    wd = torch.from_numpy(w).cuda()
    Gd = sep.to_device_colmajor(G)
    sep.eig_backward(lam, Q, wd, Gd, handle=h)
    torch.cuda.synchronize()
    reps = 5
    t = time.perf_counter()
    for _ in range(reps):
        sep.eig_backward(lam, Q, wd, Gd, handle=h)
    torch.cuda.synchronize()
    bwd = (time.perf_counter() - t) / reps
    ex["ad_n4096_forward_s"] = round(fwd, 4)
    ex["ad_n4096_backward_s"] = round(bwd, 5)
    ex["ad_n4096_bwd_over_fwd"] = round(bwd / fwd, 5)
    ex["ad_n4096_backward_tflops"] = round(4.0 * n ** 3 / bwd / 1e12, 2)
    log(f"[bench] AD n=4096: forward {fwd:.3f} s, backward {bwd * 1e3:.2f} ms "
        f"({ex['ad_n4096_backward_tflops']} TFLOP/s on 4 n^3), ratio {bwd / fwd:.5f}")
    del lam, Q, Gd

    # config 5, batched: the same backward over 1024 problems of order 64..256 (forward by
    # the batched solver, eigenvalues as it leaves them; the cotangents are N(0, 1))
    Bb = 1024
    rb = np.random.default_rng(11)
    nsb = rb.integers(64, 257, size=Bb).astype(np.int64)
    nsb += nsb % 2  # even orders (the solver's precondition)
    db, eb, offb, offqb = sep.inputs.batched_problems(nsb, seed=13)
    totq = int(offqb[-1] + nsb[-1] ** 2)
    fwd_b = float("inf")
    for _ in range(2):  # the first call also captures the solver's graph
        dvb = torch.from_numpy(db.copy()).cuda()
        evb = torch.from_numpy(eb.copy()).cuda()
        Qb = torch.zeros(totq, dtype=torch.float64, device="cuda")
        for p, nn in enumerate(nsb):
            Qb[offqb[p]:offqb[p] + nn * nn].view(nn, nn).fill_diagonal_(1.0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        sep.qr_batched_(nsb, offb, offqb, dvb, evb, Qb, handle=h)
        torch.cuda.synchronize()
        fwd_b = min(fwd_b, time.perf_counter() - t)
    lbar = torch.from_numpy(rb.standard_normal(dvb.shape[0])).cuda()
    qbar = torch.from_numpy(rb.standard_normal(totq)).cuda()
    sep.eig_backward_batched(nsb, offb, offqb, dvb, Qb, lbar, qbar, handle=h)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        sep.eig_backward_batched(nsb, offb, offqb, dvb, Qb, lbar, qbar, handle=h)
    torch.cuda.synchronize()
    bwd_b = (time.perf_counter() - t) / reps
    fl = 4.0 * float((nsb.astype(np.float64) ** 3).sum())
    ex["ad_batched_1024_forward_s"] = round(fwd_b, 4)
    ex["ad_batched_1024_backward_s"] = round(bwd_b, 5)
    ex["ad_batched_1024_bwd_over_fwd"] = round(bwd_b / fwd_b, 4)
    ex["ad_batched_1024_backward_tflops"] = round(fl / bwd_b / 1e12, 2)
    log(f"[bench] AD batched 1024 x n in [64, 256]: forward {fwd_b * 1e3:.1f} ms (default "
        f"configuration), backward {bwd_b * 1e3:.2f} ms ({ex['ad_batched_1024_backward_tflops']} "
        f"TFLOP/s on 4 sum n^3), ratio {bwd_b / fwd_b:.3f}")
    del dvb, evb, Qb, lbar, qbar


def dense_section(sep, h, torch, np, ex):
    # dense symmetric input (no reference counterpart): reduction to tridiagonal and the
    # explicit orthogonal factor at n = 4096, A = (X + X^T) / 2 with X ~ N(0, 1)
    n = 4096
    r = np.random.default_rng(3)
    X = r.standard_normal((n, n))
    Ah = np.asfortranarray((X + X.T) / 2)
    t_s, t_o = float("inf"), float("inf")
    for _ in range(2):
        Ad = sep.to_device_colmajor(Ah)
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, _, tau = sep.sytrd_(Ad, handle=h)
        torch.cuda.synchronize()
        t_s = min(t_s, time.perf_counter() - t)
        t = time.perf_counter()
        sep.orgtr(Ad, tau, handle=h)
        torch.cuda.synchronize()
        t_o = min(t_o, time.perf_counter() - t)
    # one read and one write of the trailing lower triangle per column: 8 n^3 / 3 bytes
    floor_s = 8.0 * n ** 3 / 3 / 8e12
    ex["sytrd_n4096_s"] = round(t_s, 4)
    ex["sytrd_n4096_hbm_floor_frac"] = round(floor_s / t_s, 3)
    ex["orgtr_n4096_s"] = round(t_o, 4)
    ex["orgtr_n4096_tflops"] = round(4.0 * n ** 3 / 3 / t_o / 1e12, 2)
    log(f"[bench] dense n=4096: sytrd {t_s * 1e3:.1f} ms ({100 * floor_s / t_s:.0f}% of its "
        f"HBM floor), orgtr {t_o * 1e3:.1f} ms ({ex['orgtr_n4096_tflops']} TFLOP/s on 4n^3/3)")


def batched_section(sep, h, torch, np, ex):
    # config 4: 4096 uneven problems, even n in [32, 512], d_i = i + N(0, 0.1), e ~ U[0,1)
    B = 4096
    ns = sep.inputs.batched_orders(B)
    d, e, off_d, off_q = sep.inputs.batched_problems(ns, seed=5)
    tot_q = int(off_q[-1] + ns[-1] ** 2)

    def fresh():
        dvb = torch.from_numpy(d.copy()).cuda()
        evb = torch.from_numpy(e.copy()).cuda()
        Qb = torch.zeros(tot_q, dtype=torch.float64, device="cuda")
        for p, nn in enumerate(ns):
            Qb[off_q[p]:off_q[p] + nn * nn].view(nn, nn).fill_diagonal_(1.0)
        torch.cuda.synchronize()
        return dvb, evb, Qb

    dvb, evb, Qb = fresh()
    t = time.perf_counter()
    choice = sep.tune_batched(ns, off_d, off_q, dvb, evb, handle=h)
    tune_s = time.perf_counter() - t
    best = float("inf")
    for _ in range(2):
        dvb, evb, Qb = fresh()
        t = time.perf_counter()
        conv, iters = sep.qr_batched_(ns, off_d, off_q, dvb, evb, Qb, handle=h)
        best = min(best, time.perf_counter() - t)
    # the same batch with the opt-in fast chase rotation (results equal to rounding)
    best_fast = float("inf")
    h.set_chase_rotations("fast")
    try:
        for _ in range(2):
            dvb, evb, Qb = fresh()
            t = time.perf_counter()
            conv_f, _ = sep.qr_batched_(ns, off_d, off_q, dvb, evb, Qb, handle=h)
            best_fast = min(best_fast, time.perf_counter() - t)
    finally:
        h.set_chase_rotations("exact")
    ex["batched_fast_rotations_s"] = round(best_fast, 4)
    ex["batched_fast_rotations_all_converged"] = bool(conv_f.all())
    log(f"[bench] batched B={B}, fast rotations: {best_fast * 1e3:.1f} ms")
    # the plain per-problem GPU loop (same library, one qr_algorithm! per problem) on every
    # 64th problem, extrapolated to the batch by problem count; inputs are staged on the
    # device before the clock starts, as for the batched solve
    sample = list(range(0, B, 64))
    staged = []
    for p in sample:
        nn, o = int(ns[p]), int(off_d[p])
        staged.append((sep.SymTridiagonal(torch.from_numpy(d[o:o + nn].copy()).cuda(),
                                          torch.from_numpy(e[o:o + nn - 1].copy()).cuda()),
                       torch.eye(nn, dtype=torch.float64, device="cuda").t()))
    torch.cuda.synchronize()
    t = time.perf_counter()
    for Hs, Qs in staged:
        sep.qr_algorithm_(Hs, Qs, 32, handle=h)
    torch.cuda.synchronize()
    loop_rate = len(sample) / (time.perf_counter() - t)
    del staged
    ex["batched_solves_per_s"] = round(B / best, 1)
    ex["batched_s"] = round(best, 4)
    ex["batched_all_converged"] = bool(conv.all())
    ex["batched_tuner_choice"] = int(choice)
    ex["batched_tune_s"] = round(tune_s, 3)
    ex["per_problem_loop_solves_per_s"] = round(loop_rate, 2)
    ex["batched_over_loop"] = round(B / best / loop_rate, 1)
    log(f"[bench] batched B={B}: {best * 1e3:.1f} ms = {B / best:.0f} solves/s; per-problem "
        f"loop {loop_rate:.1f} solves/s; ratio {B / best / loop_rate:.0f}x")



def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--dtype", choices=("f64", "f32"), default="f64",
                    help="element type of the apply (the headline is f64; f32 is for "
                         "profiling the Float32 kernel)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the n = 16384 solve in the extras (about 40 s)")
    ap.add_argument("--extras", default=",".join(EXTRA_SECTIONS),
                    help="comma-separated extra sections to run (of "
                         + ", ".join(EXTRA_SECTIONS) + "); profiling runs pick one")
    ap.add_argument("--cpu-rows", type=int, default=16384)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="after the timed steps, run one plain device copy of this many GiB "
                         "(the reference kernel for FETCH_SIZE / WRITE_SIZE calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the apply shards by rows of Q with no communication "
              "but the (c,s) broadcast, and multi-GPU runs are not part of this build "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    sep = load_package()
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    h = sep.Handle(0)
    stream = torch.cuda.current_stream()
    h.set_stream(stream.cuda_stream)

    n = m = args.n
    k = args.k
    f64 = args.dtype == "f64"
    tdt = torch.float64 if f64 else torch.float32
    log(f"[bench] workload: bulk Givens apply m=n={n} k={k} {args.dtype}")
    gen = torch.Generator(device="cuda").manual_seed(1)
    A = torch.randn((n, m), generator=gen, dtype=tdt, device="cuda").t()  # column-major
    grid = torch.from_numpy(sep.inputs.givens_grid(n, k, seed=2)).to(tdt).cuda()
    for _ in range(args.warmup):
        sep.bulk_wave_order_2x2_rmul_(A, grid, handle=h)
    torch.cuda.synchronize()
    h.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sep.bulk_wave_order_2x2_rmul_(A, grid, handle=h)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    apply_ms, prep_ms, launches = h.get_timing()
    h.set_timing(False)
    ms_step = wall * 1e3 / args.steps
    flops = 6.0 * m * (n - 1) * k
    value = flops / (ms_step * 1e-3) / 1e9
    kern_ms = apply_ms / max(launches, 1)
    achieved = flops * (args.steps / max(launches, 1)) / (kern_ms * 1e-3) / 1e12
    log(f"[bench] {ms_step:.3f} ms/step, {value:.1f} GFLOP/s; apply kernel {kern_ms:.3f} ms "
        f"({achieved:.2f} TFLOP/s), prep {prep_ms / args.steps:.3f} ms/step")
    # finite check on the result (rotations preserve row norms)
    assert torch.isfinite(A).all().item()

    peak = FP64_PEAK_TFLOPS if f64 else FP32_PEAK_TFLOPS
    esz = 8 if f64 else 4
    out = {
        "metric": f"Givens-apply GFLOP/s n={n} " + ("FP64" if f64 else "FP32"),
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (Q ~ N(0,1) seeded on device; theta ~ U[0,2pi) seeded, c=cos, s=sin)",
        "config": {"workload": f"bulk Givens apply (bench(Float64, {m}, {n}, {k}) analogue)",
                   "m": m, "n": n, "k": k, "flops_per_step": flops},
        "roofline": {
            "bound": "valu",
            "pipe": ("FP64 vector ALU (v_fmac_f64 with DPP row_newbcast coefficients), not "
                     "MFMA; MI355X's FP64 vector peak is 78.6 TFLOP/s") if f64 else
                    "FP32 VALU (v_fmac_f32 with DPP row_newbcast coefficients)",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": pmc_traffic_bytes() if (f64 and n == 16384 and k == 64) else None,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE x 1)",
            "algorithmic_bytes_per_launch": 2.0 * esz * m * n + 2.0 * esz * (n - 1) * k,
            "kernel": "k_apply_dpp<8,16>",
            "kernel_ms": round(kern_ms, 4),
            "algorithmic_flops_per_launch": flops * args.steps / max(launches, 1),
        },
    }

    cpu_extra = {}
    if not args.no_cpu and f64:
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline: oracle wave-order apply, {args.cpu_rows} rows x n={n}, "
            f"k={k}, {threads} threads, best of {args.cpu_reps}")
        gfs, best = cpu_baseline(n, k, args.cpu_rows, threads, args.cpu_reps)
        out["cpu_baseline"] = {
            "value": round(gfs, 2), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_rows} of {m} rows (rows are independent), n={n}, k={k}, "
                      f"min of {args.cpu_reps} runs of {best:.3f} s",
        }
        log(f"[bench] cpu {gfs:.1f} GFLOP/s")
        # the single-thread figure (the analogue of bench beside bench_parallel) on a
        # 1024-row sample, and the host it was measured on
        gfs1, best1 = cpu_baseline(n, k, 1024, 1, args.cpu_reps)
        cpu_extra = {
            "cpu_apply_1thread_gflops": round(gfs1, 2),
            "cpu_apply_1thread_sample": f"1024 rows x n={n}, k={k}, min of {args.cpu_reps} runs of "
                                        f"{best1:.3f} s",
            "cpu_host_nproc": os.cpu_count(),
            "cpu_host_model": cpu_model(),
        }
        log(f"[bench] cpu 1 thread: {gfs1:.1f} GFLOP/s ({cpu_extra['cpu_host_model']}, "
            f"{cpu_extra['cpu_host_nproc']} logical CPUs)")

    if args.calibrate_copy_gib > 0:
        nel = int(args.calibrate_copy_gib * (1 << 30)) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibrate_copy_gib} GiB done")
        del src, dst

    if not args.no_extra:
        out["extra"] = extras(sep, h, torch, np, cpu=not args.no_cpu, large=not args.no_large,
                              sections=tuple(args.extras.split(",")))
    if cpu_extra:
        out.setdefault("extra", {}).update(cpu_extra)

    print(json.dumps(out), flush=True)
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

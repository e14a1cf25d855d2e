#!/usr/bin/env python3
"""Benchmark of the on-the-fly FCI / VQE operator path on one MI355X.

Workload (the headline configuration that fits one GPU): 16 spatial orbitals, 8 alpha + 8 beta
electrons, 12,870^2 = 165,636,900 determinants, seeded synthetic integrals (n = 16 recipe of
binsim_amd/integrals.py).  A step is one FP64 sigma build H x (hvec_cuda) on a resident,
seeded vector; ``value`` is its algorithmic HBM rate  B_sigma / t  with
B_sigma = 8 * dim * (2 + s_a + d_a) bytes (s_a = 64 singles, d_a = 784 doubles per string).

Reported beside it (one run each, outside the timed steps):
  * FCI, lowest 4 roots: plain FP64 Davidson (one FP64 sigma per vector) and mixed-precision
    block Davidson (FP32 block sigma, FP64 Rayleigh-Ritz, FP64 refinement), wall seconds and
    per-root |dE| -- at 12 orbitals / 6+6 electrons to residual 1e-8 by default, and at the
    16-orbital workload to residual 1e-6 with --full (each solve takes minutes there);
  * FCI at 12 orbitals / 6+6 electrons, one root from HF to 1e-8 (the CPU-runnable case);
  * VQE energy + adjoint gradient (fused) at 12 orbitals (1,818 parameters) and at 16 orbitals
    (5,792 parameters): gradients/s and max deviation from central differences (h = 1e-6);
  * cpu_baseline: the reference's own CPU sigma (oracle/_ref/libotf_native, OpenMP) on a
    bounded sample, all host cores.

    python bench.py [--steps K] [--warmup W] [--quick]
    python bench.py --davidson-only 600     # 16-orbital ground state, plain vs mixed
    python bench.py --sigma-shape 18,7      # FP64 / FP32 sigma at another shape
    python bench.py --vqe-only 16           # VQE energy + gradient only
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# HBM bytes of one FP64 sigma build at 16 orbitals 8a8b, from two rocprofv3 --pmc runs of
# `bench.py --pmc-unit` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, summed over the
# dispatches of one sigma), each scaled by what the same counter reads for a plain copy kernel
# of 1.325 GB in the same run (profiles/sigma_fp64_16o_pmc.md)
SIGMA_TRAFFIC_16O = 266.4e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sigma_bytes(n, k, w=8):
    from math import comb
    N = comb(n, k)
    s = k * (n - k)
    d = comb(k, 2) * comb(n - k, 2)
    return w * N * N * (2 + s + d)


def seeded(torch, dim, seed):
    import numpy as np
    x = np.random.default_rng(seed).uniform(-1.0, 1.0, dim)
    x /= np.linalg.norm(x)
    return torch.from_numpy(x).to("cuda")


def build(n, na, nb, pool=False):
    from binsim_amd import otf
    from binsim_amd.ansatz import ucc_spin_orbitals
    from binsim_amd.integrals import for_norb, one_body_layout, reference_layout
    h, eri = for_norb(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    H = otf.HipOTF.from_integrals(n, one_body_layout(h), reference_layout(eri))
    P = otf.HipOTF.from_spin_orbitals(n, ucc_spin_orbitals(n, na, nb)) if pool else None
    return B, H, P, otf.HipOTF_Functions(B, ham=H, pool=P)


def timed_sigma(torch, F, x, y, steps, warmup):
    for _ in range(warmup):
        F.hvec(x, y)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall = time.perf_counter()
    t0.record()
    for _ in range(steps):
        F.hvec(x, y)
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    return t0.elapsed_time(t1) / steps, wall


def fci_block(torch, B, H, nroots, maxspace, tol, maxiter):
    from binsim_amd import davidson as D
    out = {"tol": tol, "maxspace": maxspace}
    for name, mixed in (("plain_fp64", False), ("mixed_block", True)):
        e, _, info = D.solve(B, H, nroots=nroots, mixed=mixed, tol=tol, maxiter=maxiter,
                             maxspace=maxspace, verbose=True)
        out[name] = dict(seconds=round(info["seconds"], 3), iterations=int(info["iterations"]),
                         sigma_fp64=int(info["sigma64"]), sigma_fp32=int(info["sigma32"]),
                         converged=bool(info["converged"] == 1),
                         max_residual=info["max_residual"], energies=[float(v) for v in e])
        log(f"fci {name}: {info['seconds']:.2f} s, {int(info['iterations'])} it, e = {e}")
    import numpy as np
    de = np.abs(np.array(out["plain_fp64"]["energies"]) - np.array(out["mixed_block"]["energies"]))
    out["abs_dE_per_root"] = [float(v) for v in de]
    out["speedup_mixed_vs_plain"] = round(out["plain_fp64"]["seconds"] / out["mixed_block"]["seconds"], 3)
    return out


def sigma_shape(torch, n, k, steps, warmup):
    """FP64 and single-vector FP32 sigma for n orbitals, k+k electrons, on a seeded vector,
    wall clock around ``steps`` calls after ``warmup`` (stored launch choices where present)."""
    from binsim_amd import autotune as AT
    B, H, _, F = build(n, k, k)
    entry = AT.lookup(B)
    if entry is not None:
        AT.apply(entry)
    x = seeded(torch, B.dim, 1000 + n)
    y, y32 = torch.empty_like(x), torch.empty_like(x)
    out = {"sigma_shape": f"{n} orbitals {k}a{k}b", "dim": B.dim, "steps": steps}
    for name, fn, w in (("fp64", lambda: F.hvec(x, y), 8), ("fp32", lambda: F.hvec_block([x], [y32], 1), 4)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        out[name] = dict(seconds=round(dt, 4), algorithmic_GBps=round(sigma_bytes(n, k) * w / 8 / dt / 1e9, 1))
        log(f"sigma {name} n={n} k={k}: {dt * 1e3:.1f} ms")
    out["fp32_rel_dev"] = float((y32 - y).norm() / y.norm())
    return out


def fci16_ground(torch, maxiter):
    """16-orbital 8a8b ground state to 1e-6 from the lowest diagonal: plain FP64 (the reference's
    restart-to-3 Davidson) and the mixed solver (block 1, maxspace 32), launch choices from the
    stored tuning entry."""
    from binsim_amd import autotune as AT
    from binsim_amd import davidson as D
    B, H, _, _ = build(16, 8, 8)
    entry = AT.lookup(B)
    if entry is not None:
        AT.apply(entry)
    out = {"fci16": "ground state", "tol": 1e-6, "maxiter": maxiter}
    for name, mixed, block, maxspace in (("plain_fp64", False, 0, 0), ("mixed_block1", True, 1, 32)):
        e, _, info = D.solve(B, H, nroots=1, mixed=mixed, tol=1e-6, block=block, maxspace=maxspace,
                             maxiter=maxiter, verbose=True)
        out[name] = dict(seconds=round(info["seconds"], 3), iterations=int(info["iterations"]),
                         sigma_fp64=int(info["sigma64"]), sigma_fp32=int(info["sigma32"]),
                         converged=bool(info["converged"] == 1), energy=float(e[0]))
        log(f"fci16 {name}: {info['seconds']:.2f} s, {int(info['iterations'])} it, e = {e[0]}")
    out["abs_dE"] = abs(out["plain_fp64"]["energy"] - out["mixed_block1"]["energy"])
    out["speedup_mixed_vs_plain"] = round(out["plain_fp64"]["seconds"] / out["mixed_block1"]["seconds"], 3)
    return out


def vqe(torch, n, na, nb, nfd, seed_off=2000, amp=0.05, stored_tuning=False):
    import numpy as np
    B, H, P, F = build(n, na, nb, pool=True)
    if stored_tuning:
        from binsim_amd import autotune as AT
        entry = AT.lookup(B)
        if entry is not None:
            AT.apply(entry)
    from binsim_amd.ansatz import ucc_spin_orbitals
    from binsim_amd.integrals import SEED_BASE
    npar = len(ucc_spin_orbitals(n, na, nb))
    th = np.random.default_rng(SEED_BASE + seed_off).uniform(-amp, amp, npar)
    idx = np.arange(npar)
    psi = torch.zeros(B.dim, dtype=torch.float64, device="cuda")
    lam = torch.empty_like(psi)

    def run(x):
        psi.zero_()
        psi[0] = 1.0
        return F.vqe_energy_grad(idx, x, psi, lam)

    run(th)
    torch.cuda.synchronize()
    reps = 3 if n <= 12 else 1
    t = time.perf_counter()
    for _ in range(reps):
        e, g, d2, _ = run(th)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    rng = np.random.default_rng(SEED_BASE + 3000)
    sample = np.sort(rng.choice(npar, min(nfd, npar), replace=False))
    h = 1e-6
    dev = 0.0
    for k in sample:
        xp, xm = th.copy(), th.copy()
        xp[k] += h
        xm[k] -= h
        fd = (run(xp)[0] - run(xm)[0]) / (2 * h)
        dev = max(dev, abs(fd - g[k]))
    log(f"vqe n={n}: {dt * 1e3:.1f} ms per energy+grad, fd dev {dev:.2e} over {len(sample)}")
    return dict(parameters=npar, seconds_per_eval=round(dt, 5), grads_per_s=round(1.0 / dt, 3),
                energy=e, dH2=d2, fd_max_dev=dev, fd_params=int(len(sample)))


def cpu_baseline(budget_s=30.0, dav_iters=25):
    """The reference's OpenMP CPU path (oracle/_ref) on the 12-orbital system: the sigma,
    repeated (the reported value), then a bounded Davidson sample and one VQE evaluation."""
    import numpy as np
    from oracle import glue
    from oracle.ref import Reference, available
    from binsim_amd.integrals import for_norb, one_body_layout, reference_layout
    if not available():
        return None
    R = Reference()
    n, k = 12, 6
    h, eri = for_norb(n)
    terms = R.hamiltonian_terms(one_body_layout(h), reference_layout(eri), n)
    flat = glue.flatten_network(glue.compress_by_svd(terms))
    bp, dim = R.basis(n, k, k)
    net = R.network(flat, n)
    x = np.random.default_rng(1012).uniform(-1, 1, dim)
    R.hvec(bp, net, x)
    reps, t = 0, time.perf_counter()
    while True:
        R.hvec(bp, net, x)
        reps += 1
        if time.perf_counter() - t > min(budget_s, 10.0) or reps >= 50:
            break
    dt = (time.perf_counter() - t) / reps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = dict(value=round(sigma_bytes(n, k) / dt / 1e9, 3), unit="GB/s", cores=cores,
               kind="reference",
               sample=f"reference CPU sigma (oracle/_ref libotf_native, hvec_gather_contract_otf_f64) "
                      f"on 12 orbitals / 6+6 electrons ({dim} dets), {reps} calls, "
                      f"{dt * 1e3:.1f} ms per sigma, algorithmic bytes of that shape",
               seconds_per_sigma=dt)
    # the same system's Davidson from HF (jl/davidson.jl over the reference libdiag and sigma;
    # 80 iterations to 1e-8, here the first dav_iters of them) and one VQE energy + gradient
    # (jl/vqe.jl over the reference expm / sigma / backgrad, all 1,818 parameters)
    diag = R.diags(bp, net, dim)
    hf = np.zeros(dim)
    hf[0] = 1.0
    t = time.perf_counter()
    e0, its, conv = R.davidson(lambda v: R.hvec(bp, net, v), hf, diag, 1e-8, maxiter=dav_iters)
    t_dav = time.perf_counter() - t
    from binsim_amd.ansatz import ucc_spin_orbitals
    from binsim_amd.integrals import SEED_BASE
    sos = ucc_spin_orbitals(n, k, k)
    pool = R.network(glue.flatten_network(glue.compress_pool_by_svd(
        [glue.ucc_generator(so) for so in sos])), n)
    th = np.random.default_rng(SEED_BASE + 2000).uniform(-0.05, 0.05, len(sos))
    t = time.perf_counter()
    ev, _, _ = R.energy_objective(bp, net, pool, th, dim)
    t_vqe = time.perf_counter() - t
    out["other"] = {
        "davidson_12o_from_hf": dict(iterations=its, converged=conv, seconds=round(t_dav, 3),
                                     s_per_iteration=round(t_dav / its, 4), energy=e0,
                                     note="first iterations of the 80 the solve takes to 1e-8"),
        "vqe_12o_energy_grad": dict(parameters=len(sos), seconds=round(t_vqe, 3), energy=ev),
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="skip the FCI solves and VQE checks")
    ap.add_argument("--full", action="store_true",
                    help="also run the 4-root FCI solves at 16 orbitals (plain and mixed; ~15 min)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-unit", action="store_true",
                    help="one FP64 sigma and one plain copy of the same size (counter runs)")
    ap.add_argument("--vqe-only", type=int, choices=(12, 16), default=None,
                    help="only the VQE energy + adjoint gradient at 12 or 16 orbitals (profiling runs)")
    ap.add_argument("--davidson-only", type=int, default=None, metavar="MAXITER",
                    help="only the 16-orbital ground state to 1e-6: plain FP64, then mixed (block 1, "
                         "maxspace 32), each stopped after MAXITER iterations (600: to convergence)")
    ap.add_argument("--sigma-shape", default=None, metavar="N,K",
                    help="only time FP64 and single-vector FP32 sigma for N orbitals, K+K electrons")
    ap.add_argument("--tune", action="store_true",
                    help="run the sigma autotuner on the workload first (about a minute) instead "
                         "of applying the stored choice")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU code path (the sigma "
              "would shard by destination alpha rows after an all-gather of x; not built)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    if args.vqe_only:
        m = args.vqe_only
        r = vqe(torch, m, m // 2, m // 2, nfd=0, seed_off=2000, amp=0.05 if m == 12 else 0.02,
                stored_tuning=True)
        print(json.dumps({"vqe": m, **r}))
        return
    if args.davidson_only:
        print(json.dumps(fci16_ground(torch, args.davidson_only)))
        return
    if args.sigma_shape:
        nn, kk = (int(v) for v in args.sigma_shape.split(","))
        print(json.dumps(sigma_shape(torch, nn, kk, args.steps, args.warmup)))
        return
    n, na, nb = 16, 8, 8
    t = time.perf_counter()
    B, H, _, F = build(n, na, nb)
    setup_s = time.perf_counter() - t
    from binsim_amd import autotune as AT
    if args.tune:
        entry, source = AT.tune(F, B, H, k=4, block_probe=False, path=None, log=log), "tuned now"
    else:
        entry = AT.lookup(B)
        source = "binsim_amd/tuning.json" if entry is not None else "library defaults"
        if entry is not None:
            AT.apply(entry)
    tuning = {"source": source, "device": AT.device_key(), "fp64": AT.get_config(AT.FP64),
              "fp32": AT.get_config(AT.FP32)}
    log(f"launch choices ({source}): {tuning}")
    x = seeded(torch, B.dim, 1000 + 16)
    y = torch.empty_like(x)
    if args.pmc_unit:
        F.hvec(x, y)
        torch.mul(x, 1.0, out=y)  # calibration: reads and writes 8 * dim bytes
        torch.cuda.synchronize()
        log(f"pmc unit done: one sigma, one copy of {8 * B.dim} bytes")
        return
    ms, wall = timed_sigma(torch, F, x, y, args.steps, args.warmup)
    bs = sigma_bytes(n, na)
    gbs = bs / (ms * 1e-3) / 1e9
    log(f"sigma fp64: {ms:.2f} ms/step, {gbs:.1f} GB/s algorithmic")
    # block FP32 sigma (k = 4), per vector
    xs = [seeded(torch, B.dim, 2000 + i) for i in range(4)]
    ys = [torch.empty_like(v) for v in xs]
    F.hvec_block(xs, ys, 1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    F.hvec_block(xs, ys, 1)
    e1.record()
    torch.cuda.synchronize()
    ms32 = e0.elapsed_time(e1) / 4
    del xs, ys
    result = {
        "metric": "FCI solve s + sigma GB/s, mixed vs FP64 at |dE|<1e-9 Ha; VQE adjoint grads/s",
        "value": round(gbs, 2),
        "unit": "GB/s (algorithmic HBM bytes of one FP64 sigma build)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded integrals and vectors)",
        "config": {"workload": "FCI sigma build H x, 16 orbitals 8a8b (165,636,900 determinants)",
                   "norb": n, "nalpha": na, "nbeta": nb, "dim": B.dim, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                     "traffic": SIGMA_TRAFFIC_16O,
                     "algorithmic_bytes_per_launch": bs,
                     "kernel": "one sigma build (hvec_cuda): k_gauge_in, 2 row SpMMs (k_rbspmm, "
                               "source-range passes), k_ab_src, k_finalize on the default stream, "
                               "timed with HIP events as one unit"},
        "sigma_fp32_block4_ms_per_vector": round(ms32, 3),
        "tuning": tuning,
        "setup_s": round(setup_s, 3), "wall_s_timed": round(wall, 3),
    }
    if not args.quick:
        if args.full:
            result["fci_16o_4roots"] = fci_block(torch, B, H, 4, maxspace=64, tol=1e-6, maxiter=400)
        del B, H, F, x, y
        torch.cuda.empty_cache()
        from binsim_amd import davidson as D
        B0, H0, _, _ = build(12, 6, 6)
        e12 = AT.lookup(B0)
        for prec in (AT.FP64, AT.FP32):
            AT.set_config(prec, e12["fp64" if prec == AT.FP64 else "fp32"]["config"] if e12 else {})
        hf = torch.zeros(B0.dim, dtype=torch.float64, device="cuda")
        hf[0] = 1.0
        e, _, info = D.solve(B0, H0, nroots=1, mixed=False, tol=1e-8, guess=hf)
        result["fci_12o_1root_from_hf"] = dict(seconds=round(info["seconds"], 4), energy=float(e[0]),
                                               iterations=int(info["iterations"]))
        result["fci_12o_4roots"] = fci_block(torch, B0, H0, 4, maxspace=32, tol=1e-8, maxiter=400)
        del B0, H0
        result["vqe_12o"] = vqe(torch, 12, 6, 6, nfd=8)
        torch.cuda.empty_cache()
        # the 16-orbital launch choices again (the 12-orbital ones differ)
        AT.set_config(AT.FP64, tuning["fp64"])
        AT.set_config(AT.FP32, tuning["fp32"])
        result["vqe_16o"] = vqe(torch, 16, 8, 8, nfd=2, seed_off=2000, amp=0.02)
    if not args.no_cpu:
        cb = cpu_baseline()
        if cb is not None:
            result["cpu_baseline"] = cb
    print(json.dumps(result))


if __name__ == "__main__":
    main()

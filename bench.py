#!/usr/bin/env python3
"""Benchmark of the on-the-fly FCI / VQE operator path on one MI355X.

Workload (the headline configuration that fits one GPU): 16 spatial orbitals, 8 alpha + 8 beta
electrons, 12,870^2 = 165,636,900 determinants, seeded synthetic integrals (n = 16 recipe of
binsim_amd/integrals.py).  A step is one FP64 sigma build H x (hvec_cuda) on a resident,
seeded vector; ``value`` is its algorithmic HBM rate  B_sigma / t  with
B_sigma = 8 * dim * (2 + s_a + d_a) bytes (s_a = 64 singles, d_a = 784 doubles per string).

Reported beside it (one run each, outside the timed steps, in this order; --budget seconds
bound the whole run, 520 by default; about 450 s are used):
  * FCI, lowest 4 roots at the 16-orbital workload to residual 1e-6: plain FP64 Davidson (one
    FP64 sigma per correction vector) and mixed-precision block Davidson (k = 4: FP32 block
    sigma, FP64 Rayleigh-Ritz, FP64 refinement), wall seconds, sigma counts and per-root |dE|;
  * VQE energy + adjoint gradient (fused) at 16 orbitals (5,792 parameters): gradients/s and the
    max deviation from central differences (h = 1e-6) on 64 sampled parameters (as many as the
    budget allows; fd_params says how many ran);
  * FCI at 12 orbitals / 6+6 electrons, one root from HF to 1e-8 (the CPU-runnable case), and
    the VQE energy + gradient there (1,818 parameters);
  * cpu_baseline: the reference's own CPU sigma (oracle/_ref/libotf_native, OpenMP) at 12
    orbitals on a bounded sample, all host cores, and its Davidson and VQE evaluation there.

    python bench.py [--steps K] [--warmup W] [--quick] [--no-fci16] [--fci12]
    python bench.py --davidson-only 600     # 16-orbital ground state, plain vs mixed
    python bench.py --sigma-shape 18,7      # FP64 / FP32 sigma at another shape
    python bench.py --vqe-only 16           # VQE energy + gradient only
    python bench.py --cpu-sigma 16          # the reference CPU sigma at 16 orbitals only
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

PEAK_MFMA_F64_TFLOPS = 78.6  # dense FP64 MFMA

# HBM bytes per launch of the FP64 matrix-core alpha-beta kernel at 16 orbitals 8a8b (one launch
# per destination class, two per sigma), from two rocprofv3 --pmc runs of `bench.py --pmc-unit`
# (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads
# for a plain copy kernel of 1.325 GB in the same run (profiles/ab_mma_16o_pmc.md)
AB_MMA_TRAFFIC_16O = 50.75e9


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


def fci_block(torch, B, H, nroots, maxspace, tol, maxiter, maxiter_mixed=None):
    """Plain FP64 Davidson (one FP64 sigma per correction vector) vs mixed-precision block Davidson
    (k = nroots, FP32 block sigma, FP64 Rayleigh-Ritz / residuals, FP64 refinement), both from the
    same P-space start; per-root |dE| and the wall-time ratio."""
    from binsim_amd import davidson as D
    out = {"nroots": nroots, "tol": tol, "maxspace": maxspace, "block": nroots,
           "start": "lowest eigenvectors of H in the 256 lowest-diagonal determinants (P space)"}
    for name, mixed, mi in (("plain_fp64", False, maxiter), ("mixed_block", True, maxiter_mixed or maxiter)):
        e, _, info = D.solve(B, H, nroots=nroots, mixed=mixed, tol=tol, maxiter=mi, block=nroots,
                             maxspace=maxspace, verbose=True)
        out[name] = dict(seconds=round(info["seconds"], 3), iterations=int(info["iterations"]),
                         maxiter=int(mi), sigma_fp64=int(info["sigma64"]), sigma_fp32=int(info["sigma32"]),
                         converged=bool(info["converged"] == 1),
                         max_residual=info["max_residual"], energies=[float(v) for v in e])
        if mixed:
            out[name]["fp64_refinement_from_iteration"] = int(info["switch_iter"])
        log(f"fci {name}: {info['seconds']:.2f} s, {int(info['iterations'])} it, e = {e}")
    import numpy as np
    de = np.abs(np.array(out["plain_fp64"]["energies"]) - np.array(out["mixed_block"]["energies"]))
    out["abs_dE_per_root"] = [float(v) for v in de]
    out["max_abs_dE"] = float(de.max())
    out["speedup_mixed_vs_plain"] = round(out["plain_fp64"]["seconds"] / out["mixed_block"]["seconds"], 3)
    return out


def sigma_shape(torch, n, k, steps, warmup):
    """FP64, single-vector FP32 and 4-vector FP32 sigma for n orbitals, k+k electrons, on seeded
    vectors, wall clock around ``steps`` calls after ``warmup`` (stored launch choices where
    present); the 4-vector figure's algorithmic GB/s counts all four vectors."""
    from binsim_amd import autotune as AT
    B, H, _, F = build(n, k, k)
    entry = AT.lookup(B)
    if entry is not None:
        AT.apply(entry)
    x = seeded(torch, B.dim, 1000 + n)
    y, y32 = torch.empty_like(x), torch.empty_like(x)
    g = torch.Generator(device="cuda").manual_seed(3000 + n)
    xs = [x] + [torch.rand(B.dim, dtype=torch.float64, device="cuda", generator=g) - 0.5 for _ in range(3)]
    ys = [torch.empty_like(x) for _ in range(4)]
    out = {"sigma_shape": f"{n} orbitals {k}a{k}b", "dim": B.dim, "steps": steps}
    for name, fn, w in (("fp64", lambda: F.hvec(x, y), 8), ("fp32", lambda: F.hvec_block([x], [y32], 1), 4),
                        ("fp32_block4", lambda: F.hvec_block(xs, ys, 1), 16)):
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
    out["fp32_block4"]["seconds_per_vector"] = round(out["fp32_block4"]["seconds"] / 4, 4)
    out["fp32_rel_dev"] = float((y32 - y).norm() / y.norm())
    out["fp32_block4_rel_dev"] = float((ys[0] - y).norm() / y.norm())
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


def vqe(torch, n, na, nb, nfd, seed_off=2000, amp=0.05, stored_tuning=False, fd_budget_s=None):
    """Fused energy + adjoint gradient, then central differences (h = 1e-6) on nfd parameters
    sampled with seed + 3000 (binsim_amd.vqe.fd_gradient: a running prefix state, so each trial
    state sweeps only the generators from the sampled one on, and one sigma per parameter for
    the energy difference).  With fd_budget_s the sample is cut short when the time runs out
    (fd_params says how many ran)."""
    import numpy as np
    from binsim_amd.vqe import fd_gradient
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
    hf = torch.zeros(B.dim, dtype=torch.float64, device="cuda")
    hf[0] = 1.0
    psi = hf.clone()
    lam = torch.empty_like(psi)

    def run(x):
        psi.copy_(hf)
        return F.vqe_energy_grad(idx, x, psi, lam)

    run(th)
    torch.cuda.synchronize()
    reps = 3 if n <= 12 else 1
    t = time.perf_counter()
    for _ in range(reps):
        e, g, d2, _ = run(th)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    back_to_hf = float((psi - hf).abs().max())
    rng = np.random.default_rng(SEED_BASE + 3000)
    sample = np.sort(rng.choice(npar, min(nfd, npar), replace=False))
    work = [torch.empty_like(psi) for _ in range(4)]
    t_fd = time.perf_counter()
    fd = fd_gradient(F, idx, th, sample, 1e-6, hf, work, energy=e, budget_s=fd_budget_s).tolist()
    done = sample[:len(fd)].tolist()
    t_fd = time.perf_counter() - t_fd
    dev = float(np.max(np.abs(np.array(fd) - g[done]))) if done else None
    log(f"vqe n={n}: {dt * 1e3:.1f} ms per energy+grad, fd dev {dev} over {len(done)} ({t_fd:.1f} s)")
    del work
    return dict(parameters=npar, seconds_per_eval=round(dt, 5), grads_per_s=round(1.0 / dt, 3),
                energy=e, dH2=d2, fd_max_dev=dev, fd_params=len(done), fd_seconds=round(t_fd, 2),
                fd_h=1e-6, fd_sample_seed=SEED_BASE + 3000, state_back_to_hf_max_dev=back_to_hf)


def batchgrad_time(torch, n, reps=3):
    """batchgrad_cuda (the ADAPT pool gradient: <l|G_g e^{theta_g G_g}|r> for every generator)
    at n orbitals, n/2 + n/2 electrons: the fused gradient-only pass against one launch per
    generator (sweep_mode 2), seconds per call (HIP-synchronised wall clock), and the largest
    difference between the two."""
    from binsim_amd import autotune as AT
    B, H, P, F = build(n, n // 2, n // 2, pool=True)
    entry = AT.lookup(B)
    if entry is not None:
        AT.apply(entry)
    from binsim_amd.ansatz import ucc_spin_orbitals
    ng = len(ucc_spin_orbitals(n, n // 2, n // 2))
    g = torch.Generator(device="cuda").manual_seed(77 + n)
    l = torch.rand(B.dim, dtype=torch.float64, device="cuda", generator=g) - 0.5
    r = torch.rand(B.dim, dtype=torch.float64, device="cuda", generator=g) - 0.5
    l /= l.norm()
    r /= r.norm()
    th = (torch.rand(ng, dtype=torch.float64, device="cuda", generator=g) - 0.5) * 0.04
    out = {"batchgrad": f"{n} orbitals {n // 2}a{n // 2}b", "generators": ng}
    res = {}
    base = AT.get_config(AT.FP64)
    for name, mode in (("fused", base.get("sweep_mode", 0)), ("per_generator", 2)):
        AT.set_config(AT.FP64, dict(base, sweep_mode=mode))
        gd = torch.zeros(ng, dtype=torch.float64, device="cuda")
        F.batchgrad(l, r, gd, th)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            F.batchgrad(l, r, gd, th)
        torch.cuda.synchronize()
        out[name + "_seconds"] = round((time.perf_counter() - t) / reps, 4)
        res[name] = gd.clone()
        log(f"batchgrad n={n} {name}: {out[name + '_seconds']} s")
    AT.set_config(AT.FP64, base)
    out["max_abs_diff"] = float((res["fused"] - res["per_generator"]).abs().max())
    out["speedup"] = round(out["per_generator_seconds"] / out["fused_seconds"], 2)
    return out


def rank4_time(torch, n, nsample=48):
    """tvec / backtran / batch expm (4 interleaved vectors) on a sample of generators and batch
    tran over the whole pool at n orbitals, n/2 + n/2 electrons: milliseconds per call (HIP
    events) and algorithmic GB/s from each generator's pair count (tvec: 8 B read + 8 B written
    per element of a pair plus the zero fill of dst; backtran: l, r read and written, b written;
    batch expm: nvec values of both rows read and written)."""
    from math import comb
    from binsim_amd.ansatz import ucc_spin_orbitals
    B, H, P, F = build(n, n // 2, n // 2, pool=True)
    na = nb = n // 2
    sos = ucc_spin_orbitals(n, na, nb)
    ng = len(sos)
    NA, NB = comb(n, na), comb(n, nb)

    def pairs(so):
        half = len(so) // 2
        m = {0: [0, 0], 1: [0, 0]}  # spin -> [creators, annihilators]
        for i, pos in enumerate(so):
            m[pos & 1][0 if i < half else 1] += 1
        pa = comb(n - sum(m[0]), na - m[0][1]) if sum(m[0]) else NA
        pb = comb(n - sum(m[1]), nb - m[1][1]) if sum(m[1]) else NB
        return pa * pb

    gens = list(range(0, ng, max(1, ng // nsample)))
    npairs = sum(pairs(sos[k]) for k in gens)
    g = torch.Generator(device="cuda").manual_seed(91 + n)
    l, r, x = (torch.rand(B.dim, dtype=torch.float64, device="cuda", generator=g) - 0.5 for _ in range(3))
    y, b = torch.empty_like(x), torch.empty_like(x)
    m = torch.rand(B.dim * 4, dtype=torch.float64, device="cuda", generator=g) - 0.5
    out = {"rank4": f"{n} orbitals {na}a{nb}b", "generators": ng, "sampled": len(gens)}

    def timed(fn):
        fn(gens[0])
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in gens:
            fn(k)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / len(gens)

    for name, fn, bpp, fill in (("tvec", lambda k: F.tvec(k, x, y), 32, 8),
                                ("backtran", lambda k: F.backtran(k, 0.01, l, r, b), 80, 8),
                                ("batchexpm_4", lambda k: F.batchexpm(k, 0.01, m, 4, 4), 128, 0)):
        ms = timed(fn)
        gbs = (npairs / len(gens) * bpp + fill * B.dim) / (ms * 1e-3) / 1e9
        out[name] = {"ms_per_call": round(ms, 3), "algorithmic_GBps": round(gbs, 1)}
        log(f"rank4 n={n} {name}: {ms:.3f} ms per generator, {gbs:.0f} GB/s")
    t = torch.empty(ng, dtype=torch.float64, device="cuda")
    F.batchtran(l, r, t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F.batchtran(l, r, t)
    torch.cuda.synchronize()
    out["batchtran_seconds"] = round(time.perf_counter() - t0, 4)
    log(f"rank4 n={n} batchtran over {ng} generators: {out['batchtran_seconds']} s")
    return out


def cpu_baseline(budget_s=10.0, extras=True, dav_iters=200):
    """The reference's OpenMP CPU path (oracle/_ref) on the 12-orbital system: the sigma,
    repeated (the reported value), then (extras) the Davidson ground state from HF to 1e-8 and
    one VQE energy + gradient."""
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
    if not extras:
        return out
    # the same system's Davidson from HF to 1e-8 (jl/davidson.jl over the reference libdiag and
    # sigma) and one VQE energy + gradient (jl/vqe.jl over the reference expm / sigma /
    # backgrad, all 1,818 parameters)
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
                                     s_per_iteration=round(t_dav / its, 4), energy=e0),
        "vqe_12o_energy_grad": dict(parameters=len(sos), seconds=round(t_vqe, 3), energy=ev),
    }
    return out


def cpu_sigma_shape(n, calls=1):
    """The reference's OpenMP CPU sigma (oracle/_ref) at n orbitals, n/2 + n/2 electrons, timed
    over ``calls`` calls after building the basis and network (no GPU involved)."""
    import numpy as np
    from oracle import glue
    from oracle.ref import Reference, available
    from binsim_amd.integrals import for_norb, one_body_layout, reference_layout
    if not available():
        raise SystemExit("oracle/_ref is not built")
    R = Reference()
    k = n // 2
    t = time.perf_counter()
    h, eri = for_norb(n)
    terms = R.hamiltonian_terms(one_body_layout(h), reference_layout(eri), n)
    flat = glue.flatten_network(glue.compress_by_svd(terms))
    bp, dim = R.basis(n, k, k)
    net = R.network(flat, n)
    x = np.random.default_rng(1000 + n).uniform(-1, 1, dim)
    setup = time.perf_counter() - t
    log(f"cpu sigma n={n}: setup {setup:.1f} s, dim {dim}")
    t = time.perf_counter()
    for _ in range(calls):
        R.hvec(bp, net, x)
    dt = (time.perf_counter() - t) / calls
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(cpu_sigma=f"{n} orbitals {k}a{k}b", dim=dim, calls=calls, seconds_per_sigma=round(dt, 3),
                algorithmic_GBps=round(sigma_bytes(n, k) / dt / 1e9, 3), threads=cores,
                host_cpus=os.cpu_count(), setup_s=round(setup, 1),
                kind="reference (oracle/_ref libotf_native, hvec_gather_contract_otf_f64, OpenMP)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="skip the FCI solves and VQE checks")
    ap.add_argument("--no-fci16", action="store_true",
                    help="skip the 16-orbital 4-root FCI comparison (plain vs mixed; minutes)")
    ap.add_argument("--fci12", action="store_true",
                    help="also run the 12-orbital 4-root comparison (plain vs mixed, tol 1e-8; ~20 s)")
    ap.add_argument("--budget", type=float, default=520.0,
                    help="seconds the default run may take; the FCI iteration caps and the VQE "
                         "finite-difference sample are cut to fit")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-unit", action="store_true",
                    help="one FP64 sigma and one plain copy of the same size (counter runs)")
    ap.add_argument("--vqe-only", type=int, choices=(12, 16), default=None,
                    help="only the VQE energy + adjoint gradient at 12 or 16 orbitals (profiling runs)")
    ap.add_argument("--davidson-only", type=int, default=None, metavar="MAXITER",
                    help="only the 16-orbital ground state to 1e-6: plain FP64, then mixed (block 1, "
                         "maxspace 32), each stopped after MAXITER iterations (600: to convergence)")
    ap.add_argument("--rank4", type=int, choices=(12, 16), default=None, metavar="N",
                    help="time tvec / backtran / batch expm / batch tran at N orbitals")
    ap.add_argument("--batchgrad", type=int, choices=(12, 16), default=None, metavar="N",
                    help="only time the pool gradient (batchgrad_cuda) at N orbitals: fused pass vs "
                         "one launch per generator")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only the CPU baseline at 12 orbitals (reference sigma, Davidson from HF to "
                         "1e-8, VQE energy + gradient) on the host cores")
    ap.add_argument("--cpu-sigma", type=int, default=None, metavar="N",
                    help="only time the reference CPU sigma (oracle/_ref) at N orbitals, N/2+N/2 "
                         "electrons, on the host cores (16 orbitals: about a minute on 16 cores)")
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
    t_begin = time.perf_counter()
    if args.cpu_sigma:
        print(json.dumps(cpu_sigma_shape(args.cpu_sigma)))
        return
    if args.cpu_only:
        print(json.dumps(cpu_baseline(budget_s=10.0, extras=True)))
        return
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    if args.vqe_only:
        m = args.vqe_only
        r = vqe(torch, m, m // 2, m // 2, nfd=0, seed_off=2000, amp=0.05 if m == 12 else 0.02,
                stored_tuning=True)
        print(json.dumps({"vqe": m, **r}))
        return
    if args.batchgrad:
        print(json.dumps(batchgrad_time(torch, args.batchgrad)))
        return
    if args.rank4:
        print(json.dumps(rank4_time(torch, args.rank4)))
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
    g = torch.Generator(device="cuda").manual_seed(2000)
    xs = [torch.rand(B.dim, dtype=torch.float64, device="cuda", generator=g) - 0.5 for _ in range(4)]
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
    # the dominant kernel of one FP64 sigma: the alpha-beta part on the matrix cores
    # (k_ab_mma, two launches per sigma), timed live with HIP events on the library's stream
    # (hipotf_ab_timing: events around each sigma's alpha-beta launches)
    from binsim_amd import lib as L
    import ctypes
    lib = L.load()
    tms, tsec, tlaunch = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
    L.check(lib.hipotf_ab_timing(1, None, None, None) == 0, "hipotf_ab_timing")
    for _ in range(args.steps):
        F.hvec(x, y)
    torch.cuda.synchronize()
    L.check(lib.hipotf_ab_timing(0, ctypes.byref(tms), ctypes.byref(tsec), ctypes.byref(tlaunch)) == 0,
            "hipotf_ab_timing")
    ab_section_ms = tms.value / max(1, tsec.value)
    ab_launches = tlaunch.value / max(1, tsec.value)
    # algorithmic FLOPs of the alpha-beta part: the dense product of the n(n+1)/2 x (s_a + 1)
    # integral block with the (s_a + 1) source rows, for every determinant
    ntri, nca = n * (n + 1) // 2, na * (n - na) + 1
    ab_flops = 2.0 * B.dim * nca * ntri
    per_launch_ms = ab_section_ms / ab_launches if ab_launches else ab_section_ms
    per_launch_flops = ab_flops / ab_launches if ab_launches else ab_flops
    ab_tf = per_launch_flops / (per_launch_ms * 1e-3) / 1e12
    log(f"alpha-beta: {ab_section_ms:.2f} ms per sigma, {ab_launches:.0f} matrix-core launches, "
        f"{ab_tf:.1f} TFLOP/s")
    result = {
        "metric": "FCI solve s + sigma GB/s, mixed vs FP64 at |dE|<1e-9 Ha; VQE adjoint grads/s",
        "value": round(gbs, 2),
        "unit": "GB/s (algorithmic HBM bytes of one FP64 sigma build)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded integrals and vectors)",
        "config": {"workload": "FCI sigma build H x, 16 orbitals 8a8b (165,636,900 determinants)",
                   "norb": n, "nalpha": na, "nbeta": nb, "dim": B.dim, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(ab_tf, 2), "peak": PEAK_MFMA_F64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(ab_tf / PEAK_MFMA_F64_TFLOPS, 4),
                     "traffic": AB_MMA_TRAFFIC_16O,
                     "kernel": "k_ab_mma<double, 9, 17, 16, true>: alpha-beta part of the FP64 sigma "
                               "on the FP64 MFMA units, one launch per destination class",
                     "achieved_is": "dense-product FLOPs 2 dim (s_a + 1) n(n+1)/2 per sigma divided by "
                                    "the launches, over the average launch time from HIP events on the "
                                    "library stream (hipotf_ab_timing)",
                     "avg_launch_ms": round(per_launch_ms, 3), "launches_per_sigma": ab_launches,
                     "flops_per_launch": per_launch_flops,
                     "traffic_source": "profiles/ab_mma_16o_pmc.md (FETCH_SIZE + WRITE_SIZE per "
                                       "dispatch, each calibrated against a plain copy)",
                     "sigma_hbm_model": {
                         "algorithmic_bytes": bs, "GBps": round(gbs, 2),
                         "frac_of_8TBps": round(gbs / PEAK_HBM_GBS, 4),
                         "note": "the row-stream model B_sigma = 8 dim (2 + s_a + d_a) of the whole "
                                 "sigma (the value above) counts one HBM read per source element and "
                                 "connection; the kernels reuse source rows from L2 and LDS, so it can "
                                 "exceed the HBM peak"}},
        "sigma_fp32_block4_ms_per_vector": round(ms32, 3),
        "tuning": tuning,
        "setup_s": round(setup_s, 3), "wall_s_timed": round(wall, 3),
    }

    def left():
        return args.budget - (time.perf_counter() - t_begin)

    if not args.quick:
        from binsim_amd import davidson as D
        # 16 orbitals, lowest 4 roots, residual 1e-6: plain FP64 vs mixed block (k = 4), both
        # from the same P-space start.  Both need about 256 iterations here; the iteration cap
        # only guards the time budget (measured per-iteration costs, averaged over the solves:
        # plain ~4 FP64 sigmas + 0.17 s of subspace work, mixed ~4 FP32 block-sigma vectors +
        # 0.06 s); a solve stopped by it reports converged = false.
        if not args.no_fci16:
            avail = max(0.0, left() - 65.0)  # VQE 16o, 12o checks and CPU baseline afterwards
            it_plain, it_mixed = 4 * ms * 1e-3 + 0.17, 4 * ms32 * 1e-3 + 0.06
            iters = min(600, int(avail / (it_plain + it_mixed)))
            log(f"fci 16o: {avail:.0f} s available, iteration cap {iters}")
            result["fci_16o_4roots"] = fci_block(torch, B, H, 4, maxspace=32, tol=1e-6,
                                                 maxiter=max(iters, 1))
        del x, y
        del B, H, F
        torch.cuda.empty_cache()
        # VQE at 16 orbitals: 5,792 parameters, FD on 64 sampled ones (as many as the time allows)
        fd_budget = max(5.0, left() - 40.0)
        result["vqe_16o"] = vqe(torch, 16, 8, 8, nfd=64, seed_off=2000, amp=0.02, fd_budget_s=fd_budget)
        torch.cuda.empty_cache()
        # 12 orbitals (the CPU-runnable system): one root from HF (tol 1e-8) and the VQE
        # energy + gradient with 1,818 parameters
        B0, H0, _, _ = build(12, 6, 6)
        e12 = AT.lookup(B0)
        for prec in (AT.FP64, AT.FP32):
            AT.set_config(prec, e12["fp64" if prec == AT.FP64 else "fp32"]["config"] if e12 else {})
        hf = torch.zeros(B0.dim, dtype=torch.float64, device="cuda")
        hf[0] = 1.0
        e, _, info = D.solve(B0, H0, nroots=1, mixed=False, tol=1e-8, guess=hf)
        result["fci_12o_1root_from_hf"] = dict(seconds=round(info["seconds"], 4), energy=float(e[0]),
                                               iterations=int(info["iterations"]))
        if args.fci12:
            result["fci_12o_4roots"] = fci_block(torch, B0, H0, 4, maxspace=32, tol=1e-8, maxiter=400)
        del B0, H0, hf
        result["vqe_12o"] = vqe(torch, 12, 6, 6, nfd=8)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        cb = cpu_baseline(budget_s=min(10.0, max(3.0, left() - 25.0)), extras=left() > 35.0)
        if cb is not None:
            result["cpu_baseline"] = cb
    result["bench_seconds"] = round(time.perf_counter() - t_begin, 1)
    print(json.dumps(result))


if __name__ == "__main__":
    main()

"""Finite-difference check of the fused VQE adjoint gradient (jl/vqe.jl:1-21 energy_objective).

Central differences  (E(theta + h e_j) - E(theta - h e_j)) / 2h  for a sample of parameters,
from the energy-only entry point (hipotf_vqe_energy).  The sampled positions are visited in
increasing order with a running prefix state U(theta_0..theta_{j-1})|ref> advanced by
expm_cuda, so each evaluation sweeps only the generators from j on (and one sigma) instead of
the whole ansatz.  Parameters j with equal generators elsewhere in the list are handled the
same way: only position j is perturbed.
"""
from __future__ import annotations

import numpy as np


def fd_gradient(F, idxs, thetas, sample, h: float, ref_state, work=None, energy=None,
                budget_s=None):
    """Central-difference gradient entries (E(theta + h e_j) - E(theta - h e_j)) / 2h for the
    positions j in ``sample``, E = <psi|H|psi> / <psi|psi>.

    F: HipOTF_Functions with ham and pool; idxs / thetas: the ansatz (host arrays);
    ref_state: device reference state (unchanged); work: optional list of four device vectors
    of the same size; energy: E at theta if known (else one evaluation); budget_s: stop after
    this many seconds.  Returns an array aligned with the first entries of sorted(sample)
    (all of them unless the budget ran out).

    A running prefix state U(theta_{<j}) |ref> is advanced by fused forward sweeps; per j the
    two trial states psi_+- sweep the tail from j on.  For real symmetric H,
    <psi_+|H|psi_+> - <psi_-|H|psi_-> = <d|H|s> with d = psi_+ - psi_-, s = psi_+ + psi_-, so
    the difference takes one sigma; with n_+- = <psi_+-|psi_+->,
    E_+ - E_- = <d|H|s> / n + E (n_- - n_+) / n, n = (n_+ + n_-) / 2, to first order in the
    norm deviations (~1e-15 after thousands of rotations), E the energy at theta.
    """
    import time

    import torch

    t_start = time.perf_counter()
    idxs = np.asarray(idxs, dtype=np.int64)
    th = np.asarray(thetas, dtype=np.float64)
    pos = sorted(int(j) for j in sample)
    if work is None:
        work = [torch.empty_like(ref_state) for _ in range(4)]
    pre, tp, tm, lam = work
    e0 = energy
    if e0 is None:
        tp.copy_(ref_state)
        e0 = F.vqe_energy(idxs, th, tp, lam)[0]
    pre.copy_(ref_state)
    done = 0
    out = []
    for j in pos:
        if budget_s is not None and time.perf_counter() - t_start > budget_s:
            break
        if j > done:
            F.vqe_sweep(idxs[done:j], th[done:j], pre, lam)
        done = j
        t = th[j:].copy()
        t[0] = th[j] + h
        tp.copy_(pre)
        F.vqe_sweep(idxs[j:], t, tp, lam)
        t[0] = th[j] - h
        tm.copy_(pre)
        F.vqe_sweep(idxs[j:], t, tm, lam)
        n_p, n_m = float(torch.dot(tp, tp)), float(torch.dot(tm, tm))
        torch.sub(tp, tm, out=tm)   # d
        tp.mul_(2.0).sub_(tm)       # s = 2 psi_+ - d
        F.hvec(tp, lam)
        n = 0.5 * (n_p + n_m)
        diff = float(torch.dot(tm, lam)) / n + e0 * (n_m - n_p) / n
        out.append(diff / (2.0 * h))
    return np.array(out)

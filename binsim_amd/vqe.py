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


def fd_gradient(F, idxs, thetas, sample, h: float, ref_state, work=None):
    """Central-difference gradient entries for the positions in ``sample``.

    F: HipOTF_Functions with ham and pool; idxs / thetas: the ansatz (host arrays);
    ref_state: device reference state (unchanged); work: optional list of three device vectors
    of the same size (prefix, trial state, sigma output).  Returns an array aligned with
    sorted(sample).
    """
    import torch

    idxs = np.asarray(idxs, dtype=np.int64)
    th = np.asarray(thetas, dtype=np.float64)
    pos = sorted(int(j) for j in sample)
    if work is None:
        work = [torch.empty_like(ref_state) for _ in range(3)]
    pre, trial, lam = work
    pre.copy_(ref_state)
    done = 0
    out = []
    for j in pos:
        for i in range(done, j):  # prefix: U(theta_i) for i < j
            F.expm(int(idxs[i]), float(th[i]), pre)
        done = j
        e = []
        for sgn in (1.0, -1.0):
            t = th[j:].copy()
            t[0] += sgn * h
            trial.copy_(pre)
            e.append(F.vqe_energy(idxs[j:], t, trial, lam)[0])
        out.append((e[0] - e[1]) / (2.0 * h))
    return np.array(out)

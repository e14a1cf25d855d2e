"""VQE around the fused energy + adjoint gradient (jl/vqe.jl:1-21 energy_objective).

* run_vqe / minimize_fg: the optimisation loop of jl/vqe.jl:40-69,163-209 (run_vqe,
  optimze_fg!): L-BFGS with a history of 10 on (energy, gradient) evaluations, x written to a
  checkpoint on every objective call (jl/vqe.jl:170-174) and read back by load_x
  (jl/vqe.jl:122-129) to resume an interrupted run.  The reference uses Optim.jl's LBFGS(m=10)
  with a More-Thuente line search; this loop uses scipy's L-BFGS-B (maxcor 10, no bounds) and its
  own line search, so the iterates are not the reference's (parity of the path: the objective
  is pinned, the optimiser trajectory is not).
* The checkpoint is plain binary (float64, native order) plus a JSON sidecar (<path>.json: count,
  dtype, objective calls so far), each replaced atomically.
* fd_gradient: central differences  (E(theta + h e_j) - E(theta - h e_j)) / 2h  for a sample of
  parameters, from the energy-only entry point (hipotf_vqe_energy).  The sampled positions are
  visited in increasing order with a running prefix state U(theta_0..theta_{j-1})|ref> advanced
  by the forward sweep, so each evaluation sweeps only the generators from j on (and one sigma)
  instead of the whole ansatz.  Parameters j with equal generators elsewhere in the list are
  handled the same way: only position j is perturbed.
"""
from __future__ import annotations

import json
import os

import numpy as np


def save_x(path: str, x, **meta) -> None:
    """Write x (float64) to ``path`` and its sidecar ``path + ".json"``, each atomically."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(x.tobytes())
    os.replace(tmp, path)
    side = dict(meta, n=int(x.size), dtype="float64")
    with open(tmp, "w") as f:
        json.dump(side, f)
    os.replace(tmp, path + ".json")


def load_x(path: str) -> np.ndarray:
    """The parameters saved by save_x (jl/vqe.jl:122-129 load_x); raises if the file is
    missing or its length disagrees with the sidecar."""
    if not os.path.isfile(path):
        raise FileNotFoundError(path)
    x = np.fromfile(path, dtype=np.float64)
    side_path = path + ".json"
    if os.path.isfile(side_path):
        with open(side_path) as f:
            n = json.load(f).get("n")
        if n is not None and n != x.size:
            raise ValueError(f"{path}: {x.size} values, sidecar says {n}")
    return x


def minimize_fg(fg, x0, save_path: str = "", maxiter: int = 1000, gtol: float = 1e-6, m: int = 10,
                ftol: float = 2.220446049250313e-09, log=None):
    """Minimise f with fg(x) -> (f, grad) by L-BFGS (history m; stops when the largest gradient
    entry is below gtol or the relative decrease of f below ftol).  With save_path, x is saved
    before every objective call, so an interrupted run resumes from load_x(save_path) (without
    the L-BFGS history, which is rebuilt).  Returns (f_min, x_min, info) with
    info = {nit, nfev, success, message}."""
    from scipy.optimize import minimize

    calls = [0]

    def f(x):
        if save_path:
            save_x(save_path, x, calls=calls[0])
        e, g = fg(x)
        calls[0] += 1
        if log is not None:
            log(f"  f: {e: .12f}  |g|: {np.linalg.norm(g):.3e}")
        return e, np.asarray(g, dtype=np.float64)

    res = minimize(f, np.asarray(x0, dtype=np.float64), jac=True, method="L-BFGS-B",
                   options=dict(maxiter=maxiter, gtol=gtol, maxcor=m, ftol=ftol))
    return float(res.fun), np.asarray(res.x), dict(nit=int(res.nit), nfev=int(res.nfev),
                                                     success=bool(res.success), message=str(res.message))


def run_vqe(F, idxs, x0, ref_state, save_path: str = "", maxiter: int = 1000, gtol: float = 1e-6,
            m: int = 10, ftol: float = 2.220446049250313e-09, work=None, log=None):
    """jl/vqe.jl:163-209 run_vqe on the device: minimise the energy of U(x)|ref> over x with the
    fused energy + adjoint gradient (F.vqe_energy_grad), the state reset to ref_state before
    every evaluation as the reference does.  Returns (e_opt, psi, x_opt, info): psi the device
    state U(x_opt)|ref> (a new vector unless work supplies two), info as minimize_fg's plus
    the last evaluation's dH2."""
    import torch

    idxs = np.asarray(idxs, dtype=np.int64)
    if work is None:
        work = [torch.empty_like(ref_state) for _ in range(2)]
    psi, lam = work
    last = {}

    def fg(x):
        psi.copy_(ref_state)
        e, g, d2, _ = F.vqe_energy_grad(idxs, x, psi, lam)
        last["dH2"] = d2
        return e, g

    e_opt, x_opt, info = minimize_fg(fg, x0, save_path=save_path, maxiter=maxiter, gtol=gtol, m=m,
                                     ftol=ftol, log=log)
    info["dH2"] = last.get("dH2")
    psi.copy_(ref_state)
    F.vqe_sweep(idxs, x_opt, psi, lam)
    return e_opt, psi, x_opt, info


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

"""The reference's Davidson outer loop (jl/davidson.jl:1-185) driving libhipotf's device
libdiag entry points -- what a Julia GPU driver (``hipbinsim.jl``) would run.

Vectors live on the device (torch CUDA tensors); V / AV are passed as host arrays of device
pointers, the small subspace matrix stays on the host and is diagonalised with numpy, exactly
as the Julia loop calls LAPACK ``eigen``.  Same defaults: ncv = 3, maxspace = ncv + 20,
kappa = 0.717, shift = max(|r| 1e-2, 1e-4), restart to the ncv lowest Ritz vectors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


def _pp(vs):
    arr = (C.c_void_p * max(len(vs), 1))(*[int(v.data_ptr()) for v in vs])
    return C.cast(arr, L._dpp)


def davidson(aop, v0, diags, tol: float = 1e-5, ncv: int = 3, maxspace: int | None = None,
             maxiter: int = 100000, verbose: bool = False):
    """Returns (e_best, x, iterations).  ``aop(v, Av)`` writes H v into Av (device tensors)."""
    import torch

    lib = L.load()
    maxspace = ncv + 20 if maxspace is None else maxspace
    assert maxspace > ncv
    N = v0.numel()
    V = [v0.clone()]
    AV = [torch.empty_like(v0)]
    lib.inplace_normalize_f64(N, V[0].data_ptr())
    aop(V[0], AV[0])
    heff = np.zeros((maxspace, maxspace))
    heff[0, 0] = lib.fast_dot_f64(N, V[0].data_ptr(), AV[0].data_ptr())
    axt = torch.empty_like(v0)
    dim = 1
    for it in range(1, maxiter + 1):
        es, vs = np.linalg.eigh(heff[:dim, :dim])
        k = int(np.argmin(es))
        e = float(es[k])
        c = np.ascontiguousarray(vs[:, k])
        nr = lib.build_ritz_and_residual_f64(N, dim, c.ctypes.data_as(L._dp), e, _pp(V[:dim]),
                                             _pp(AV[:dim]), axt.data_ptr())
        L.check(np.isfinite(nr), "build_ritz_and_residual_f64")
        if verbose:
            print(f"  Step {it:03d}:  residual: {nr:.4e},  e: {e:.14f}", flush=True)
        if nr < tol:
            xt = torch.empty_like(v0)
            lib.build_ritz_vector_f64(N, dim, c.ctypes.data_as(L._dp), _pp(V[:dim]), xt.data_ptr())
            return e, xt, it
        if dim >= maxspace:
            nsave = min(ncv, dim)
            order = np.argsort(es)
            Vn, AVn = [], []
            for i in range(nsave):
                ci = np.ascontiguousarray(vs[:, order[i]])
                t1, t2 = torch.empty_like(v0), torch.empty_like(v0)
                lib.build_ritz_vector_f64(N, dim, ci.ctypes.data_as(L._dp), _pp(V[:dim]), t1.data_ptr())
                lib.build_ritz_vector_f64(N, dim, ci.ctypes.data_as(L._dp), _pp(AV[:dim]), t2.data_ptr())
                Vn.append(t1)
                AVn.append(t2)
            for i in range(nsave):
                V[i].copy_(Vn[i])
                AV[i].copy_(AVn[i])
            dim = nsave
            hb = np.zeros(maxspace * maxspace)
            lib.rebuild_heff_f64(N, dim, _pp(V[:dim]), _pp(AV[:dim]), hb.ctypes.data_as(L._dp), maxspace)
            heff[:, :] = hb.reshape(maxspace, maxspace).T
        shift = max(nr * 1e-2, 1e-4)
        lib.apply_preconditioner_inplace_f64(N, axt.data_ptr(), diags.data_ptr(), e, shift)
        nb = float(torch.linalg.vector_norm(axt))
        na = lib.mgs_orthogonalize_f64(N, dim, _pp(V[:dim]), axt.data_ptr())
        if na < 0.717 * nb:
            na = lib.mgs_orthogonalize_f64(N, dim, _pp(V[:dim]), axt.data_ptr())
        if na < 1e-12:
            xt = torch.empty_like(v0)
            lib.build_ritz_vector_f64(N, dim, c.ctypes.data_as(L._dp), _pp(V[:dim]), xt.data_ptr())
            return e, xt, it
        lib.inplace_normalize_f64(N, axt.data_ptr())
        dim += 1
        if len(V) < dim:
            V.append(torch.empty_like(v0))
            AV.append(torch.empty_like(v0))
        V[dim - 1].copy_(axt)
        aop(V[dim - 1], AV[dim - 1])
        col = np.zeros(dim)
        lib.compute_heff_col_f64(N, dim, _pp(V[:dim]), AV[dim - 1].data_ptr(), col.ctypes.data_as(L._dp))
        heff[:dim, dim - 1] = col
        heff[dim - 1, :dim] = col
    raise RuntimeError("davidson: maximum iterations reached")


def solve(basis, ham, nroots: int = 1, mixed: bool = False, tol: float = 1e-8,
          switch_tol: float = 1e-5, maxiter: int = 500, maxspace: int = 0, verbose: bool = False,
          guess=None, want_vectors: bool = False, block: int = 0):
    """Device-resident Davidson (hipotf_davidson_f64).  Returns (energies, vectors, info)."""
    import torch

    lib = L.load()
    dim = basis.dim
    e = np.zeros(nroots)
    info = np.zeros(7)
    vecs = torch.empty(nroots * dim, dtype=torch.float64, device="cuda") if want_vectors else None
    rc = lib.hipotf_davidson_f64(basis.ptr, ham.ptr, nroots, 1 if mixed else 0, block, tol, switch_tol,
                                 maxiter, maxspace, 1 if verbose else 0,
                                 0 if guess is None else int(guess.data_ptr()), e.ctypes.data_as(L._dp),
                                 0 if vecs is None else int(vecs.data_ptr()), info.ctypes.data_as(L._dp))
    L.check(rc != 1, "hipotf_davidson_f64")
    keys = ["iterations", "sigma64", "sigma32", "max_residual", "seconds", "switch_iter", "converged"]
    return e, vecs, dict(zip(keys, info.tolist()))

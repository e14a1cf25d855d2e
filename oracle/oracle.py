"""ORACLE (test infrastructure only) -- Python front end of the clean-room restatement.

Builds the reference's operator networks from integrals / spin-orbital lists with
:mod:`oracle.glue` and applies them with ``liboracle.so`` (oracle.c).  Also restates the
Julia drivers the reference runs on top of the operator library:

* :func:`davidson` -- jl/davidson.jl:1-185 (ncv=3, maxspace=ncv+20, kappa=0.717, the shift
  rule and restart), over the libdiag operations of oracle.c;
* :func:`energy_objective` -- jl/vqe.jl:1-21.

Parity status of each function is recorded in the header of tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import glue

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)


class _Net(C.Structure):
    _fields_ = [("ngs", C.c_int64), ("axs", _u32p), ("bxs", _u32p), ("ranks", _i64p),
                ("num_zas", _i64p), ("num_zbs", _i64p), ("flat_zas", _u32p), ("flat_zbs", _u32p),
                ("flat_wa", _dp), ("flat_wb", _dp)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER(_Net)
        L.oracle_hvec.argtypes = [C.c_int, C.c_int, C.c_int, P, _dp, _dp]
        L.oracle_diags.argtypes = [C.c_int, C.c_int, C.c_int, P, _dp]
        L.oracle_expm.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int64, C.c_double, _dp]
        L.oracle_grad.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int64, C.c_double, _dp, _dp]
        L.oracle_grad.restype = C.c_double
        L.oracle_backgrad.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int64, C.c_double, _dp, _dp]
        L.oracle_backgrad.restype = C.c_double
        L.oracle_tvec.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int64, _dp, _dp]
        L.oracle_backtran.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int64, C.c_double, _dp, _dp, _dp]
        L.oracle_tran.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int64, _dp, _dp]
        L.oracle_tran.restype = C.c_double
        L.oracle_batch_expm.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int64, C.c_double, _dp,
                                        C.c_int64, C.c_int]
        pp = C.POINTER(_dp)
        L.oracle_ritz_and_residual.argtypes = [C.c_int64, C.c_int, _dp, C.c_double, pp, pp, _dp]
        L.oracle_ritz_and_residual.restype = C.c_double
        L.oracle_ritz_vector.argtypes = [C.c_int64, C.c_int, _dp, pp, _dp]
        L.oracle_precondition.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double]
        L.oracle_mgs.argtypes = [C.c_int64, C.c_int, pp, _dp]
        L.oracle_mgs.restype = C.c_double
        L.oracle_normalize.argtypes = [C.c_int64, _dp]
        L.oracle_normalize.restype = C.c_double
        L.oracle_dot.argtypes = [C.c_int64, _dp, _dp]
        L.oracle_dot.restype = C.c_double
        L.oracle_heff_col.argtypes = [C.c_int64, C.c_int, pp, _dp, _dp]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Network:
    """A flattened SVD-group network (the arrays build_network_otf_f64 receives)."""

    def __init__(self, groups):
        self.groups = groups
        self.flat = glue.flatten_network(groups)
        f = self.flat
        self._s = _Net(f["ngs"], _p(f["axs"], _u32p), _p(f["bxs"], _u32p), _p(f["ranks"], _i64p),
                       _p(f["num_zas"], _i64p), _p(f["num_zbs"], _i64p), _p(f["flat_zas"], _u32p),
                       _p(f["flat_zbs"], _u32p), _p(f["flat_wa"]), _p(f["flat_wb"]))

    @property
    def ref(self):
        return C.byref(self._s)


def hamiltonian_network(h1_flat, g2_flat, n, tol=1e-12):
    return Network(glue.compress_by_svd(glue.jw_hamiltonian_terms(h1_flat, g2_flat, n, tol)))


def pool_network(spin_orbitals):
    return Network(glue.compress_pool_by_svd([glue.ucc_generator(so) for so in spin_orbitals]))


class System:
    """Basis (n, na, nb) with the oracle's operators bound to it."""

    def __init__(self, n, na, nb, ham: Network | None = None, pool: Network | None = None):
        self.n, self.na, self.nb = n, na, nb
        from math import comb
        self.NA, self.NB = comb(n, na), comb(n, nb)
        self.dim = self.NA * self.NB
        self.ham, self.pool = ham, pool

    def hvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().oracle_hvec(self.n, self.na, self.nb, self.ham.ref, _p(x), _p(y))
        return y

    def diags(self):
        d = np.zeros(self.dim)
        lib().oracle_diags(self.n, self.na, self.nb, self.ham.ref, _p(d))
        return d

    def expm(self, g, theta, v):
        lib().oracle_expm(self.n, self.na, self.nb, self.pool.ref, g, theta, _p(v))

    def grad(self, g, theta, l, r):
        return lib().oracle_grad(self.n, self.na, self.nb, self.pool.ref, g, theta, _p(l), _p(r))

    def backgrad(self, g, theta, l, r):
        return lib().oracle_backgrad(self.n, self.na, self.nb, self.pool.ref, g, theta, _p(l), _p(r))

    def tvec(self, g, src):
        """G_g src (op/tvec.hpp:373-445)"""
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.empty_like(src)
        lib().oracle_tvec(self.n, self.na, self.nb, self.pool.ref, g, _p(src), _p(dst))
        return dst

    def backtran(self, g, theta, l, r):
        """l, r rotated in place by exp(theta G_g); returns G_g l (op/backtran.hpp:383-452)"""
        b = np.empty_like(l)
        lib().oracle_backtran(self.n, self.na, self.nb, self.pool.ref, g, theta, _p(l), _p(r), _p(b))
        return b

    def batch_tran(self, ngen, l, r):
        """[sum of tran_update over the pairs of G_k] for k < ngen (op/batchtran.hpp:517-528)"""
        l = np.ascontiguousarray(l, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        return np.array([lib().oracle_tran(self.n, self.na, self.nb, self.pool.ref, k, _p(l), _p(r))
                         for k in range(ngen)])

    def batch_expm(self, g, theta, m, ld, nvec):
        """exp(theta G_g) on columns 0..nvec-1 of the det-major matrix m (row stride ld),
        in place (op/batchexpm.hpp:376-444)"""
        assert m.dtype == np.float64 and m.flags.c_contiguous and m.size >= self.dim * ld
        lib().oracle_batch_expm(self.n, self.na, self.nb, self.pool.ref, g, theta, _p(m), ld, nvec)


def hf_vector(sysm: System):
    v = np.zeros(sysm.dim)
    v[0] = 1.0  # lowest-bit strings are index 0 of both spins (jl/network.jl:81-116)
    return v


def energy_objective(sysm: System, idxs, x, lv, rv):
    """jl/vqe.jl:1-21 with 0-based pool indices."""
    for i, t in zip(idxs, x):
        sysm.expm(i, t, lv)
    rv[:] = sysm.hvec(lv)
    lnorm = float(lv @ lv)
    rnorm = float(rv @ rv)
    e = float(lv @ rv) / lnorm
    d2 = max(0.0, rnorm / lnorm - e * e)
    g = np.empty(len(x))
    for k in range(len(x) - 1, -1, -1):
        g[k] = sysm.backgrad(idxs[k], x[k], lv, rv) * 2.0 / lnorm
    return e, g, d2


def _ptrs(vs):
    arr = (_dp * max(len(vs), 1))()
    for i, v in enumerate(vs):
        arr[i] = _p(v)
    return arr


def davidson(aop, v0, diags, tol=1e-5, ncv=3, maxspace=None, maxiter=100000):
    """jl/davidson.jl:1-185 over oracle.c's libdiag restatement.  Returns (e, x, iters)."""
    L = lib()
    maxspace = ncv + 20 if maxspace is None else maxspace
    N = len(v0)
    V = [np.array(v0, dtype=np.float64)]
    AV = [np.empty(N)]
    L.oracle_normalize(N, _p(V[0]))
    AV[0][:] = aop(V[0])
    heff = np.zeros((maxspace, maxspace))
    heff[0, 0] = L.oracle_dot(N, _p(V[0]), _p(AV[0]))
    axt = np.empty(N)
    dim = 1
    for it in range(1, maxiter + 1):
        es, vs = np.linalg.eigh(heff[:dim, :dim])
        k = int(np.argmin(es))
        e = float(es[k])
        c = np.ascontiguousarray(vs[:, k])
        nr = L.oracle_ritz_and_residual(N, dim, _p(c), e, _ptrs(V[:dim]), _ptrs(AV[:dim]), _p(axt))
        if nr < tol:
            xt = np.empty(N)
            L.oracle_ritz_vector(N, dim, _p(c), _ptrs(V[:dim]), _p(xt))
            return e, xt, it
        if dim >= maxspace:
            nsave = min(ncv, dim)
            order = np.argsort(es)
            Vn = [sum(vs[j, order[i]] * V[j] for j in range(dim)) for i in range(nsave)]
            AVn = [sum(vs[j, order[i]] * AV[j] for j in range(dim)) for i in range(nsave)]
            for i in range(nsave):
                V[i][:] = Vn[i]
                AV[i][:] = AVn[i]
            dim = nsave
            for cc in range(dim):
                for rr in range(dim):
                    heff[rr, cc] = L.oracle_dot(N, _p(V[rr]), _p(AV[cc]))
        shift = max(nr * 1e-2, 1e-4)
        L.oracle_precondition(N, _p(axt), _p(diags), e, shift)
        nb = float(np.linalg.norm(axt))
        na = L.oracle_mgs(N, dim, _ptrs(V[:dim]), _p(axt))
        if na < 0.717 * nb:
            na = L.oracle_mgs(N, dim, _ptrs(V[:dim]), _p(axt))
        if na < 1e-12:
            xt = np.empty(N)
            L.oracle_ritz_vector(N, dim, _p(c), _ptrs(V[:dim]), _p(xt))
            return e, xt, it
        L.oracle_normalize(N, _p(axt))
        dim += 1
        if len(V) < dim:
            V.append(np.empty(N))
            AV.append(np.empty(N))
        V[dim - 1][:] = axt
        AV[dim - 1][:] = aop(V[dim - 1])
        col = np.empty(dim)
        L.oracle_heff_col(N, dim, _ptrs(V[:dim]), _p(AV[dim - 1]), _p(col))
        heff[:dim, dim - 1] = col
        heff[dim - 1, :dim] = col
    raise RuntimeError("davidson: maxiter reached")

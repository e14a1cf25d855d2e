"""ORACLE (test infrastructure only) -- ctypes bindings to the reference's own CPU libraries.

``oracle/Makefile`` compiles them from the reference sources into ``oracle/_ref``; this
module only loads what is there.  Used by tests/ (to pin the restatement), by
``tests/golden/make_golden.py`` and by the CPU-baseline leg of ``bench.py``; never by the
product.

Signatures follow src/include/op/declare.hpp:13-132, src/include/diag/declare.hpp:4-75,
src/include/basis/declare.hpp:4-81 and src/include/ham/declare.hpp:7-27.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)


class HamResult(C.Structure):
    _fields_ = [("axs", C.c_void_p), ("azs", C.c_void_p), ("bxs", C.c_void_p),
                ("bzs", C.c_void_p), ("cs", _dp), ("ncs", C.c_size_t)]


def available() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, f)) for f in
               ("libotf_native.so", "libdiag.so", "libbasis.so", "libham_real.so"))


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Reference:
    def __init__(self):
        if not available():
            raise RuntimeError("oracle/_ref is not built (run make -C oracle ref)")
        self.otf = C.CDLL(os.path.join(REF_DIR, "libotf_native.so"))
        self.diag = C.CDLL(os.path.join(REF_DIR, "libdiag.so"))
        self.basis_lib = C.CDLL(os.path.join(REF_DIR, "libbasis.so"))
        self.ham = C.CDLL(os.path.join(REF_DIR, "libham_real.so"))
        self.libc = C.CDLL(None)
        o, d, b, h = self.otf, self.diag, self.basis_lib, self.ham
        h.generate_hamiltonian_real_64_128_f64.restype = HamResult
        h.generate_hamiltonian_real_64_128_f64.argtypes = [C.c_double, _dp, _dp, C.c_int,
                                                           C.c_double, C.c_bool]
        b.create_basis_manager.restype = C.c_void_p
        b.create_basis_manager.argtypes = [C.c_int64] * 4 + [_i64p, C.c_int64]
        b.get_subspace_dim.restype = C.c_int64
        b.get_subspace_dim.argtypes = [C.c_void_p]
        b.destroy_basis_manager.argtypes = [C.c_void_p]
        b.set_det_coeff_f64.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double, _dp]
        o.build_network_otf_f64.restype = C.c_void_p
        o.build_network_otf_f64.argtypes = [_i64p, C.c_int64, C.c_int64, _u32p, _u32p, _i64p,
                                            _i64p, _i64p, _u32p, _u32p, _dp, _dp]
        o.destroy_network_otf_f64.argtypes = [C.c_void_p]
        o.hvec_gather_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp]
        o.get_diags_elements_f64.argtypes = [C.c_void_p, C.c_void_p, _dp]
        o.expm_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, _dp]
        o.grad_contract_otf_f64.restype = C.c_double
        o.grad_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, _dp, _dp]
        o.backgrad_contract_otf_f64.restype = C.c_double
        o.backgrad_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double,
                                                _dp, _dp]
        o.tvec_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, _dp, _dp]
        o.tvec_contract_otf_f64.restype = None
        o.backtran_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, _dp, _dp, _dp]
        o.backtran_contract_otf_f64.restype = None
        o.batch_expm_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, _dp,
                                                  C.c_int, C.c_int]
        o.batch_expm_contract_otf_f64.restype = None
        o.batch_tran_contract_otf_f64.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, _dp]
        o.batch_tran_contract_otf_f64.restype = None
        pp = C.POINTER(_dp)
        d.build_ritz_and_residual_f64.restype = C.c_double
        d.build_ritz_and_residual_f64.argtypes = [C.c_int64, C.c_int, _dp, C.c_double, pp, pp, _dp]
        d.build_ritz_vector_f64.argtypes = [C.c_int64, C.c_int, _dp, pp, _dp]
        d.apply_preconditioner_inplace_f64.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double]
        d.mgs_orthogonalize_f64.restype = C.c_double
        d.mgs_orthogonalize_f64.argtypes = [C.c_int64, C.c_int, pp, _dp]
        d.inplace_normalize_f64.restype = C.c_double
        d.inplace_normalize_f64.argtypes = [C.c_int64, _dp]
        d.fast_dot_f64.restype = C.c_double
        d.fast_dot_f64.argtypes = [C.c_int64, _dp, _dp]
        d.compute_heff_col_f64.argtypes = [C.c_int64, C.c_int, pp, _dp, _dp]
        d.rebuild_heff_f64.argtypes = [C.c_int64, C.c_int, pp, pp, _dp, C.c_int]
        self.libc.free.argtypes = [C.c_void_p]

    # -- ham.hpp ------------------------------------------------------------------------
    def hamiltonian_terms(self, h1_flat, g2_flat, n, tol=1e-12):
        h1 = np.ascontiguousarray(h1_flat, dtype=np.float64)
        g2 = np.ascontiguousarray(g2_flat, dtype=np.float64)
        r = self.ham.generate_hamiltonian_real_64_128_f64(0.0, _ptr(h1, _dp), _ptr(g2, _dp),
                                                          n, tol, False)
        m = r.ncs

        def take(p, dt):
            a = np.ctypeslib.as_array(C.cast(p, C.POINTER(dt)), shape=(max(m, 1),))[:m].copy()
            self.libc.free(C.cast(p, C.c_void_p))
            return a

        axs, azs = take(r.axs, C.c_uint32), take(r.azs, C.c_uint32)
        bxs, bzs = take(r.bxs, C.c_uint32), take(r.bzs, C.c_uint32)
        cs = take(r.cs, C.c_double)
        return axs, bxs, azs, bzs, cs

    # -- basis.hpp ----------------------------------------------------------------------
    def basis(self, norb, na, nb):
        orbsym = np.zeros(max(norb, 1), dtype=np.int64)
        p = self.basis_lib.create_basis_manager(norb, na, nb, 0, _ptr(orbsym, _i64p), 16)
        return p, int(self.basis_lib.get_subspace_dim(p))

    def free_basis(self, p):
        self.basis_lib.destroy_basis_manager(p)

    # -- otf.hpp ------------------------------------------------------------------------
    def network(self, flat, norb):
        orbsym = np.zeros(max(norb, 1), dtype=np.int64)
        f = flat
        keep = [f[k] for k in ("axs", "bxs", "ranks", "num_zas", "num_zbs", "flat_zas",
                               "flat_zbs", "flat_wa", "flat_wb")]
        p = self.otf.build_network_otf_f64(
            _ptr(orbsym, _i64p), norb, f["ngs"], _ptr(keep[0], _u32p), _ptr(keep[1], _u32p),
            _ptr(keep[2], _i64p), _ptr(keep[3], _i64p), _ptr(keep[4], _i64p),
            _ptr(keep[5], _u32p), _ptr(keep[6], _u32p), _ptr(keep[7], _dp), _ptr(keep[8], _dp))
        return p

    def free_network(self, p):
        self.otf.destroy_network_otf_f64(p)

    def hvec(self, basis, net, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.otf.hvec_gather_contract_otf_f64(basis, net, _ptr(x, _dp), _ptr(y, _dp))
        return y

    def diags(self, basis, net, dim):
        d = np.zeros(dim)
        self.otf.get_diags_elements_f64(basis, net, _ptr(d, _dp))
        return d

    def expm(self, basis, net, idx, theta, v):
        self.otf.expm_contract_otf_f64(basis, net, idx, theta, _ptr(v, _dp))

    def grad(self, basis, net, idx, theta, lv, rv):
        return self.otf.grad_contract_otf_f64(basis, net, idx, theta, _ptr(lv, _dp), _ptr(rv, _dp))

    def backgrad(self, basis, net, idx, theta, lv, rv):
        return self.otf.backgrad_contract_otf_f64(basis, net, idx, theta, _ptr(lv, _dp),
                                                  _ptr(rv, _dp))

    def tvec(self, basis, net, idx, src, dst):
        self.otf.tvec_contract_otf_f64(basis, net, idx, _ptr(src, _dp), _ptr(dst, _dp))

    def backtran(self, basis, net, idx, theta, lv, rv, bv):
        self.otf.backtran_contract_otf_f64(basis, net, idx, theta, _ptr(lv, _dp), _ptr(rv, _dp), _ptr(bv, _dp))

    def batch_expm(self, basis, net, idx, theta, m, ld, nvec):
        self.otf.batch_expm_contract_otf_f64(basis, net, idx, theta, _ptr(m, _dp), ld, nvec)

    def batch_tran(self, basis, net, lv, rv, trans):
        self.otf.batch_tran_contract_otf_f64(basis, net, _ptr(lv, _dp), _ptr(rv, _dp), _ptr(trans, _dp))

    # -- davidson.hpp -------------------------------------------------------------------
    @staticmethod
    def ptr_array(vecs):
        arr = (_dp * max(len(vecs), 1))()
        for i, v in enumerate(vecs):
            arr[i] = _ptr(v, _dp)
        return arr

    def davidson(self, aop, v0, diags, tol, maxiter=100000):
        """jl/davidson.jl:1-185 (one root, ncv = 3, maxspace = ncv + 20) driving this library's
        libdiag ops.  Returns (energy, iterations, converged); stops after ``maxiter``."""
        d = self.diag
        N = len(v0)
        ncv, maxspace = 3, 23
        P = lambda a: _ptr(a, _dp)  # noqa: E731
        V, AV = [np.array(v0, dtype=np.float64)], [np.empty(N)]
        d.inplace_normalize_f64(N, P(V[0]))
        AV[0][:] = aop(V[0])
        heff = np.zeros((maxspace, maxspace))
        heff[0, 0] = d.fast_dot_f64(N, P(V[0]), P(AV[0]))
        axt = np.empty(N)
        dim, e = 1, float(heff[0, 0])
        for it in range(1, maxiter + 1):
            es, vs = np.linalg.eigh(heff[:dim, :dim])
            k = int(np.argmin(es))
            e = float(es[k])
            c = np.ascontiguousarray(vs[:, k])
            nr = d.build_ritz_and_residual_f64(N, dim, P(c), e, self.ptr_array(V[:dim]),
                                               self.ptr_array(AV[:dim]), P(axt))
            if nr < tol:
                return e, it, True
            if dim >= maxspace:  # restart to the ncv lowest Ritz vectors (davidson.jl:82-114)
                order = np.argsort(es)
                Vn = [sum(vs[j, order[i]] * V[j] for j in range(dim)) for i in range(ncv)]
                AVn = [sum(vs[j, order[i]] * AV[j] for j in range(dim)) for i in range(ncv)]
                for i in range(ncv):
                    V[i][:], AV[i][:] = Vn[i], AVn[i]
                dim = ncv
                heff_c = np.zeros(maxspace * maxspace)
                d.rebuild_heff_f64(N, dim, self.ptr_array(V[:dim]), self.ptr_array(AV[:dim]),
                                   P(heff_c), maxspace)
                heff[:, :] = heff_c.reshape(maxspace, maxspace).T
            shift = max(nr * 1e-2, 1e-4)
            d.apply_preconditioner_inplace_f64(N, P(axt), P(diags), e, shift)
            nb_ = float(np.linalg.norm(axt))
            na_ = d.mgs_orthogonalize_f64(N, dim, self.ptr_array(V[:dim]), P(axt))
            if na_ < 0.717 * nb_:
                na_ = d.mgs_orthogonalize_f64(N, dim, self.ptr_array(V[:dim]), P(axt))
            if na_ < 1e-12:
                return e, it, True
            d.inplace_normalize_f64(N, P(axt))
            dim += 1
            if len(V) < dim:
                V.append(np.empty(N))
                AV.append(np.empty(N))
            V[dim - 1][:] = axt
            AV[dim - 1][:] = aop(V[dim - 1])
            col = np.zeros(dim)
            d.compute_heff_col_f64(N, dim, self.ptr_array(V[:dim]), P(AV[dim - 1]), P(col))
            heff[:dim, dim - 1] = col
            heff[dim - 1, :dim] = col
        return e, maxiter, False

    def energy_objective(self, basis, ham, pool, x, dim):
        """jl/vqe.jl:1-21 from the Hartree-Fock state (index 0) with 0-based pool indices
        0..len(x)-1, on this library's expm / hvec / backgrad.  Returns (energy, grad, dH2)."""
        lv = np.zeros(dim)
        lv[0] = 1.0
        for i, t in enumerate(x):
            self.expm(basis, pool, i, float(t), lv)
        rv = self.hvec(basis, ham, lv)
        lnorm, rnorm = float(lv @ lv), float(rv @ rv)
        e = float(lv @ rv) / lnorm
        d2 = max(0.0, rnorm / lnorm - e * e)
        g = np.zeros(len(x))
        for i in range(len(x) - 1, -1, -1):
            g[i] = self.backgrad(basis, pool, i, float(x[i]), lv, rv) * 2.0 / lnorm
        return e, g, d2

"""Host-side mirror of the reference's GPU operator interface, over libhipotf.so.

Same objects and closures as jl/network.jl (``BasisManager``, ``get_hf``) and
jl/cuda/cunetwork.jl (``CuBasisManager``, ``CuOTF``, ``CuOTF_Functions``) with Python
conventions: generator indices are 0-based (the Julia closures subtract 1 themselves,
cunetwork.jl:125), and device vectors are anything exposing ``data_ptr()`` (a torch CUDA
tensor) or a raw integer device address.  Every call goes to the HIP library; errors raise
:class:`binsim_amd.lib.HipOTFError`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


def _ptr(v) -> int:
    if v is None:
        return 0
    if isinstance(v, int):
        return v
    if hasattr(v, "data_ptr"):
        return int(v.data_ptr())
    if isinstance(v, np.ndarray):
        return int(v.ctypes.data)
    raise TypeError(f"cannot take a device address of {type(v)}")


def _np(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class BasisManager:
    """Host basis (jl/network.jl:17-45): colex alpha/beta strings, one symmetry block."""

    def __init__(self, norb: int, na: int, nb: int, orbsym=None, total_sym: int = 0,
                 num_irreps: int = 16):
        lib = L.load()
        self.norb, self.nelec = int(norb), (int(na), int(nb))
        self.orbsym = _np(np.zeros(norb) if orbsym is None else orbsym, np.int64)
        self.ptr = lib.create_basis_manager(norb, na, nb, total_sym,
                                            self.orbsym.ctypes.data_as(L._i64p), num_irreps)
        L.check(bool(self.ptr), "create_basis_manager")
        self.dim = int(lib.get_subspace_dim(self.ptr))

    def __del__(self):
        if getattr(self, "ptr", None) and L is not None and L.load is not None:
            L.load().destroy_basis_manager(self.ptr)
            self.ptr = None

    def get_hf(self) -> np.ndarray:
        """jl/network.jl:81-116: unit vector on the lowest-orbital alpha/beta strings (host)."""
        v = np.zeros(self.dim)
        na, nb = self.nelec
        L.call_void("set_det_coeff_f64", self.ptr, (1 << na) - 1, (1 << nb) - 1, 1.0,
                    v.ctypes.data)
        return v


class HipBasisManager:
    """Device basis (jl/cuda/cunetwork.jl:13-30)."""

    def __init__(self, basis: BasisManager):
        self.host = basis
        self.ptr = L.load().build_basisdev_f64(basis.ptr)
        L.check(bool(self.ptr), "build_basisdev_f64")
        self.dim = basis.dim

    def __del__(self):
        if getattr(self, "ptr", None) and L is not None and L.load is not None:
            L.load().destroy_basisdev_f64(self.ptr)
            self.ptr = None


class HipOTF:
    """Device operator handle (jl/cuda/cunetwork.jl:38-53): a Hamiltonian or a generator pool."""

    def __init__(self, ptr, kind: str):
        L.check(bool(ptr), f"building the {kind} operator")
        self.ptr, self.kind = ptr, kind

    @classmethod
    def from_integrals(cls, norb: int, one_body_mo, two_body_mo, energy_nuc: float = 0.0):
        """Same buffers as generate_hamiltonian_real_64_128_f64 (ham/declare.hpp:7-27)."""
        h1 = _np(one_body_mo, np.float64).reshape(-1)
        g2 = _np(two_body_mo, np.float64).reshape(-1)
        assert h1.size == norb ** 2 and g2.size == norb ** 4
        p = L.load().hipotf_build_ham_f64(norb, h1.ctypes.data_as(L._dp), g2.ctypes.data_as(L._dp),
                                          float(energy_nuc))
        return cls(p, "hamiltonian")

    @classmethod
    def from_spin_orbitals(cls, norb: int, spin_orbitals):
        """Pool from generate_ci_spin_orbitals-style lists (jl/ansatz.jl:67-167)."""
        offs = np.zeros(len(spin_orbitals) + 1, dtype=np.int32)
        for i, so in enumerate(spin_orbitals):
            offs[i + 1] = offs[i] + len(so)
        orbs = np.array([o for so in spin_orbitals for o in so], dtype=np.int32)
        if orbs.size == 0:
            orbs = np.zeros(1, dtype=np.int32)
        p = L.load().hipotf_build_pool_f64(norb, len(spin_orbitals), offs.ctypes.data_as(L._i32p),
                                           orbs.ctypes.data_as(L._i32p))
        return cls(p, "pool")

    @classmethod
    def from_network(cls, norb: int, flat: dict):
        """A reference SVD network (the flat arrays of jl/network.jl:313-339)."""
        lib = L.load()
        orbsym = np.zeros(norb, dtype=np.int64)
        f = flat
        host = lib.build_network_otf_f64(
            orbsym.ctypes.data_as(L._i64p), norb, f["ngs"], f["axs"].ctypes.data_as(L._u32p),
            f["bxs"].ctypes.data_as(L._u32p), f["ranks"].ctypes.data_as(L._i64p),
            f["num_zas"].ctypes.data_as(L._i64p), f["num_zbs"].ctypes.data_as(L._i64p),
            f["flat_zas"].ctypes.data_as(L._u32p), f["flat_zbs"].ctypes.data_as(L._u32p),
            f["flat_wa"].ctypes.data_as(L._dp), f["flat_wb"].ctypes.data_as(L._dp))
        L.check(bool(host), "build_network_otf_f64")
        try:
            p = lib.build_networkdev_f64(host)
            L.check(bool(p), "build_networkdev_f64")
        finally:
            lib.destroy_network_otf_f64(host)
        return cls(p, "network")

    def __del__(self):
        if getattr(self, "ptr", None) and L is not None and L.load is not None:
            L.load().destroy_networkdev_f64(self.ptr)
            self.ptr = None


class HipOTF_Functions:
    """The closures of CuOTF_Functions (jl/cuda/cunetwork.jl:74-160) over libhipotf."""

    def __init__(self, basis: HipBasisManager, ham: HipOTF | None = None,
                 pool: HipOTF | None = None, prepare: bool = True):
        self.basis, self.ham, self.pool = basis, ham, pool
        lib = L.load()
        if prepare:
            for op in (ham, pool):
                if op is not None:
                    L.check(lib.hipotf_prepare(basis.ptr, op.ptr) == 0, "hipotf_prepare")

    # -- Hamiltonian --------------------------------------------------------------------
    def hvec(self, v, Hv):
        L.call_void("hvec_cuda", self.basis.ptr, self.ham.ptr, _ptr(v), _ptr(Hv))

    def hvec_block(self, vs, Hvs, precision: int = 0):
        n = len(vs)
        src = (C.c_void_p * n)(*[_ptr(v) for v in vs])
        dst = (C.c_void_p * n)(*[_ptr(v) for v in Hvs])
        rc = L.load().hipotf_hvec_block(self.basis.ptr, self.ham.ptr, n,
                                        C.cast(src, L._dpp), C.cast(dst, L._dpp), precision)
        L.check(rc == 0, "hipotf_hvec_block")

    def get_diags(self, dv):
        L.call_void("get_diags_elements_cuda", self.basis.ptr, self.ham.ptr, _ptr(dv))

    # -- pool ---------------------------------------------------------------------------
    def expm(self, idx: int, theta: float, v):
        L.call_void("expm_cuda", self.basis.ptr, self.pool.ptr, int(idx), float(theta), _ptr(v))

    def grad(self, idx: int, theta: float, lv, rv) -> float:
        lib = L.load()
        lib.hipotf_clear_error()
        g = lib.grad_cuda(self.basis.ptr, self.pool.ptr, int(idx), float(theta), _ptr(lv), _ptr(rv))
        L.check(not L.last_error(), "grad_cuda")
        return g

    def backgrad(self, idx: int, theta: float, lv, rv) -> float:
        lib = L.load()
        lib.hipotf_clear_error()
        g = lib.backgrad_cuda(self.basis.ptr, self.pool.ptr, int(idx), float(theta), _ptr(lv),
                              _ptr(rv))
        L.check(not L.last_error(), "backgrad_cuda")
        return g

    def tvec(self, idx: int, lv, rv):
        """rv = G_idx lv (jl/network.jl:572-574 f_tvec)"""
        L.call_void("tvec_cuda", self.basis.ptr, self.pool.ptr, int(idx), _ptr(lv), _ptr(rv))

    def backtran(self, idx: int, theta: float, lv, rv, tlv):
        """lv, rv <- exp(theta G_idx) lv, rv; tlv = G_idx lv (f_backtran)"""
        L.call_void("backtran_cuda", self.basis.ptr, self.pool.ptr, int(idx), float(theta), _ptr(lv),
                    _ptr(rv), _ptr(tlv))

    def batchexpm(self, idx: int, theta: float, mat, N: int, j: int):
        """exp(theta G_idx) on the first j of the N interleaved vectors of mat (det-major,
        mat[i * N + k]) (f_batchexpm)"""
        L.call_void("batchexpm_cuda", self.basis.ptr, self.pool.ptr, int(idx), float(theta), _ptr(mat),
                    int(N), int(j))

    def batchtran(self, lv, rv, trans):
        """trans[k] = -<lv|G_k|rv> for every pool generator (f_batchtran)"""
        L.call_void("batchtran_cuda", self.basis.ptr, self.pool.ptr, _ptr(lv), _ptr(rv), _ptr(trans))

    def batchgrad(self, lv, rv, grads, thetas):
        L.call_void("batchgrad_cuda", self.basis.ptr, self.pool.ptr, _ptr(thetas), _ptr(lv),
                    _ptr(rv), _ptr(grads))

    def vqe_energy_grad(self, idxs, thetas, psi, lam):
        """Fused jl/vqe.jl energy_objective: returns (energy, grads, dH2, |psi|^2)."""
        idx = _np(idxs, np.int64)
        th = _np(thetas, np.float64)
        g = np.empty(len(th))
        out3 = np.empty(3)
        rc = L.load().hipotf_vqe_energy_grad(
            self.basis.ptr, self.ham.ptr, self.pool.ptr, len(th), idx.ctypes.data_as(L._i64p),
            th.ctypes.data_as(L._dp), _ptr(psi), _ptr(lam), g.ctypes.data_as(L._dp),
            out3.ctypes.data_as(L._dp))
        L.check(rc == 0, "hipotf_vqe_energy_grad")
        return float(out3[0]), g, float(out3[1]), float(out3[2])

    def vqe_energy(self, idxs, thetas, psi, lam):
        """Energy only (no gradient): returns (energy, dH2, |psi|^2); psi ends as U(theta)|ref>."""
        idx = _np(idxs, np.int64)
        th = _np(thetas, np.float64)
        out3 = np.empty(3)
        rc = L.load().hipotf_vqe_energy(
            self.basis.ptr, self.ham.ptr, self.pool.ptr, len(th), idx.ctypes.data_as(L._i64p),
            th.ctypes.data_as(L._dp), _ptr(psi), _ptr(lam), out3.ctypes.data_as(L._dp))
        L.check(rc == 0, "hipotf_vqe_energy")
        return float(out3[0]), float(out3[1]), float(out3[2])

    def vqe_sweep(self, idxs, thetas, psi, lam):
        """Forward sweep only (hipotf_vqe_energy with no Hamiltonian): psi <- U(theta) psi,
        the fused rotation segments of the VQE evaluation; lam is workspace."""
        idx = _np(idxs, np.int64)
        th = _np(thetas, np.float64)
        rc = L.load().hipotf_vqe_energy(
            self.basis.ptr, None, self.pool.ptr, len(th), idx.ctypes.data_as(L._i64p),
            th.ctypes.data_as(L._dp), _ptr(psi), _ptr(lam), None)
        L.check(rc == 0, "hipotf_vqe_energy")


def sync_device():
    L.call_void("sync_device_cuda")

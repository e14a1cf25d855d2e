"""The C-ABI library loads, exports every symbol include/hipotf.h declares, and its host-only
entry points (basis, network container, errors) behave as the reference's do.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import binsim_amd.lib as L
from binsim_amd import otf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hipotf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_0-9]+)\s*\(", text)) - {"if", "defined"})


def test_every_declared_symbol_is_exported_and_bound():
    lib = L.load()
    syms = declared_symbols()
    assert len(syms) >= 35
    for s in syms:
        assert hasattr(lib, s), s
        assert s in L.SIGNATURES, s
    assert sorted(L.SIGNATURES) == syms


def test_host_basis_matches_reference_semantics():
    b = otf.BasisManager(6, 3, 2)
    assert b.dim == 20 * 15
    lib = L.load()
    assert lib.get_num_symmetry_blocks(b.ptr) == 1
    hf = b.get_hf()
    assert hf[0] == 1.0 and hf.sum() == 1.0
    # set_det_coeff accumulates at ia*Nb + ib for colex-ranked strings (basis.hpp:298-330)
    v = np.zeros(b.dim)
    lib.set_det_coeff_f64(b.ptr, 0b110001, 0b000101, 2.5, v.ctypes.data)
    lib.set_det_coeff_f64(b.ptr, 0b110001, 0b000101, 0.5, v.ctypes.data)
    # rank of 110001 among 3-of-6 colex = C(0,1)+C(4,2)+C(5,3) = 16; of 000101 = C(0,1)+C(2,2) = 1
    assert v[16 * 15 + 1] == 3.0 and v.sum() == 3.0
    # strings outside the basis are ignored
    lib.set_det_coeff_f64(b.ptr, 0b000011, 0b000101, 1.0, v.ctypes.data)
    assert v.sum() == 3.0
    # the complex form (basis/declare.hpp:46-56) on a complex128 array: same index, both parts
    z = np.zeros(b.dim, dtype=np.complex128)
    lib.set_det_coeff_c64(b.ptr, 0b110001, 0b000101, L.C64(1.5, -2.0), z.ctypes.data)
    lib.set_det_coeff_c64(b.ptr, 0b110001, 0b000101, L.C64(0.5, 0.25), z.ctypes.data)
    lib.set_det_coeff_c64(b.ptr, 0b000011, 0b000101, L.C64(9.0, 9.0), z.ctypes.data)
    assert z[16 * 15 + 1] == 2.0 - 1.75j and np.count_nonzero(z) == 1


def test_empty_and_error_paths():
    lib = L.load()
    orb = np.zeros(4, dtype=np.int64)
    p = lib.create_basis_manager(4, 2, 2, 1, orb.ctypes.data_as(L._i64p), 16)
    assert p and lib.get_subspace_dim(p) == 0   # total_sym != 0 -> empty space
    lib.destroy_basis_manager(p)
    orb[1] = 3
    p = lib.create_basis_manager(4, 2, 2, 0, orb.ctypes.data_as(L._i64p), 16)
    assert not p and "single-block" in L.last_error()
    assert lib.get_subspace_dim(None) == -1 and "NULL" in L.last_error()
    lib.destroy_basis_manager(None)  # NULL-safe like the reference destroy_* (cuotf.cu:14-23)
    lib.destroy_basisdev_f64(None)
    lib.destroy_networkdev_f64(None)
    with pytest.raises(L.HipOTFError):
        otf.HipOTF.from_spin_orbitals(4, [[2, 2, 0, 1]])   # repeated creator -> T = 0
    h = np.zeros((3, 3))
    h[0, 1] = 1.0  # not symmetric
    with pytest.raises(L.HipOTFError, match="symmetric"):
        otf.HipOTF.from_integrals(3, h, np.zeros(81))


def test_host_network_container():
    from oracle import oracle as O
    from binsim_amd.integrals import chemist_integrals, one_body_layout, reference_layout
    h, eri = chemist_integrals(4, 3)
    net = O.hamiltonian_network(one_body_layout(h), reference_layout(eri), 4)
    f = net.flat
    lib = L.load()
    orb = np.zeros(4, dtype=np.int64)
    p = lib.build_network_otf_f64(orb.ctypes.data_as(L._i64p), 4, f["ngs"],
                                  f["axs"].ctypes.data_as(L._u32p), f["bxs"].ctypes.data_as(L._u32p),
                                  f["ranks"].ctypes.data_as(L._i64p), f["num_zas"].ctypes.data_as(L._i64p),
                                  f["num_zbs"].ctypes.data_as(L._i64p), f["flat_zas"].ctypes.data_as(L._u32p),
                                  f["flat_zbs"].ctypes.data_as(L._u32p), f["flat_wa"].ctypes.data_as(L._dp),
                                  f["flat_wb"].ctypes.data_as(L._dp))
    assert p
    lib.destroy_network_otf_f64(p)

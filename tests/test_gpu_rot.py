"""GPU parity of the Givens-rotation operators (expm / grad / backgrad / batchgrad) and the
fused VQE energy + adjoint gradient against the oracle (network restatement, jl/vqe.jl:1-21).

Tolerances: states and gradient values within 1e-12 absolute (they are O(1) and O(1e-2));
energies within 1e-12.
"""
import numpy as np
import pytest

from binsim_amd import otf
from binsim_amd.ansatz import ucc_spin_orbitals
from oracle import oracle as O
from tests.conftest import gpu_vec, host
from tests.helpers import integrals, oracle_pool, oracle_system

pytestmark = pytest.mark.gpu

CASES = [(4, 2, 2), (6, 3, 3), (6, 3, 2), (6, 1, 4), (8, 4, 4), (8, 3, 5)]


def _dev(n, na, nb, network_pool=False):
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    H = otf.HipOTF.from_integrals(n, h1, g2, 0.0)
    sos = ucc_spin_orbitals(n, na, nb)
    P = (otf.HipOTF.from_network(n, oracle_pool(n, na, nb).flat) if network_pool
         else otf.HipOTF.from_spin_orbitals(n, sos))
    return B, H, P, otf.HipOTF_Functions(B, ham=H, pool=P), sos


@pytest.mark.parametrize("n,na,nb", CASES)
@pytest.mark.parametrize("network_pool", [False, True])
def test_expm_sweep_and_backgrad(torch_cuda, n, na, nb, network_pool):
    print(f"rot n={n} na={na} nb={nb} net={network_pool}", flush=True)
    S = oracle_system(n, na, nb, pool=True)
    B, H, P, F, sos = _dev(n, na, nb, network_pool)
    th = np.random.default_rng(n + na).uniform(-0.4, 0.4, len(sos))
    v = O.hf_vector(S)
    vd = gpu_vec(torch_cuda, v)
    for i, t in enumerate(th):
        S.expm(i, t, v)
        F.expm(i, t, vd)
    assert np.max(np.abs(host(vd) - v)) < 1e-12
    # grad on a few generators
    r = S.hvec(v)
    rd = gpu_vec(torch_cuda, r)
    for i in range(0, len(sos), max(1, len(sos) // 7)):
        assert abs(F.grad(i, 0.3, vd, rd) - S.grad(i, 0.3, v, r)) < 1e-12
    # reverse sweep
    for i in range(len(sos) - 1, -1, -1):
        g_ref = S.backgrad(i, th[i], v, r)
        g = F.backgrad(i, th[i], vd, rd)
        assert abs(g - g_ref) < 1e-12, i
    assert np.max(np.abs(host(vd) - v)) < 1e-12
    assert np.max(np.abs(host(rd) - r)) < 1e-12


@pytest.mark.parametrize("n,na,nb", CASES + [(12, 6, 6)])
def test_fused_vqe_energy_grad(torch_cuda, n, na, nb):
    """Full sweeps for n <= 8 exercise every fused segment kind (pure-beta rows, alpha-pattern
    row pairs, pure-alpha rows via the beta-major copy); at 12 orbitals the first 60."""
    S = oracle_system(n, na, nb, pool=True)
    B, H, P, F, sos = _dev(n, na, nb)
    npar = len(sos) if n <= 8 else 60
    idxs = np.arange(npar)
    th = np.random.default_rng(7).uniform(-0.05, 0.05, npar)
    lv, rv = O.hf_vector(S), np.zeros(S.dim)
    e_ref, g_ref, d2_ref = O.energy_objective(S, idxs, th, lv, rv)
    psi = gpu_vec(torch_cuda, O.hf_vector(S))
    lam = torch_cuda.zeros_like(psi)
    e, g, d2, nrm = F.vqe_energy_grad(idxs, th, psi, lam)
    assert abs(e - e_ref) < 1e-12
    assert abs(d2 - d2_ref) < 1e-10
    assert np.max(np.abs(g - g_ref)) < 1e-12
    assert np.max(np.abs(host(psi) - lv)) < 1e-12  # reverse sweep uncomputes the state


@pytest.mark.parametrize("sweep_mode", [0, 1, 2, 3, 4, 5, 7])
def test_vqe_config2_matches_reference_fixture(torch_cuda, sweep_mode):
    """12 orbitals, 6+6 electrons, all 1,818 UCC parameters, theta ~ U[-0.05, 0.05]: the fused
    energy + adjoint gradient equal the values the compiled reference produced (golden), for
    each sweep form: fused segments with rows in LDS (0; everything fits at this size), fused
    segments updating the rows in global memory (1), one launch per generator (2), and
    segments with one row and half of the next per block in LDS and the rest in global memory
    (3; the forms 16 orbitals takes: whole, split and global rows); 4, 5, 7 are 0, 1, 3 with
    the adjoint sweep's alpha-pair segments on interleaved (l, r) rows."""
    print(f"vqe fixture sweep_mode={sweep_mode}", flush=True)
    import os
    from binsim_amd import autotune as AT
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config0_n12.npz"))
    B, H, P, F, sos = _dev(12, 6, 6)
    th = f["vqe_thetas"]
    assert len(th) == len(sos) == 1818
    psi = torch_cuda.zeros(B.dim, dtype=torch_cuda.float64, device="cuda")
    psi[0] = 1.0
    lam = torch_cuda.zeros_like(psi)
    saved = AT.get_config(AT.FP64)
    try:
        AT.set_config(AT.FP64, dict(saved, sweep_mode=sweep_mode))
        e, g, d2, nrm = F.vqe_energy_grad(np.arange(len(th)), th, psi, lam)
    finally:
        AT.set_config(AT.FP64, saved)
    assert abs(e - float(f["vqe_energy"])) < 1e-11
    assert abs(d2 - float(f["vqe_d2h"])) < 1e-9
    assert np.max(np.abs(g - f["vqe_grad"])) < 1e-11
    assert abs(float(host(psi)[0]) - 1.0) < 1e-12 and abs(nrm - 1.0) < 1e-12


@pytest.mark.parametrize("n,na,nb,mode", [(6, 3, 3, 0), (8, 3, 5, 0), (8, 3, 5, 2), (12, 6, 6, 0),
                                         (12, 6, 6, 3)])
def test_batchgrad(torch_cuda, n, na, nb, mode):
    """Pool gradient of every generator at one (l, r) pair (batchgrad_cuda, the fused
    gradient-only pass; sweep_mode 2 runs one launch per generator, 3 stages one row and half of
    the next) against the oracle's grad of each generator."""
    from binsim_amd import autotune as AT
    print(f"batchgrad n={n} na={na} nb={nb} sweep_mode={mode}", flush=True)
    S = oracle_system(n, na, nb, pool=True)
    B, H, P, F, sos = _dev(n, na, nb)
    rng = np.random.default_rng(3 + n)
    l, r = rng.uniform(-1, 1, S.dim), rng.uniform(-1, 1, S.dim)
    l /= np.linalg.norm(l)
    r /= np.linalg.norm(r)
    th = rng.uniform(-0.5, 0.5, len(sos))
    gd = torch_cuda.zeros(len(sos), dtype=torch_cuda.float64, device="cuda")
    saved = AT.get_config(AT.FP64)
    try:
        AT.set_config(AT.FP64, dict(saved, sweep_mode=mode))
        F.batchgrad(gpu_vec(torch_cuda, l), gpu_vec(torch_cuda, r), gd, gpu_vec(torch_cuda, th))
    finally:
        AT.set_config(AT.FP64, saved)
    ref = np.array([S.grad(i, th[i], l, r) for i in range(len(sos))])
    assert np.max(np.abs(host(gd) - ref)) < 1e-12

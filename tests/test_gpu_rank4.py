"""GPU parity of the pool operators the reference has on the CPU only -- tvec, backtran, batch expm,
batch tran (tvec_cuda, backtran_cuda, batchexpm_cuda, batchtran_cuda) -- against the oracle
(oracle/oracle.c, pinned against the compiled reference in tests/test_oracle_rank4.py).

Tolerances: tvec, backtran and batch expm do the oracle's per-element arithmetic up to fused
multiply-adds and the oracle's pair coefficient, a product of SVD phase weights that is +-1 only
to rounding (1e-14 absolute on O(1) values); batch tran sums in another order (1e-12 relative
to the largest).
Both pool forms: built from the spin-orbital lists, and decoded from the reference's SVD network.
"""
import numpy as np
import pytest

from binsim_amd import otf
from binsim_amd.ansatz import ucc_spin_orbitals
from tests.conftest import gpu_vec, host
from tests.helpers import oracle_pool, oracle_system

pytestmark = pytest.mark.gpu

CASES = [(4, 2, 2), (6, 3, 3), (6, 1, 4), (8, 3, 5), (10, 5, 4), (12, 6, 6)]


def _dev(n, na, nb, network_pool):
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    sos = ucc_spin_orbitals(n, na, nb)
    P = (otf.HipOTF.from_network(n, oracle_pool(n, na, nb).flat) if network_pool
         else otf.HipOTF.from_spin_orbitals(n, sos))
    return B, P, otf.HipOTF_Functions(B, pool=P), len(sos)


def _gens(ng, k=6):
    return sorted(set(list(range(0, ng, max(1, ng // k))) + [ng - 1]))


@pytest.mark.parametrize("n,na,nb", CASES)
@pytest.mark.parametrize("network_pool", [False, True])
def test_tvec_backtran_batchexpm(torch_cuda, n, na, nb, network_pool):
    print(f"rank4 n={n} na={na} nb={nb} net={network_pool}", flush=True)
    S = oracle_system(n, na, nb, pool=True)
    B, P, F, ng = _dev(n, na, nb, network_pool)
    rng = np.random.default_rng(n + 3 * na + nb)
    x = rng.uniform(-1, 1, S.dim)
    xd = gpu_vec(torch_cuda, x)
    yd = gpu_vec(torch_cuda, np.full(S.dim, 7.0))  # overwritten, zero outside the pairs
    for g in _gens(ng):
        F.tvec(g, xd, yd)
        assert np.max(np.abs(host(yd) - S.tvec(g, x))) < 1e-14, g
    l, r = rng.uniform(-1, 1, S.dim), rng.uniform(-1, 1, S.dim)
    ld, rd = gpu_vec(torch_cuda, l), gpu_vec(torch_cuda, r)
    bd = gpu_vec(torch_cuda, np.full(S.dim, 7.0))
    for g in _gens(ng):
        th = -0.31 + 0.002 * g
        b = S.backtran(g, th, l, r)
        F.backtran(g, th, ld, rd, bd)
        assert np.max(np.abs(host(ld) - l)) < 1e-14 and np.max(np.abs(host(rd) - r)) < 1e-14, g
        assert np.max(np.abs(host(bd) - b)) < 1e-14, g
    for N, j in ((1, 1), (4, 3), (8, 8)):
        m = rng.uniform(-1, 1, (S.dim, N))
        md = gpu_vec(torch_cuda, m.reshape(-1))
        for g in _gens(ng, 3):
            S.batch_expm(g, 0.23, m, N, j)
            F.batchexpm(g, 0.23, md, N, j)
        assert np.max(np.abs(host(md).reshape(S.dim, N) - m)) < 1e-14, (N, j)


@pytest.mark.parametrize("n,na,nb", CASES)
@pytest.mark.parametrize("network_pool", [False, True])
def test_batchtran(torch_cuda, n, na, nb, network_pool):
    S = oracle_system(n, na, nb, pool=True)
    B, P, F, ng = _dev(n, na, nb, network_pool)
    rng = np.random.default_rng(5 + n)
    l, r = rng.uniform(-1, 1, S.dim), rng.uniform(-1, 1, S.dim)
    ref = S.batch_tran(ng, l, r)
    td = gpu_vec(torch_cuda, np.full(ng, 7.0))
    F.batchtran(gpu_vec(torch_cuda, l), gpu_vec(torch_cuda, r), td)
    assert np.max(np.abs(host(td) - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_rank4_argument_errors(torch_cuda):
    import binsim_amd.lib as L
    B, P, F, ng = _dev(6, 3, 3, False)
    x = gpu_vec(torch_cuda, np.ones(B.dim))
    with pytest.raises(L.HipOTFError, match="different"):
        F.tvec(0, x, x)
    with pytest.raises(L.HipOTFError, match="out of range"):
        F.tvec(ng, x, gpu_vec(torch_cuda, np.zeros(B.dim)))
    with pytest.raises(L.HipOTFError, match="nvec"):
        F.batchexpm(0, 0.1, x, 1, 2)

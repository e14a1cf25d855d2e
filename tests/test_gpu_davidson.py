"""GPU parity of the Davidson layer: the eight libdiag entry points against the oracle's
restatement (src/include/diag/davidson.hpp), the davidson.jl mirror against the oracle's
davidson restatement, and the device driver (plain FP64 and mixed-precision block) against
dense diagonalisation and the reference-driven config-0 energy.

Tolerances: vector ops 1e-12 relative; energies 1e-10 Ha (residual tol 1e-8 or tighter);
mixed vs plain FP64 |dE| < 1e-9 Ha.
"""
import ctypes as C
import os

import numpy as np
import pytest

import binsim_amd.lib as L
from binsim_amd import davidson as D
from binsim_amd import otf
from oracle import oracle as O
from tests.conftest import gpu_vec, host
from tests.helpers import integrals, oracle_system, rel_err

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pp(vs):
    return C.cast((C.c_void_p * len(vs))(*[int(v.data_ptr()) for v in vs]), L._dpp)


def _opp(vs):
    arr = (O._dp * len(vs))(*[v.ctypes.data_as(O._dp) for v in vs])
    return arr


def test_libdiag_ops_match_oracle(torch_cuda):
    lib, ol = L.load(), O.lib()
    rng = np.random.default_rng(0)
    N, dim = 100_003, 7
    Vh = [rng.uniform(-1, 1, N) for _ in range(dim)]
    AVh = [rng.uniform(-1, 1, N) for _ in range(dim)]
    Vd = [gpu_vec(torch_cuda, v) for v in Vh]
    AVd = [gpu_vec(torch_cuda, v) for v in AVh]
    c = rng.uniform(-1, 1, dim)
    e = -1.2345
    out_h = np.empty(N)
    out_d = torch_cuda.empty(N, dtype=torch_cuda.float64, device="cuda")
    r_h = ol.oracle_ritz_and_residual(N, dim, c.ctypes.data_as(O._dp), e, _opp(Vh), _opp(AVh),
                                      out_h.ctypes.data_as(O._dp))
    r_d = lib.build_ritz_and_residual_f64(N, dim, c.ctypes.data_as(L._dp), e, _pp(Vd), _pp(AVd),
                                          out_d.data_ptr())
    assert abs(r_h - r_d) < 1e-12 * r_h and rel_err(host(out_d), out_h) < 1e-13
    ol.oracle_ritz_vector(N, dim, c.ctypes.data_as(O._dp), _opp(Vh), out_h.ctypes.data_as(O._dp))
    lib.build_ritz_vector_f64(N, dim, c.ctypes.data_as(L._dp), _pp(Vd), out_d.data_ptr())
    assert rel_err(host(out_d), out_h) < 1e-13
    diags = rng.uniform(-3, 3, N)
    diags[:50] = e  # exercises the shift clamp
    t_h = Vh[0].copy()
    t_d = gpu_vec(torch_cuda, t_h)
    ol.oracle_precondition(N, t_h.ctypes.data_as(O._dp), diags.ctypes.data_as(O._dp), e, 1e-3)
    lib.apply_preconditioner_inplace_f64(N, t_d.data_ptr(), gpu_vec(torch_cuda, diags).data_ptr(), e, 1e-3)
    assert rel_err(host(t_d), t_h) < 1e-14
    m_h = AVh[1].copy()
    m_d = gpu_vec(torch_cuda, m_h)
    n_h = ol.oracle_mgs(N, dim, _opp(Vh), m_h.ctypes.data_as(O._dp))
    n_d = lib.mgs_orthogonalize_f64(N, dim, _pp(Vd), m_d.data_ptr())
    assert abs(n_h - n_d) < 1e-12 * n_h and rel_err(host(m_d), m_h) < 1e-12
    assert abs(lib.inplace_normalize_f64(N, m_d.data_ptr()) - n_d) < 1e-12 * n_d
    assert abs(np.linalg.norm(host(m_d)) - 1.0) < 1e-14
    assert abs(lib.fast_dot_f64(N, Vd[0].data_ptr(), AVd[0].data_ptr()) - Vh[0] @ AVh[0]) < 1e-10
    col = np.empty(dim)
    lib.compute_heff_col_f64(N, dim, _pp(Vd), AVd[3].data_ptr(), col.ctypes.data_as(L._dp))
    assert np.max(np.abs(col - np.array([v @ AVh[3] for v in Vh]))) < 1e-10
    ld = 9
    heff = np.zeros(ld * ld)
    lib.rebuild_heff_f64(N, dim, _pp(Vd), _pp(AVd), heff.ctypes.data_as(L._dp), ld)
    want = np.array([[Vh[r] @ AVh[cc] for cc in range(dim)] for r in range(dim)])
    got = heff.reshape(ld, ld).T[:dim, :dim]
    assert np.max(np.abs(got - want)) < 1e-10


@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (8, 4, 4), (8, 3, 5)])
def test_davidson_mirror_matches_oracle(torch_cuda, n, na, nb):
    """binsim_amd.davidson (the davidson.jl loop over the device libdiag) == oracle.davidson."""
    S = oracle_system(n, na, nb)
    hf = O.hf_vector(S)
    e_ref, x_ref, it_ref = O.davidson(S.hvec, hf, S.diags(), tol=1e-8)
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    F = otf.HipOTF_Functions(B, ham=otf.HipOTF.from_integrals(n, h1, g2))
    d = torch_cuda.zeros(S.dim, dtype=torch_cuda.float64, device="cuda")
    F.get_diags(d)
    e, x, it = D.davidson(F.hvec, gpu_vec(torch_cuda, hf), d, tol=1e-8)
    assert abs(e - e_ref) < 1e-10
    assert abs(it - it_ref) <= 1
    assert abs(abs(float(host(x) @ x_ref)) - 1.0) < 1e-9


def test_driver_config0_energy_matches_reference(torch_cuda):
    """Device driver, one root from HF, tol 1e-8 vs the reference-driven value (fixture)."""
    f = np.load(os.path.join(GOLDEN, "config0_n12.npz"))
    n = 12
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, 6, 6))
    H = otf.HipOTF.from_integrals(n, h1, g2)
    hf = torch_cuda.zeros(B.dim, dtype=torch_cuda.float64, device="cuda")
    hf[0] = 1.0
    e, _, info = D.solve(B, H, nroots=1, mixed=False, tol=1e-8, guess=hf)
    assert info["converged"] == 1
    assert abs(e[0] - float(f["e0"])) < 1e-10, (e[0], float(f["e0"]))
    # From HF, FP64 arithmetic keeps the alpha<->beta exchange parity of the start (as the
    # reference does); FP32 rounding does not, so the mixed solver is compared with the plain
    # one from a start that already mixes both parities.
    g = torch_cuda.from_numpy(np.random.default_rng(4).uniform(-1, 1, B.dim)).to("cuda")
    e64, _, i64 = D.solve(B, H, nroots=1, mixed=False, tol=1e-8, guess=g)
    e32, _, i32 = D.solve(B, H, nroots=1, mixed=True, tol=1e-8, guess=g)
    assert i64["converged"] == 1 and i32["converged"] == 1
    assert e64[0] <= float(f["e0"]) + 1e-10
    assert abs(e64[0] - e32[0]) < 1e-9
    assert i32["sigma32"] > 0


@pytest.mark.parametrize("n,na,nb,k", [(6, 3, 3, 3), (7, 3, 3, 4), (8, 3, 5, 4)])
def test_driver_multiroot_plain_and_mixed(torch_cuda, n, na, nb, k):
    S = oracle_system(n, na, nb)
    H = np.zeros((S.dim, S.dim))
    for j in range(S.dim):
        e = np.zeros(S.dim)
        e[j] = 1.0
        H[:, j] = S.hvec(e)
    w = np.linalg.eigvalsh(H)
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    Hd = otf.HipOTF.from_integrals(n, h1, g2)
    e0, v0, i0 = D.solve(B, Hd, nroots=k, mixed=False, tol=1e-8, want_vectors=True)
    e1, v1, i1 = D.solve(B, Hd, nroots=k, mixed=True, tol=1e-8, want_vectors=True)
    assert i0["converged"] == 1 and i1["converged"] == 1
    # both variants find exactly the k lowest eigenvalues of H (no root skipped), and agree
    assert np.allclose(e0, w[:k], rtol=0, atol=1e-9), (e0, w[:k])
    assert np.allclose(e1, w[:k], rtol=0, atol=1e-9), (e1, w[:k])
    assert np.max(np.abs(e0 - e1)) < 1e-9
    # and their vectors are eigenvectors to the residual tolerance
    for ee, vv in ((e0, v0), (e1, v1)):
        X = host(vv).reshape(k, S.dim)
        for i in range(k):
            assert np.linalg.norm(H @ X[i] - ee[i] * X[i]) < 1e-7
            assert abs(np.linalg.norm(X[i]) - 1.0) < 1e-10


def _dense(S):
    H = np.zeros((S.dim, S.dim))
    for j in range(S.dim):
        e = np.zeros(S.dim)
        e[j] = 1.0
        H[:, j] = S.hvec(e)
    return H


@pytest.mark.parametrize("mixed", [False, True], ids=["plain", "mixed"])
def test_driver_maxiter_returns_latest_ritz_pairs(torch_cuda, mixed):
    """maxiter reached (status 2): energies and vectors are the latest Ritz pairs."""
    n, na, nb, k = 8, 4, 4, 2
    S = oracle_system(n, na, nb)
    H = _dense(S)
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    Hd = otf.HipOTF.from_integrals(n, h1, g2)
    e, v, info = D.solve(B, Hd, nroots=k, mixed=mixed, tol=1e-12, maxiter=3, want_vectors=True)
    assert info["converged"] == 0 and info["iterations"] == 3
    X = host(v).reshape(k, S.dim)
    # the mixed solver's estimates after 3 iterations come from its FP32 sigma (Rayleigh-Ritz
    # with the FP32 images of the subspace), so they carry FP32 rounding
    tol = 1e-6 * abs(e[0]) if mixed else 1e-10
    for i in range(k):
        assert abs(np.linalg.norm(X[i]) - 1.0) < 1e-10
        assert abs(X[i] @ H @ X[i] - e[i]) < tol
    assert e[0] >= np.linalg.eigvalsh(H)[0] - tol


def test_driver_breakdown_status(torch_cuda):
    """No new direction while the residual is above tol (status 3, info converged = 2): started
    from an exact eigenvector with tol below rounding, the first correction lies in the subspace."""
    n, na, nb = 6, 3, 3
    S = oracle_system(n, na, nb)
    H = _dense(S)
    w, U = np.linalg.eigh(H)
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    Hd = otf.HipOTF.from_integrals(n, h1, g2)
    lib = L.load()
    e = np.zeros(1)
    info = np.zeros(7)
    vecs = torch_cuda.empty(S.dim, dtype=torch_cuda.float64, device="cuda")
    g = gpu_vec(torch_cuda, U[:, 0])
    rc = lib.hipotf_davidson_f64(B.ptr, Hd.ptr, 1, 0, 0, 1e-30, 1e-5, 50, 0, 0, int(g.data_ptr()),
                                 e.ctypes.data_as(L._dp), int(vecs.data_ptr()), info.ctypes.data_as(L._dp))
    assert rc == 3 and info[6] == 2 and info[0] == 1
    assert abs(e[0] - w[0]) < 1e-12
    assert abs(abs(float(host(vecs) @ U[:, 0])) - 1.0) < 1e-12

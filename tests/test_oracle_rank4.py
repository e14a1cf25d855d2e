"""Pins of the oracle's tvec / backtran / batch expm / batch tran (oracle/oracle.c, restating
op/tvec.hpp:373-445, op/backtran.hpp:383-452, op/batchexpm.hpp:376-444, op/batchtran.hpp:517-528)
against identities that do not use the oracle's own formulas, and against the compiled reference
(oracle/_ref) where it is built.

* tvec is the generator itself: the central difference of expm at theta = 0, and antisymmetric.
* backtran rotates l and r exactly as expm does and returns tvec of the rotated l.
* batch tran is the generator's gradient at theta = 0 and equals -<l, G_k r>.
* batch expm on a det-major matrix equals expm on each of its first nvec columns and leaves
  the other columns alone.
"""
import numpy as np
import pytest

from binsim_amd.ansatz import ucc_spin_orbitals
from oracle import oracle as O
from oracle import ref as REF

SHAPES = [(4, 2, 2), (6, 3, 3), (6, 1, 4), (7, 4, 2)]


def _system(n, na, nb):
    sos = ucc_spin_orbitals(n, na, nb)
    return O.System(n, na, nb, pool=O.pool_network(sos)), len(sos)


def _vec(dim, seed):
    return np.random.default_rng(seed).uniform(-1, 1, dim)


@pytest.mark.parametrize("n,na,nb", SHAPES)
def test_tvec_is_the_generator(n, na, nb):
    S, ng = _system(n, na, nb)
    x, y = _vec(S.dim, 1), _vec(S.dim, 2)
    h = 1e-5
    for g in range(0, ng, max(1, ng // 7)):
        gx = S.tvec(g, x)
        vp, vm = x.copy(), x.copy()
        S.expm(g, h, vp)
        S.expm(g, -h, vm)
        assert np.max(np.abs((vp - vm) / (2 * h) - gx)) < 1e-8
        assert abs(y @ gx + S.tvec(g, y) @ x) < 1e-12 * S.dim  # <y|G x> = -<G y|x>
        assert np.any(gx != 0)


@pytest.mark.parametrize("n,na,nb", SHAPES)
def test_backtran_rotates_and_returns_tvec(n, na, nb):
    S, ng = _system(n, na, nb)
    for g in range(0, ng, max(1, ng // 5)):
        th = 0.3 + 0.01 * g
        l, r = _vec(S.dim, 3 + g), _vec(S.dim, 4 + g)
        el, er = l.copy(), r.copy()
        S.expm(g, th, el)
        S.expm(g, th, er)
        b = S.backtran(g, th, l, r)
        assert np.array_equal(l, el) and np.array_equal(r, er)
        assert np.max(np.abs(b - S.tvec(g, el))) < 1e-15


@pytest.mark.parametrize("n,na,nb", SHAPES)
def test_batch_tran_is_grad_at_zero(n, na, nb):
    S, ng = _system(n, na, nb)
    l, r = _vec(S.dim, 5), _vec(S.dim, 6)
    t = S.batch_tran(ng, l, r)
    g0 = np.array([S.grad(k, 0.0, l, r) for k in range(ng)])
    assert np.max(np.abs(t - g0)) < 1e-13
    gt = np.array([-(l @ S.tvec(k, r)) for k in range(ng)])
    assert np.max(np.abs(t - gt)) < 1e-12


@pytest.mark.parametrize("n,na,nb", SHAPES)
@pytest.mark.parametrize("ld,nvec", [(1, 1), (3, 2), (4, 4)])
def test_batch_expm_matches_single_vectors(n, na, nb, ld, nvec):
    S, ng = _system(n, na, nb)
    m = np.random.default_rng(7).uniform(-1, 1, (S.dim, ld))
    m0 = m.copy()
    g, th = ng // 2, 0.41
    S.batch_expm(g, th, m, ld, nvec)
    for k in range(ld):
        v = m0[:, k].copy()
        if k < nvec:
            S.expm(g, th, v)
        assert np.max(np.abs(m[:, k] - v)) < 1e-15


needs_ref = pytest.mark.skipif(not REF.available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("n,na,nb", SHAPES)
def test_rank4_ops_match_reference(n, na, nb):
    R = REF.Reference()
    sos = ucc_spin_orbitals(n, na, nb)
    pn = O.pool_network(sos)
    S = O.System(n, na, nb, pool=pn)
    bp, dim = R.basis(n, na, nb)
    rp = R.network(pn.flat, n)
    try:
        ng = len(sos)
        x, l, r = _vec(dim, 11), _vec(dim, 12), _vec(dim, 13)
        for g in range(0, ng, max(1, ng // 6)):
            d = np.full(dim, 9.0)
            R.tvec(bp, rp, g, x, d)
            assert np.max(np.abs(S.tvec(g, x) - d)) < 1e-14
            l1, r1, l2, r2 = l.copy(), r.copy(), l.copy(), r.copy()
            b2 = np.full(dim, 9.0)
            b1 = S.backtran(g, -0.37, l1, r1)
            R.backtran(bp, rp, g, -0.37, l2, r2, b2)
            assert max(np.max(np.abs(l1 - l2)), np.max(np.abs(r1 - r2)), np.max(np.abs(b1 - b2))) < 1e-14
            m1 = np.random.default_rng(g).uniform(-1, 1, (dim, 3))
            m2 = m1.copy()
            S.batch_expm(g, 0.29, m1, 3, 2)
            R.batch_expm(bp, rp, g, 0.29, m2, 3, 2)
            assert np.max(np.abs(m1 - m2)) < 1e-14
        t2 = np.zeros(ng)
        R.batch_tran(bp, rp, l, r, t2)
        assert np.max(np.abs(S.batch_tran(ng, l, r) - t2)) < 1e-12
    finally:
        R.free_network(rp)
        R.free_basis(bp)

"""CPU pins of the oracle (the checker every GPU parity test compares against).

Parity status of each oracle function and what pins it:

* jw_hamiltonian_terms (ham.hpp restatement): equal term set to the compiled reference
  (oracle/_ref) and, independently, the network it produces equals brute-force second
  quantisation (Slater-Condon with interleaved Jordan-Wigner signs) on small systems.
* oracle_hvec / oracle_diags: golden vectors from the compiled reference (tests/golden),
  the dense brute-force H above, the closed-shell HF energy, Hermiticity.
* oracle_expm / grad / backgrad: golden vectors; scipy.linalg.expm of the dense generator;
  norm preservation and exp(tG)exp(-tG) = 1; central finite differences of energy_objective.
* ucc order (glue.ci_spin_orbitals and binsim_amd.ansatz): counts 72/225/225/1296 and
  128/1568/4096 and the explicit head of the list read from jl/ansatz.jl:67-167.
* davidson (jl/davidson.jl restatement): lowest eigenvalue of the dense H (numpy eigh) and
  the reference-driven value in the config-0 fixture.
* Planted mistakes (a flipped sign in the oracle's JW generation, a transposed integral
  layout) are caught by the dense-H pin.
"""
import os
from itertools import combinations

import numpy as np
import pytest
import scipy.linalg

from binsim_amd.ansatz import ucc_spin_orbitals
from binsim_amd.integrals import chemist_integrals, one_body_layout, reference_layout
from oracle import glue
from oracle import oracle as O
from oracle import ref as REF
from tests.helpers import rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------------------------------
# independent brute force (second quantisation in the interleaved Jordan-Wigner order)
# ---------------------------------------------------------------------------------------
def _strings(n, k):
    return sorted(sum(1 << i for i in c) for c in combinations(range(n), k))


def _apply(ops, s):
    sign = 1
    for pos, cre in reversed(ops):
        occ = (s >> pos) & 1
        if cre == occ:
            return 0, 0
        if bin(s & ((1 << pos) - 1)).count("1") & 1:
            sign = -sign
        s ^= 1 << pos
    return sign, s


def _index(n, na, nb):
    A, B = _strings(n, na), _strings(n, nb)
    return {glue.unzip_even(a) | glue.unzip_odd(b): i * len(B) + j
            for i, a in enumerate(A) for j, b in enumerate(B)}


def dense_h(n, na, nb, h, eri):
    idx = _index(n, na, nb)
    H = np.zeros((len(idx), len(idx)))
    for s, j in idx.items():
        for p in range(n):
            for q in range(n):
                for sg in (0, 1):
                    sgn, t = _apply([(2 * p + sg, 1), (2 * q + sg, 0)], s)
                    if sgn and t in idx:
                        H[idx[t], j] += sgn * h[p, q]
        for p in range(n):
            for q in range(n):
                for r in range(n):
                    for ss in range(n):
                        g = eri[p, ss, q, r]
                        for s1 in (0, 1):
                            for s2 in (0, 1):
                                sgn, t = _apply([(2 * p + s1, 1), (2 * q + s2, 1), (2 * r + s2, 0),
                                                 (2 * ss + s1, 0)], s)
                                if sgn and t in idx:
                                    H[idx[t], j] += 0.5 * sgn * g
    return H


def dense_generator(n, na, nb, spin_orbital):
    idx = _index(n, na, nb)
    mid = len(spin_orbital) // 2
    ops = [(o, 1) for o in spin_orbital[:mid]] + [(o, 0) for o in spin_orbital[mid:]]
    T = np.zeros((len(idx), len(idx)))
    for s, j in idx.items():
        sgn, t = _apply(ops, s)
        if sgn and t in idx:
            T[idx[t], j] += sgn
    return T - T.T


def _matrix(apply, dim):
    M = np.zeros((dim, dim))
    for j in range(dim):
        e = np.zeros(dim)
        e[j] = 1.0
        M[:, j] = apply(e)
    return M


BRUTE = [(4, 2, 2), (4, 1, 2), (5, 2, 2), (5, 3, 2), (4, 0, 3)]


@pytest.mark.parametrize("n,na,nb", BRUTE)
def test_oracle_hamiltonian_equals_second_quantisation(n, na, nb):
    h, eri = chemist_integrals(n, 31 + n)
    net = O.hamiltonian_network(one_body_layout(h), reference_layout(eri), n)
    S = O.System(n, na, nb, ham=net)
    H = _matrix(S.hvec, S.dim)
    Hd = dense_h(n, na, nb, h, eri)
    assert np.max(np.abs(H - Hd)) < 1e-12
    assert np.max(np.abs(S.diags() - np.diag(Hd))) < 1e-12


def test_planted_mistakes_are_caught():
    """A sign flipped in the oracle's JW generation, or the integral layout transposed,
    must fail the dense-H pin (the pin is not a retyping of the oracle's formula)."""
    n, na, nb = 4, 2, 2
    h, eri = chemist_integrals(n, 35)
    Hd = dense_h(n, na, nb, h, eri)
    axs, bxs, azs, bzs, cs = glue.jw_hamiltonian_terms(one_body_layout(h), reference_layout(eri), n)
    cs_bad = cs.copy()
    cs_bad[np.argmax(np.abs(cs) * (axs > 0))] *= -1.0
    bad = O.Network(glue.compress_by_svd((axs, bxs, azs, bzs, cs_bad)))
    Sb = O.System(n, na, nb, ham=bad)
    assert np.max(np.abs(_matrix(Sb.hvec, Sb.dim) - Hd)) > 1e-3
    wrong_layout = np.ascontiguousarray(np.transpose(eri, (0, 1, 2, 3)).reshape(-1))  # (pq|rs) flat
    bad2 = O.hamiltonian_network(one_body_layout(h), wrong_layout, n)
    S2 = O.System(n, na, nb, ham=bad2)
    assert np.max(np.abs(_matrix(S2.hvec, S2.dim) - Hd)) > 1e-3


@pytest.mark.parametrize("n,k", [(6, 3), (8, 4), (12, 6)])
def test_hf_energy_closed_form(n, k):
    h, eri = chemist_integrals(n, 77 + n)
    net = O.hamiltonian_network(one_body_layout(h), reference_layout(eri), n)
    if n == 12:
        S = O.System(n, k, k, ham=net)
    else:
        S = O.System(n, k, k, ham=net)
    d0 = S.diags()[0]
    occ = range(k)
    ehf = 2 * sum(h[i, i] for i in occ) + sum(2 * eri[i, i, j, j] - eri[i, j, j, i]
                                              for i in occ for j in occ)
    assert abs(d0 - ehf) < 1e-11


@pytest.mark.parametrize("n,na,nb", [(4, 2, 2), (5, 2, 3), (4, 1, 3)])
def test_oracle_expm_equals_matrix_exponential(n, na, nb):
    sos = ucc_spin_orbitals(n, na, nb)
    S = O.System(n, na, nb, pool=O.pool_network(sos))
    rng = np.random.default_rng(5)
    for g, so in enumerate(sos):
        G = dense_generator(n, na, nb, so)
        v = rng.uniform(-1, 1, S.dim)
        want = scipy.linalg.expm(0.37 * G) @ v
        got = v.copy()
        S.expm(g, 0.37, got)
        assert np.max(np.abs(got - want)) < 1e-13, so
        # inverse rotation restores the vector; the norm is preserved
        back = got.copy()
        S.expm(g, -0.37, back)
        assert np.max(np.abs(back - v)) < 1e-13
        assert abs(np.linalg.norm(got) - np.linalg.norm(v)) < 1e-13
        # grad(theta, l, r) = <r| G e^{theta G} |l>  (src/include/core/math.hpp:148-162)
        l, r = rng.uniform(-1, 1, S.dim), rng.uniform(-1, 1, S.dim)
        want_g = r @ (G @ scipy.linalg.expm(0.21 * G) @ l)
        assert abs(S.grad(g, 0.21, l, r) - want_g) < 1e-12


def test_energy_objective_matches_finite_differences():
    n, na, nb = 6, 3, 3
    h, eri = chemist_integrals(n, 99)
    sos = ucc_spin_orbitals(n, na, nb)
    S = O.System(n, na, nb, ham=O.hamiltonian_network(one_body_layout(h), reference_layout(eri), n),
                 pool=O.pool_network(sos))
    idxs = np.arange(len(sos))
    th = np.random.default_rng(2).uniform(-0.05, 0.05, len(sos))

    def energy(x):
        lv, rv = O.hf_vector(S), np.zeros(S.dim)
        e, g, _ = O.energy_objective(S, idxs, x, lv, rv)
        return e, g

    e0, g0 = energy(th)
    hstep = 1e-6
    for k in list(range(0, len(sos), 9))[:14]:
        xp, xm = th.copy(), th.copy()
        xp[k] += hstep
        xm[k] -= hstep
        fd = (energy(xp)[0] - energy(xm)[0]) / (2 * hstep)
        assert abs(fd - g0[k]) < 1e-8, k


def test_ucc_order_and_counts():
    ops12 = ucc_spin_orbitals(12, 6, 6)
    assert ops12 == glue.ci_spin_orbitals(12, 6, 6)
    kinds = {"a": 0, "b": 0, "aa": 0, "bb": 0, "ab": 0}
    for so in ops12:
        na_ = sum(1 for o in so if o % 2 == 0)
        nb_ = len(so) - na_
        key = {(2, 0): "a", (0, 2): "b", (4, 0): "aa", (0, 4): "bb", (2, 2): "ab"}[(na_, nb_)]
        kinds[key] += 1
    assert kinds == {"a": 36, "b": 36, "aa": 225, "bb": 225, "ab": 1296}
    assert len(ops12) == 1818
    ops16 = ucc_spin_orbitals(16, 8, 8)
    assert len(ops16) == 5792
    # head of the list: beta singles first (ka = 0, kb = 1), occ outer / vir inner
    assert ops12[0] == [13, 1] and ops12[1] == [15, 1] and ops12[6] == [13, 3]
    # first alpha single (occ 0 -> vir 6) follows all 36 + 225 pure-beta entries
    assert ops12[261] == [12, 0] and ops12[262] == [12, 13, 0, 1]


@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (6, 2, 4), (7, 3, 3)])
def test_oracle_davidson_lowest_eigenvalue(n, na, nb):
    h, eri = chemist_integrals(n, 55 + n)
    net = O.hamiltonian_network(one_body_layout(h), reference_layout(eri), n)
    S = O.System(n, na, nb, ham=net)
    H = _matrix(S.hvec, S.dim)
    w, U = np.linalg.eigh(H)
    # from a random start: the global ground state
    v0 = np.random.default_rng(n).uniform(-1, 1, S.dim)
    e, x, it = O.davidson(S.hvec, v0, S.diags(), tol=1e-9)
    assert abs(e - w[0]) < 1e-10
    assert rel_err(H @ x, e * x) < 1e-8
    # from HF (index 0, as run_fci does): H and the diagonal preconditioner keep the
    # alpha<->beta exchange parity of the start, so the result is the lowest eigenvalue
    # whose eigenvector overlaps that sector (HF itself suffices to identify it)
    hf = np.zeros(S.dim)
    hf[0] = 1.0
    e_hf, x_hf, _ = O.davidson(S.hvec, hf, S.diags(), tol=1e-9)
    reach = np.abs(U[0, :]) > 1e-8
    assert np.min(np.abs(w - e_hf)) < 1e-10
    assert e_hf <= w[reach][0] + 1e-10


# ---------------------------------------------------------------------------------------
# golden vectors written from the compiled reference
# ---------------------------------------------------------------------------------------
SMALL_GOLDEN = [(6, 3, 3), (6, 3, 2), (8, 4, 4), (8, 3, 5)]


@pytest.mark.parametrize("n,na,nb", SMALL_GOLDEN)
def test_oracle_matches_golden(n, na, nb):
    f = np.load(os.path.join(GOLDEN, f"small_n{n}_{na}{nb}.npz"))
    h, eri = chemist_integrals(n, int(f["seed"]))
    sos = ucc_spin_orbitals(n, na, nb)
    S = O.System(n, na, nb, ham=O.hamiltonian_network(one_body_layout(h), reference_layout(eri), n),
                 pool=O.pool_network(sos))
    x = np.random.default_rng(int(f["x_seed"])).uniform(-1.0, 1.0, S.dim)
    x /= np.linalg.norm(x)
    assert rel_err(S.hvec(x), f["sigma"]) < 1e-13
    assert np.max(np.abs(S.diags() - f["diag"])) < 1e-12
    v = O.hf_vector(S)
    for i, t in enumerate(f["thetas"]):
        S.expm(i, t, v)
    assert np.max(np.abs(v - f["state"])) < 1e-13
    r = S.hvec(v)
    g = np.array([S.grad(i, 0.3, v, r) for i in range(len(sos))])
    assert np.max(np.abs(g - f["grads_at_0p3"])) < 1e-13
    back = np.zeros(len(sos))
    for i in range(len(sos) - 1, -1, -1):
        back[i] = S.backgrad(i, f["thetas"][i], v, r)
    assert np.max(np.abs(back - f["backgrad"])) < 1e-13


@pytest.mark.slow
def test_oracle_matches_golden_config0():
    from binsim_amd.integrals import for_norb
    f = np.load(os.path.join(GOLDEN, "config0_n12.npz"))
    h, eri = for_norb(12)
    S = O.System(12, 6, 6, ham=O.hamiltonian_network(one_body_layout(h), reference_layout(eri), 12))
    x = np.random.default_rng(1012).uniform(-1.0, 1.0, S.dim)
    x /= np.linalg.norm(x)
    sg = S.hvec(x)
    y = np.random.default_rng(4242).uniform(-1.0, 1.0, S.dim)
    y /= np.linalg.norm(y)
    assert abs(sg @ sg - f["sigma_norm2"]) < 1e-12 * f["sigma_norm2"]
    assert abs(y @ sg - f["sigma_dot_y"]) < 1e-12 * np.sqrt(f["sigma_norm2"])
    assert np.max(np.abs(sg[f["sample_idx"]] - f["sample_val"])) < 1e-12 * np.max(np.abs(sg))
    d = S.diags()
    assert np.max(np.abs(d[:256] - f["diag_head"])) < 1e-12


# ---------------------------------------------------------------------------------------
# the restatement against the compiled reference itself (when oracle/_ref is built)
# ---------------------------------------------------------------------------------------
needs_ref = pytest.mark.skipif(not REF.available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("n", [4, 6, 8])
def test_jw_terms_match_reference(n):
    R = REF.Reference()
    h, eri = chemist_integrals(n, 7 + n)
    tr = R.hamiltonian_terms(one_body_layout(h), reference_layout(eri), n)
    tg = glue.jw_hamiltonian_terms(one_body_layout(h), reference_layout(eri), n)
    for a, b in zip(tr[:4], tg[:4]):
        assert np.array_equal(a, b)
    assert np.max(np.abs(tr[4] - tg[4])) < 1e-13


@needs_ref
@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (6, 1, 4), (7, 4, 2)])
def test_oracle_matches_reference(n, na, nb):
    R = REF.Reference()
    h, eri = chemist_integrals(n, 3 + n)
    h1, g2 = one_body_layout(h), reference_layout(eri)
    sos = ucc_spin_orbitals(n, na, nb)
    hn = O.hamiltonian_network(h1, g2, n)
    pn = O.pool_network(sos)
    S = O.System(n, na, nb, ham=hn, pool=pn)
    bp, dim = R.basis(n, na, nb)
    rh, rp = R.network(hn.flat, n), R.network(pn.flat, n)
    x = np.random.default_rng(1).uniform(-1, 1, dim)
    assert rel_err(S.hvec(x), R.hvec(bp, rh, x)) < 1e-13
    assert np.max(np.abs(S.diags() - R.diags(bp, rh, dim))) < 1e-12
    th = np.random.default_rng(2).uniform(-0.5, 0.5, len(sos))
    v1 = np.zeros(dim)
    v1[0] = 1
    v2 = v1.copy()
    for i, t in enumerate(th):
        S.expm(i, t, v1)
        R.expm(bp, rp, i, t, v2)
    assert np.max(np.abs(v1 - v2)) < 1e-13
    r1, r2 = S.hvec(v1), R.hvec(bp, rh, v2)
    for i in range(len(sos) - 1, -1, -1):
        assert abs(S.backgrad(i, th[i], v1, r1) - R.backgrad(bp, rp, i, th[i], v2, r2)) < 1e-13
    R.free_network(rh)
    R.free_network(rp)
    R.free_basis(bp)


@needs_ref
def test_reference_driven_davidson_and_objective():
    """The reference-library drivers (fixtures and the CPU baseline use them) agree with the
    restatement: the same Davidson iterations and energy from HF, and the same VQE objective."""
    n, na, nb = 6, 3, 3
    R = REF.Reference()
    h, eri = chemist_integrals(n, 55 + n)
    h1, g2 = one_body_layout(h), reference_layout(eri)
    sos = ucc_spin_orbitals(n, na, nb)
    hn, pn = O.hamiltonian_network(h1, g2, n), O.pool_network(sos)
    S = O.System(n, na, nb, ham=hn, pool=pn)
    bp, dim = R.basis(n, na, nb)
    rh, rp = R.network(hn.flat, n), R.network(pn.flat, n)
    hf = np.zeros(dim)
    hf[0] = 1.0
    e_o, _, it_o = O.davidson(S.hvec, hf, S.diags(), tol=1e-9)
    e_r, it_r, conv = R.davidson(lambda v: R.hvec(bp, rh, v), hf, R.diags(bp, rh, dim), 1e-9)
    assert conv and it_r == it_o and abs(e_r - e_o) < 1e-11
    # a bounded run stops after maxiter and says so
    assert R.davidson(lambda v: R.hvec(bp, rh, v), hf, R.diags(bp, rh, dim), 1e-9, maxiter=2)[1:] == (2, False)
    th = np.random.default_rng(3).uniform(-0.3, 0.3, len(sos))
    lv, rv = hf.copy(), np.empty(dim)
    e1, g1, d1 = O.energy_objective(S, list(range(len(sos))), th, lv, rv)
    e2, g2_, d2 = R.energy_objective(bp, rh, rp, th, dim)
    assert abs(e1 - e2) < 1e-12 and abs(d1 - d2) < 1e-11
    assert np.max(np.abs(g1 - g2_)) < 1e-12
    R.free_network(rh)
    R.free_network(rp)
    R.free_basis(bp)

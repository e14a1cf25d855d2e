"""Pins of the oracle's spin-model and constraint builders (oracle/glue.py, restating
jl/hamiltonian.jl:404-484 and binqubitaabb.jl:436-463) against matrices built here from scratch.

* A qubit term (x, z, c) is c * prod_q X_q^{x_q} Z_q^{z_q} (X left of Z on each qubit, the
  product order the reference's multiply assumes, binqubitaabb.jl:313-374); the dense matrix of
  a term list is built qubit by qubit with numpy kron, independently of the glue's algebra.
* Heisenberg / Ising chains equal the same chains assembled from explicit 2x2 Pauli matrices
  (Y = [[0, -i], [i, 0]]): this fixes the sign of the real encoding of Y Y.
* N, S^2, S_z: N and S_z are diagonal with the occupation counts; S^2 commutes with both and
  its spectrum is s(s+1); apply_constraint is the matrix sum H + c0 (N - n)^2 + c1 S^2 + c2 S_z
  formed with matrix products.
* The oracle's network sigma of these operators on the determinant basis equals the dense
  qubit matrix restricted to the determinants (interleaved qubit order), so the whole chain --
  SVD grouping, network, oracle hvec -- is pinned for them.
"""
from functools import reduce

import numpy as np
import pytest

from oracle import glue
from oracle import oracle as O
from tests.helpers import integrals

I2 = np.eye(2)
X = np.array([[0.0, 1.0], [1.0, 0.0]])
Z = np.array([[1.0, 0.0], [0.0, -1.0]])
Y = np.array([[0.0, -1j], [1j, 0.0]])


def dense(terms, nq):
    """sum_t c_t prod_q X^{x_q} Z^{z_q}; qubit q is bit q of the basis index."""
    out = np.zeros((1 << nq, 1 << nq), dtype=complex)
    for x, z, c in terms:
        # kron lists the highest qubit first so that qubit q is bit q of the index
        mats = [np.linalg.matrix_power(X, (x >> q) & 1) @ np.linalg.matrix_power(Z, (z >> q) & 1)
                for q in reversed(range(nq))]
        out += c * reduce(np.kron, mats)
    return out


def pauli(nq, ops):
    """explicit Pauli string, ops = {qubit: 2x2 matrix}"""
    return reduce(np.kron, [ops.get(q, I2) for q in reversed(range(nq))])


@pytest.mark.parametrize("nq,pbc", [(2, False), (3, False), (4, True)])
def test_heisenberg_chain_equals_pauli_matrices(nq, pbc):
    jx, jy, jz = 0.7, -1.3, 0.4
    got = dense(glue.heisenberg_terms(nq, jx, jy, jz, pbc), nq)
    bonds = [(i, i + 1) for i in range(nq - 1)] + ([(nq - 1, 0)] if pbc else [])
    ref = sum(jx * pauli(nq, {i: X, j: X}) + jy * pauli(nq, {i: Y, j: Y}) + jz * pauli(nq, {i: Z, j: Z})
              for i, j in bonds)
    assert np.allclose(got, ref, atol=1e-13)
    if nq == 2 and not pbc:
        w = np.linalg.eigvalsh(dense(glue.heisenberg_terms(2), 2))
        assert np.allclose(sorted(w), [-3.0, 1.0, 1.0, 1.0])


@pytest.mark.parametrize("nq,pbc", [(3, False), (4, True)])
def test_ising_chain_equals_pauli_matrices(nq, pbc):
    J, h = 1.1, 0.35
    got = dense(glue.ising_terms(nq, J, h, pbc), nq)
    bonds = [(i, i + 1) for i in range(nq - 1)] + ([(nq - 1, 0)] if pbc else [])
    ref = sum(-J * pauli(nq, {i: Z, j: Z}) for i, j in bonds) + sum(-h * pauli(nq, {i: X}) for i in range(nq))
    assert np.allclose(got, ref, atol=1e-13)


@pytest.mark.parametrize("norb", [2, 3])
def test_spin_operators_spectrum(norb):
    nq = 2 * norb
    N, s2, sz = (dense(t, nq) for t in glue.spin_operators(norb))
    idx = np.arange(1 << nq)
    na = np.array([bin(i & int("01" * norb, 2)).count("1") for i in idx])   # alpha: even qubits
    nb = np.array([bin(i & int("10" * norb, 2)).count("1") for i in idx])
    assert np.allclose(N, np.diag(na + nb), atol=1e-13)
    assert np.allclose(sz, np.diag(0.5 * (na - nb)), atol=1e-13)
    assert np.allclose(s2 @ N, N @ s2, atol=1e-12) and np.allclose(s2 @ sz, sz @ s2, atol=1e-12)
    assert np.allclose(s2, s2.conj().T, atol=1e-13)
    w = np.linalg.eigvalsh(s2)
    allowed = [s * (s + 1) for s in np.arange(0, norb + 1, 0.5)]
    assert all(min(abs(v - a) for a in allowed) < 1e-10 for v in w)
    assert abs(w.max() - (norb / 2) * (norb / 2 + 1)) < 1e-10  # highest spin: norb unpaired electrons


def test_apply_constraint_is_the_matrix_sum():
    norb, nelec, c = 2, (1, 1), (0.7, 0.3, -0.2)
    _, _, h1, g2 = integrals(norb)
    h = [(int(glue.unzip_even(int(ax)) | glue.unzip_odd(int(bx))),
          int(glue.unzip_even(int(az)) | glue.unzip_odd(int(bz))), float(cc))
         for ax, bx, az, bz, cc in zip(*glue.jw_hamiltonian_terms(h1, g2, norb))]
    nq = 2 * norb
    N, s2, sz = (dense(t, nq) for t in glue.spin_operators(norb))
    shift = N - sum(nelec) * np.eye(1 << nq)
    ref = dense(glue.add_terms(h), nq) + c[0] * shift @ shift + c[1] * s2 + c[2] * sz
    got = dense(glue.apply_constraint(h, norb, nelec, c), nq)
    assert np.allclose(got, ref, atol=1e-12)


def _restricted(terms, n, na, nb):
    """dense qubit matrix restricted to the (na, nb) determinants, in the basis order"""
    A = [s for s in range(1 << n) if bin(s).count("1") == na]
    B = [s for s in range(1 << n) if bin(s).count("1") == nb]
    q = [glue.unzip_even(a) | glue.unzip_odd(b) for a in A for b in B]
    D = dense(terms, 2 * n)
    return np.real(D[np.ix_(q, q)])


@pytest.mark.parametrize("kind", ["heisenberg", "ising", "s2_penalty"])
@pytest.mark.parametrize("n,na,nb", [(3, 1, 2), (3, 2, 1), (4, 2, 2)])
def test_oracle_network_sigma_equals_restricted_matrix(kind, n, na, nb):
    if kind == "heisenberg":
        terms = glue.heisenberg_terms(2 * n, 0.8, 1.1, -0.6)
    elif kind == "ising":
        terms = glue.ising_terms(2 * n, 1.0, 0.4)
    else:
        _, _, h1, g2 = integrals(n)
        h = [(int(glue.unzip_even(int(ax)) | glue.unzip_odd(int(bx))),
              int(glue.unzip_even(int(az)) | glue.unzip_odd(int(bz))), float(cc))
             for ax, bx, az, bz, cc in zip(*glue.jw_hamiltonian_terms(h1, g2, n))]
        terms = glue.apply_constraint(h, n, (na, nb), (0.0, 0.5, 0.0))
    net = O.Network(glue.compress_by_svd(glue.term_arrays(terms)))
    S = O.System(n, na, nb, ham=net)
    M = _restricted(terms, n, na, nb)
    x = np.random.default_rng(5).uniform(-1, 1, S.dim)
    assert np.max(np.abs(S.hvec(x) - M @ x)) < 1e-12 * max(1.0, np.abs(M).max()) * S.dim

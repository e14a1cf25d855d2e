"""Shared builders for the parity tests (integrals, oracle systems, device handles)."""
from __future__ import annotations

import functools

import numpy as np

from binsim_amd.ansatz import ucc_spin_orbitals
from binsim_amd.integrals import chemist_integrals, for_norb, one_body_layout, reference_layout

SMALL = [(4, 2, 2), (5, 2, 3), (6, 3, 3), (6, 3, 2), (6, 1, 4), (7, 3, 3), (8, 4, 4), (8, 3, 5),
         (5, 0, 2), (5, 3, 0), (4, 4, 4)]


@functools.lru_cache(maxsize=None)
def integrals(n: int, seed: int | None = None):
    h, eri = for_norb(n) if seed is None else chemist_integrals(n, seed)
    return h, eri, one_body_layout(h), reference_layout(eri)


@functools.lru_cache(maxsize=None)
def oracle_ham(n: int, seed: int | None = None):
    from oracle import oracle as O
    _, _, h1, g2 = integrals(n, seed)
    return O.hamiltonian_network(h1, g2, n)


@functools.lru_cache(maxsize=None)
def oracle_pool(n: int, na: int, nb: int):
    from oracle import oracle as O
    return O.pool_network(ucc_spin_orbitals(n, na, nb))


def oracle_system(n, na, nb, seed=None, pool=False):
    from oracle import oracle as O
    return O.System(n, na, nb, ham=oracle_ham(n, seed), pool=oracle_pool(n, na, nb) if pool else None)


def seeded_vector(dim: int, seed: int):
    x = np.random.default_rng(seed).uniform(-1.0, 1.0, dim)
    return x / np.linalg.norm(x)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

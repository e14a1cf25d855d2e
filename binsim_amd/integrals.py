"""Seeded synthetic molecular integrals (real, 8-fold symmetric) in the reference layout.

Recipe (identical bytes on every x86-64 host with this image's glibc):

* RNG: SplitMix64 seeded with ``seed``; a uniform is ``(next() >> 11) * 2**-53``.
  Normal deviates come from Box--Muller on ``(u1, u2) = (1 - u, u')`` and BOTH outputs
  ``r*cos(2*pi*u2)`` then ``r*sin(2*pi*u2)`` are used, in that order.
* one-electron: draw an n*n matrix G of N(0,1) in row-major order;
  ``h = diag(eps) + 0.1 * G * exp(-|p-q|/3)`` with ``eps_p = -1.5 + 2.5 p/(n-1)``, then
  ``h <- (h + h^T)/2``.
* two-electron: for k = 1..2n draw the upper triangle (p <= q, row-major) of a symmetric
  L^k with entries ``0.2 * N(0,1) * exp(-|p-q|/4)``; ``(pq|rs) = sum_k L^k_pq L^k_rs``
  (accumulated in k order) and 0.5 is added to every (pp|pp).  This is positive
  semi-definite with 8-fold permutational symmetry.  E_core = 0.
* the reference consumes ``two_body_mo[p,q,r,s] = (ps|qr)`` as a Julia column-major array,
  i.e. flat index ``p + q n + r n^2 + s n^3`` (src/include/ham/ham.hpp:236-239,
  py/mole_pbc_int.py:10).  :func:`reference_layout` produces exactly that buffer.

Seeds are ``20261202 + id`` where id is the first workload using that orbital count:
n=12 -> 20261202, n=16 -> 20261203, n=18 -> 20261205.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

_MASK64 = (1 << 64) - 1
SEED_BASE = 20261202
SEEDS = {12: SEED_BASE + 0, 16: SEED_BASE + 1, 18: SEED_BASE + 3}


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) with a Box--Muller normal stream."""

    def __init__(self, seed: int):
        self.state = seed & _MASK64
        self._spare = None

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        if self._spare is not None:
            v, self._spare = self._spare, None
            return v
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        t = 2.0 * math.pi * u2
        self._spare = r * math.sin(t)
        return r * math.cos(t)


def chemist_integrals(n: int, seed: int):
    """Return (h[n,n], eri[n,n,n,n]) with eri[p,q,r,s] = (pq|rs) (chemist notation)."""
    rng = SplitMix64(seed)
    eps = np.array([-1.5 + 2.5 * p / (n - 1) if n > 1 else -1.5 for p in range(n)])
    g = np.empty((n, n))
    for p in range(n):
        for q in range(n):
            g[p, q] = rng.normal()
    idx = np.arange(n)
    dist = np.abs(idx[:, None] - idx[None, :]).astype(float)
    decay1 = np.array([[math.exp(-d / 3.0) for d in row] for row in dist])
    decay2 = np.array([[math.exp(-d / 4.0) for d in row] for row in dist])
    h = np.diag(eps) + 0.1 * g * decay1
    h = (h + h.T) / 2.0
    eri = np.zeros((n * n, n * n))
    for _ in range(2 * n):
        L = np.zeros((n, n))
        for p in range(n):
            for q in range(p, n):
                v = 0.2 * rng.normal() * decay2[p, q]
                L[p, q] = v
                L[q, p] = v
        Lf = L.reshape(-1)
        eri += np.multiply.outer(Lf, Lf)
    eri = eri.reshape(n, n, n, n)
    for p in range(n):
        eri[p, p, p, p] += 0.5
    return h, eri


def reference_layout(eri: np.ndarray) -> np.ndarray:
    """Flat float64 buffer with buf[p + q n + r n^2 + s n^3] = (ps|qr)."""
    # two_body_mo[p,q,r,s] = (ps|qr) = eri[p,s,q,r]
    tb = np.transpose(eri, (0, 2, 3, 1))  # tb[p,q,r,s] = eri[p,s,q,r]
    return np.ascontiguousarray(tb.transpose(3, 2, 1, 0)).reshape(-1)  # column-major flatten


def one_body_layout(h: np.ndarray) -> np.ndarray:
    """Flat float64 buffer with buf[p + q n] = h[p,q] (Julia column-major)."""
    return np.ascontiguousarray(h.T).reshape(-1)


def digest(h: np.ndarray, eri: np.ndarray) -> str:
    m = hashlib.sha256()
    m.update(np.ascontiguousarray(h, dtype="<f8").tobytes())
    m.update(np.ascontiguousarray(eri, dtype="<f8").tobytes())
    return m.hexdigest()


def for_norb(n: int):
    """Integrals of the benchmark workloads, keyed by orbital count."""
    return chemist_integrals(n, SEEDS.get(n, SEED_BASE + 100 + n))

"""CPU check of the colex rank update used by the alpha-beta sigma kernel's scatter
(binsim_amd/csrc/ab_src.hpp): for a string J with occupied orbitals p_0 < .. < p_{k-1}, moving the
electron of slot io (orbital r = p_io) to the iv-th empty orbital t changes the colex rank
sum_j C(p_j, j+1) by a per-io term plus a per-iv term, and the sign of <J - r + t| E_tr |J> is the
parity of the occupied orbitals strictly between r and t.  Brute force over every string and
slot up to 12 orbitals."""
from itertools import combinations
from math import comb

import pytest


def _update(s, n, io, iv):
    """The kernel's arithmetic, term by term (same prefix sums and the same case split)."""
    occ = [p for p in range(n) if s >> p & 1]
    vir = [p for p in range(n) if not s >> p & 1]
    A = B = 0
    up = dn = None
    for j, p in enumerate(occ):
        c0, c1, c2 = comb(p, j), comb(p, j + 1), comb(p, j + 2)
        if j == io:
            up, dn = -c1 - (A + c0 - c1), -c1 + B
        A += c0 - c1
        B += c2 - c1
    t = vir[iv]
    m = t - iv
    F, G = comb(t, m), comb(t, m + 1)
    for j, p in enumerate(occ):
        if p < t:
            c1 = comb(p, j + 1)
            F += comb(p, j) - c1
            G -= comb(p, j + 2) - c1
    above = m > io
    return (up + F if above else dn + G), (m + io + (1 if above else 0)) & 1


@pytest.mark.parametrize("n", range(2, 13))
def test_colex_rank_update(n):
    for k in range(1, n):
        strs = sorted(sum(1 << p for p in c) for c in combinations(range(n), k))
        rank = {s: i for i, s in enumerate(strs)}
        for J, s in enumerate(strs):
            occ = [p for p in range(n) if s >> p & 1]
            vir = [p for p in range(n) if not s >> p & 1]
            for io, r in enumerate(occ):
                for iv, t in enumerate(vir):
                    d, neg = _update(s, n, io, iv)
                    lo, hi = min(r, t), max(r, t)
                    between = bin(s & (((1 << hi) - 1) ^ ((1 << (lo + 1)) - 1))).count("1") & 1
                    assert J + d == rank[s ^ (1 << r) ^ (1 << t)]
                    assert neg == between

"""UCC excitation lists in the reference's parameter order.

The order is the one generate_ci_spin_orbitals produces (jl/ansatz.jl:67-167, generalize =
false, every orbsym 0): for ka = 0..2 alpha excitation ranks, each (alpha occupied combo,
alpha virtual combo) in lexicographic order, then kb = 0..2-ka beta ranks likewise; entries are
[alpha virtuals, beta virtuals, alpha occupied, beta occupied] as interleaved spin-orbitals
(2p alpha, 2p+1 beta); the identity entry (ka = kb = 0) is dropped.
"""
from __future__ import annotations

from itertools import combinations


def _excitations(occ, vir, k):
    return [(o, v) for o in combinations(occ, k) for v in combinations(vir, k)]


def ucc_spin_orbitals(norb: int, na: int, nb: int, kmax: int = 2):
    ex_a = {k: _excitations(range(na), range(na, norb), k) for k in range(min(na, kmax) + 1)}
    ex_b = {k: _excitations(range(nb), range(nb, norb), k) for k in range(min(nb, kmax) + 1)}
    ops = []
    for ka in range(min(na, kmax) + 1):
        for oa, va in ex_a[ka]:
            for kb in range(min(nb, kmax - ka) + 1):
                for ob, vb in ex_b[kb]:
                    if ka == 0 and kb == 0:
                        continue
                    ops.append([2 * p for p in va] + [2 * p + 1 for p in vb]
                               + [2 * p for p in oa] + [2 * p + 1 for p in ob])
    return ops

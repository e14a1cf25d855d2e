"""The seeded synthetic integrals are bit-reproducible and have the advertised structure."""
import numpy as np
import pytest

from binsim_amd.integrals import chemist_integrals, digest, for_norb, one_body_layout, reference_layout

# SHA-256 over (h, eri) little-endian float64; regenerated identically on every host
DIGESTS = {12: "810970b6f717159a", 16: "0b80e2be24d9def8", 18: "387a269b0578a064"}


@pytest.mark.parametrize("n", sorted(DIGESTS))
def test_digest_pinned(n):
    h, eri = for_norb(n)
    assert digest(h, eri)[:16] == DIGESTS[n]


@pytest.mark.parametrize("n", [4, 7, 12])
def test_symmetry_and_psd(n):
    h, eri = chemist_integrals(n, 123)
    assert np.array_equal(h, h.T)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.max(np.abs(eri - np.transpose(eri, perm))) < 1e-15
    assert np.linalg.eigvalsh(eri.reshape(n * n, n * n)).min() > -1e-12
    eps = [-1.5 + 2.5 * p / (n - 1) for p in range(n)]
    assert np.max(np.abs(np.diag(h) - eps)) < 0.5


def test_reference_layout():
    n = 5
    h, eri = chemist_integrals(n, 1)
    g2 = reference_layout(eri)
    h1 = one_body_layout(h)
    for p, q, r, s in [(0, 1, 2, 3), (4, 0, 1, 1), (2, 2, 3, 4)]:
        # two_body_mo[p,q,r,s] = (ps|qr) at flat p + q n + r n^2 + s n^3
        assert g2[p + q * n + r * n * n + s * n ** 3] == eri[p, s, q, r]
    assert h1[1 + 3 * n] == h[1, 3]

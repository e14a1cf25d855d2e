"""The sigma build at the benchmark shapes themselves, checked through what holds at any size.

* 16 orbitals, 8+8 electrons (165,636,900 determinants, the benchmark workload) and 18
  orbitals, 7+7 electrons (1,012,766,976 determinants): the FP64 sigma of the seeded benchmark
  vector against the reference CPU library's sigma of the same vector
  (tests/golden/config1_n16.npz, config3_n18.npz: sum, squared norm, dot with a second seeded
  vector, 4096 sampled entries; the diagonal's head and the same samples), within 1e-12
  relative.
* 16 orbitals and 18 orbitals / 7+7 electrons (1,012,766,976 determinants, 8.1 GB vectors):
  Hermiticity <y|Hx> = <Hy|x> (1e-12 relative), the closed-shell HF energy
  2 sum_i h_ii + sum_ij [2 (ii|jj) - (ij|ji)] at index 0 of the diagonal and of H e_0 (1e-12),
  FP32 sigma within 2e-5 relative of FP64, and (16 orbitals) bitwise reproducibility of the
  deterministic alpha-beta scatter.
"""
import os

import numpy as np
import pytest

from binsim_amd import autotune as AT
from binsim_amd import otf
from binsim_amd.integrals import for_norb, one_body_layout, reference_layout

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _system(n, k):
    h, eri = for_norb(n)
    B = otf.HipBasisManager(otf.BasisManager(n, k, k))
    H = otf.HipOTF.from_integrals(n, one_body_layout(h), reference_layout(eri))
    return h, eri, B, H, otf.HipOTF_Functions(B, ham=H)


def _seeded_host(torch, dim, seed):
    x = np.random.default_rng(seed).uniform(-1.0, 1.0, dim)
    x /= np.linalg.norm(x)
    return torch.from_numpy(x).to("cuda")


def _hf_energy(h, eri, k):
    occ = range(k)
    e = 2.0 * sum(h[i, i] for i in occ)
    e += sum(2.0 * eri[i, i, j, j] - eri[i, j, j, i] for i in occ for j in occ)
    return e


def _properties(torch, n, k, det_check):
    print(f"full-size properties n={n} k={k}", flush=True)
    h, eri, B, H, F = _system(n, k)
    dim = B.dim
    g = torch.Generator(device="cuda").manual_seed(7 + n)
    x = torch.rand(dim, dtype=torch.float64, device="cuda", generator=g) - 0.5
    y = torch.rand(dim, dtype=torch.float64, device="cuda", generator=g) - 0.5
    hx, hy = torch.empty_like(x), torch.empty_like(x)
    F.hvec(x, hx)
    F.hvec(y, hy)
    a, b = float(torch.dot(y, hx)), float(torch.dot(hy, x))
    assert abs(a - b) <= 1e-12 * max(abs(a), float(hx.norm() * y.norm()) * 1e-3)
    # FP32 sigma of the same vector, alone and as the first of a pair (16 orbitals: the pair on
    # the class-split kernel too)
    h32 = torch.empty_like(x)
    F.hvec_block([x], [h32], 1)
    assert float((h32 - hx).norm() / hx.norm()) < 2e-5
    p32, q32 = torch.empty_like(x), torch.empty_like(x)
    kernels = (1, 4) if n <= 16 else (1,)
    saved = AT.get_config(AT.FP32)
    try:
        for kern in kernels:
            AT.set_config(AT.FP32, dict(saved, ab_kernel=kern, ab_nvs=2))
            F.hvec_block([x, y], [p32, q32], 1)
            assert float((p32 - hx).norm() / hx.norm()) < 2e-5
            assert float((q32 - hy).norm() / hy.norm()) < 2e-5
    finally:
        AT.set_config(AT.FP32, saved)
    del y, hy, h32, p32, q32
    # closed-shell HF determinant (index 0): diagonal and H e_0
    ehf = _hf_energy(h, eri, k)
    d = torch.zeros_like(x)
    F.get_diags(d)
    assert abs(float(d[0]) - ehf) <= 1e-12 * abs(ehf)
    e0 = torch.zeros_like(x)
    e0[0] = 1.0
    F.hvec(e0, hx)
    assert abs(float(hx[0]) - ehf) <= 1e-12 * abs(ehf)
    if det_check:
        saved = AT.get_config(AT.FP64)
        try:
            AT.set_config(AT.FP64, dict(saved, deterministic=1))
            F.hvec(x, hx)
            hx2 = torch.empty_like(x)
            F.hvec(x, hx2)
            assert torch.equal(hx, hx2)
        finally:
            AT.set_config(AT.FP64, saved)


@pytest.mark.parametrize("n,k,fixture", [(16, 8, "config1_n16.npz"), (18, 7, "config3_n18.npz")])
def test_sigma_full_size_matches_reference(torch_cuda, n, k, fixture):
    torch = torch_cuda
    print(f"full-size reference sigma n={n} k={k}", flush=True)
    f = np.load(os.path.join(GOLDEN, fixture))
    _, _, B, H, F = _system(n, k)
    dim = B.dim
    x = _seeded_host(torch, dim, 1000 + n)
    s = torch.empty_like(x)
    F.hvec(x, s)
    del x
    scale = float(np.sqrt(f["sigma_norm2"]))
    assert abs(float(s.sum()) - float(f["sigma_sum"])) <= 1e-12 * scale * np.sqrt(dim)
    assert abs(float(torch.dot(s, s)) / float(f["sigma_norm2"]) - 1.0) < 1e-12
    y = _seeded_host(torch, dim, 4242)
    dy = float(torch.dot(y, s))
    assert abs(dy - float(f["sigma_dot_y"])) <= 1e-12 * scale
    del y
    idx = torch.from_numpy(f["sample_idx"]).to("cuda")
    got = s[idx].cpu().numpy()
    assert np.max(np.abs(got - f["sample_val"])) <= 1e-12 * float(s.abs().max())
    d = torch.zeros_like(s)
    F.get_diags(d)
    dmax = float(np.max(np.abs(f["diag_head"])))
    assert np.max(np.abs(d[:256].cpu().numpy() - f["diag_head"])) <= 1e-12 * dmax
    assert np.max(np.abs(d[idx].cpu().numpy() - f["diag_sample"])) <= 1e-12 * dmax
    del s, d
    torch.cuda.empty_cache()


def test_sigma_16_orbitals_properties(torch_cuda):
    _properties(torch_cuda, 16, 8, det_check=True)


def test_sigma_18_orbitals_properties(torch_cuda):
    _properties(torch_cuda, 18, 7, det_check=False)
    torch_cuda.cuda.empty_cache()

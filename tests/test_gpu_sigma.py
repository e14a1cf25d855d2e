"""GPU parity of the sigma build and diagonal against the oracle (network restatement).

Tolerance: FP64 sigma within 1e-12 relative (2-norm and per-element max / max|sigma|), the
bar the operator path is held to; the FP32 path within 2e-5 relative of the FP64 result.
"""
import numpy as np
import pytest

from binsim_amd import otf
from tests.conftest import gpu_vec, host
from tests.helpers import SMALL, integrals, oracle_system, oracle_ham, rel_err, seeded_vector

pytestmark = pytest.mark.gpu


def _dev(n, na, nb, seed=None):
    _, _, h1, g2 = integrals(n, seed)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    H = otf.HipOTF.from_integrals(n, h1, g2, 0.0)
    return B, H, otf.HipOTF_Functions(B, ham=H)


# beyond 16 orbitals (36-bit interleaved determinants).  The alpha-beta part runs on the
# fallback kernels, except where the source-lane kernel's 512-entry-row form takes it: FP32 with
# at most 8 beta electrons and 12 holes ((18, 1, 6), (17, 2, 5)), FP64 with at most 7 and 11
# ((17, 1, 6) in one launch; (18, 1, 7) with its 254 KB destination row in two halves)
WIDE = [(17, 3, 3), (18, 2, 2), (18, 3, 2), (18, 1, 6), (17, 2, 5), (17, 1, 6), (18, 1, 7)]


@pytest.mark.parametrize("n,na,nb", SMALL + [(12, 6, 6)] + WIDE)
def test_hvec_matches_oracle(torch_cuda, n, na, nb):
    print(f"hvec n={n} na={na} nb={nb}", flush=True)
    S = oracle_system(n, na, nb)
    x = seeded_vector(S.dim, 11 + n)
    ref = S.hvec(x)
    B, H, F = _dev(n, na, nb)
    xd = gpu_vec(torch_cuda, x)
    yd = torch_cuda.empty_like(xd)
    F.hvec(xd, yd)
    y = host(yd)
    assert rel_err(y, ref) < 1e-12
    assert np.max(np.abs(y - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("n,na,nb", SMALL + [(12, 6, 6)] + WIDE)
def test_diags_match_oracle(torch_cuda, n, na, nb):
    S = oracle_system(n, na, nb)
    ref = S.diags()
    B, H, F = _dev(n, na, nb)
    d = gpu_vec(torch_cuda, np.full(S.dim, 0.25))  # accumulates into the caller's buffer
    F.get_diags(d)
    assert np.max(np.abs(host(d) - 0.25 - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (8, 3, 5), (12, 6, 6), (18, 3, 2), (18, 1, 6)])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 8, 9])
def test_block_sigma(torch_cuda, n, na, nb, k):
    S = oracle_system(n, na, nb)
    xs = [seeded_vector(S.dim, 100 + j) for j in range(k)]
    refs = [S.hvec(x) for x in xs]
    B, H, F = _dev(n, na, nb)
    xd = [gpu_vec(torch_cuda, x) for x in xs]
    yd = [torch_cuda.empty_like(v) for v in xd]
    F.hvec_block(xd, yd, precision=0)
    for y, r in zip(yd, refs):
        assert rel_err(host(y), r) < 1e-12
    F.hvec_block(xd, yd, precision=1)
    for y, r in zip(yd, refs):
        assert rel_err(host(y), r) < 2e-5


@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (8, 3, 5), (9, 4, 5), (10, 5, 2), (12, 6, 6), (12, 5, 6)])
@pytest.mark.parametrize("k", [2, 3, 4])
def test_class_split_fp32_pairs(torch_cuda, n, na, nb, k):
    """FP32 block sigma with vector pairs on the class-split alpha-beta kernel (ab_kernel 4; an
    odd last vector on the single-vector source-lane kernel) against the oracle."""
    from binsim_amd import autotune as AT
    S = oracle_system(n, na, nb)
    xs = [seeded_vector(S.dim, 500 + j) for j in range(k)]
    refs = [S.hvec(x) for x in xs]
    B, H, F = _dev(n, na, nb)
    xd = [gpu_vec(torch_cuda, x) for x in xs]
    yd = [torch_cuda.empty_like(v) for v in xd]
    saved = AT.get_config(AT.FP32)
    try:
        AT.set_config(AT.FP32, dict(saved, ab_kernel=4))
        F.hvec_block(xd, yd, precision=1)
    finally:
        AT.set_config(AT.FP32, saved)
    for y, r in zip(yd, refs):
        assert rel_err(host(y), r) < 2e-5


# matrix-core alpha-beta kernel (ab_kernel 5): every n <= 16, including 16 orbitals with the
# full 8-electron beta space (12,870 strings in the LDS row) and FP64 in two connection passes
MMA = [(4, 2, 2), (6, 3, 3), (8, 3, 5), (10, 5, 2), (12, 6, 6), (12, 5, 6), (14, 3, 4), (16, 2, 2),
       (16, 3, 2), (16, 1, 8), (16, 2, 8)]


@pytest.mark.parametrize("n,na,nb", MMA)
def test_mfma_alpha_beta(torch_cuda, n, na, nb):
    """FP64 sigma (one vector) and FP32 block sigma (3 interleaved vectors) with the alpha-beta
    part on the matrix-core kernel, against the oracle."""
    from binsim_amd import autotune as AT
    print(f"mfma alpha-beta n={n} na={na} nb={nb}", flush=True)
    S = oracle_system(n, na, nb)
    xs = [seeded_vector(S.dim, 700 + j) for j in range(3)]
    refs = [S.hvec(x) for x in xs]
    B, H, F = _dev(n, na, nb)
    xd = [gpu_vec(torch_cuda, x) for x in xs]
    yd = [torch_cuda.empty_like(v) for v in xd]
    s64, s32 = AT.get_config(AT.FP64), AT.get_config(AT.FP32)
    try:
        AT.set_config(AT.FP64, dict(s64, ab_kernel=5, deterministic=0))
        AT.set_config(AT.FP32, dict(s32, ab_kernel=5, deterministic=0))
        F.hvec(xd[0], yd[0])
        y = host(yd[0])
        assert rel_err(y, refs[0]) < 1e-12
        assert np.max(np.abs(y - refs[0])) < 1e-12 * np.max(np.abs(refs[0]))
        F.hvec_block(xd, yd, precision=1)
        for y, r in zip(yd, refs):
            assert rel_err(host(y), r) < 2e-5
    finally:
        AT.set_config(AT.FP64, s64)
        AT.set_config(AT.FP32, s32)


@pytest.mark.parametrize("n,na,nb", [(17, 1, 6), (18, 1, 7)])
def test_sigma_512_entry_forms_both_precisions(torch_cuda, n, na, nb):
    """The source-lane kernel's 512-entry forms with at most 7 beta electrons and 11 holes:
    FP64 (one launch at (17, 1, 6); the destination row in two halves at (18, 1, 7)) and the
    single-vector FP32 form, against the oracle."""
    print(f"512-entry forms n={n} na={na} nb={nb}", flush=True)
    S = oracle_system(n, na, nb)
    x = seeded_vector(S.dim, 300 + n)
    ref = S.hvec(x)
    B, H, F = _dev(n, na, nb)
    xd = gpu_vec(torch_cuda, x)
    yd = torch_cuda.empty_like(xd)
    F.hvec_block([xd], [yd], precision=0)
    assert rel_err(host(yd), ref) < 1e-12
    # the ordered scatter is bitwise reproducible, also when the row is done in parts
    from binsim_amd import autotune as AT
    saved = AT.get_config(AT.FP64)
    try:
        AT.set_config(AT.FP64, dict(saved, deterministic=1))
        y2 = torch_cuda.empty_like(xd)
        F.hvec(xd, yd)
        F.hvec(xd, y2)
        assert torch_cuda.equal(yd, y2)
        assert rel_err(host(y2), ref) < 1e-12
    finally:
        AT.set_config(AT.FP64, saved)
    F.hvec_block([xd], [yd], precision=1)
    assert rel_err(host(yd), ref) < 2e-5


@pytest.mark.parametrize("n,na,nb", [(4, 2, 2), (6, 3, 2), (8, 4, 4)])
def test_network_handle_decodes_to_same_operator(torch_cuda, n, na, nb):
    """build_network_otf_f64 + build_networkdev_f64 (the drop-in path) == integral path."""
    S = oracle_system(n, na, nb)
    x = seeded_vector(S.dim, 5)
    ref = S.hvec(x)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    Hn = otf.HipOTF.from_network(n, oracle_ham(n).flat)
    F = otf.HipOTF_Functions(B, ham=Hn)
    xd = gpu_vec(torch_cuda, x)
    yd = torch_cuda.empty_like(xd)
    F.hvec(xd, yd)
    assert rel_err(host(yd), ref) < 1e-12

"""Launch-choice table (hipotf_set_tuning / hipotf_get_tuning) and the autotuner's JSON store.

The CPU tests exercise the host-only parts: validation on both sides of the C ABI, per-precision
state, and the JSON round trip.  The GPU test tunes a 10-orbital system and checks that the
chosen configuration reproduces the default configuration's sigma and that the stored entry
is applied by ``load_or_tune``.
"""
import json

import numpy as np
import pytest

import binsim_amd.lib as L
from binsim_amd import autotune as AT


@pytest.fixture
def restore_tuning():
    saved = [AT.get_config(p) for p in (AT.FP64, AT.FP32)]
    yield
    for p, c in zip((AT.FP64, AT.FP32), saved):
        AT.set_config(p, c)


def test_defaults_and_per_precision_state(restore_tuning):
    d = AT.get_config(AT.FP64)
    assert set(d) == set(AT.FIELDS)
    assert d["ab_kernel"] == 0 and d["ab_split"] == 1 and d["rs_passes"] == 0 and d["rs_block"] == 8
    assert d["sweep_mode"] == 4
    AT.set_config(AT.FP32, dict(ab_kernel=1, ab_split=2, ab_db=2, ab_nvs=1, rs_passes=4, rs_block=16))
    assert AT.get_config(AT.FP32) == dict(ab_kernel=1, ab_split=2, ab_db=2, ab_nvs=1, rs_passes=4,
                                          deterministic=0, rs_block=16, sweep_mode=4)
    assert AT.get_config(AT.FP64) == d  # the FP64 table is untouched


def test_invalid_values_rejected_by_library_and_host(restore_tuning):
    lib = L.load()
    before = AT.get_config(AT.FP64)
    for bad in ([6, 1, 1, 2, 1, 0, 8, 0], [0, 3, 1, 2, 1, 0, 8, 0], [0, 1, 0, 2, 1, 0, 8, 0],
                [0, 1, 1, 3, 1, 0, 8, 0], [0, 1, 1, 2, 3, 0, 8, 0], [0, 1, 1, 2, 1, 2, 8, 0],
                [0, 1, 1, 2, 1, 0, 2, 0], [0, 1, 1, 2, 1, 0, 8, 8]):
        v = np.array(bad, dtype=np.int32)
        assert lib.hipotf_set_tuning(0, v.ctypes.data_as(L._i32p), 8) != 0
        assert "tuning" in L.last_error()
    v = np.array([0, 1, 1, 2, 1, 0, 8, 0], dtype=np.int32)
    assert lib.hipotf_set_tuning(2, v.ctypes.data_as(L._i32p), 8) != 0     # bad precision
    assert lib.hipotf_set_tuning(0, v.ctypes.data_as(L._i32p), 7) != 0     # wrong length
    assert lib.hipotf_set_tuning(0, None, 8) != 0
    assert AT.get_config(AT.FP64) == before  # a rejected call changes nothing
    with pytest.raises(ValueError):
        AT.validate(dict(rs_passes=3))
    with pytest.raises(ValueError):
        AT.validate(dict(tile=8))


def test_candidates_cover_kernel_families_and_waves():
    c64 = AT.candidates(AT.FP64, 1)
    c32 = AT.candidates(AT.FP32, 4)
    assert {c["ab_kernel"] for c in c64} == {1, 2, 3, 5}
    assert {c["ab_split"] for c in c64 if c["ab_kernel"] == 1} == {1, 2}
    assert {c["ab_nvs"] for c in c32 if c["ab_kernel"] == 1} == {1, 2}
    assert {4, 5} <= {c["ab_kernel"] for c in c32}
    assert len(c32) == 12 and len(c64) == 7


def test_json_store_round_trip(tmp_path):
    p = str(tmp_path / "t.json")
    assert AT.load_file(p) == {}
    entry = {"fp64": {"config": AT.validate(dict(ab_kernel=1)), "ms": 1.5},
             "fp32": {"config": AT.validate(dict(ab_nvs=1)), "ms": float("inf"), "k": 4}}
    AT.save_file({"devA": {"n10_a5_b5": AT._jsonable(entry)}}, p)
    got = AT.load_file(p)
    assert got["devA"]["n10_a5_b5"]["fp64"]["config"]["ab_kernel"] == 1
    assert got["devA"]["n10_a5_b5"]["fp32"]["ms"] is None  # non-finite timings stored as null
    json.loads(open(p).read())  # plain JSON
    with open(p, "w") as f:
        f.write("[1, 2]")
    with pytest.raises(ValueError):
        AT.load_file(p)


@pytest.mark.gpu
def test_tuned_sigma_matches_default_and_is_applied(torch_cuda, tmp_path, restore_tuning):
    torch = torch_cuda
    from binsim_amd import otf
    from tests.helpers import integrals

    n, na, nb = 10, 5, 5
    _, _, h1, g2 = integrals(n)
    basis = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    ham = otf.HipOTF.from_integrals(n, h1, g2)
    F = otf.HipOTF_Functions(basis, ham)
    g = torch.Generator(device="cuda").manual_seed(5)
    xs = [torch.randn(basis.dim, dtype=torch.float64, device="cuda", generator=g) for _ in range(4)]
    ref64 = torch.empty_like(xs[0])
    F.hvec(xs[0], ref64)
    ref32 = [torch.empty_like(x) for x in xs]
    F.hvec_block(xs, ref32, 1)
    torch.cuda.synchronize()
    path = str(tmp_path / "tuning.json")
    entry = AT.tune(F, basis, ham, k=4, nroots=1, path=path, reps=2)
    assert AT.get_config(AT.FP64) == entry["fp64"]["config"]
    assert AT.get_config(AT.FP32) == entry["fp32"]["config"]
    assert entry["davidson"]["block"] in (1, 2, 4, 8)
    # every kernel family and launch variant reproduced the first candidate's sigma
    assert not entry["fp64"]["rejected"] and not entry["fp32"]["rejected"]
    y = torch.empty_like(xs[0])
    F.hvec(xs[0], y)
    ys = [torch.empty_like(x) for x in xs]
    F.hvec_block(xs, ys, 1)
    torch.cuda.synchronize()
    assert float((y - ref64).abs().max()) <= 1e-12 * float(ref64.abs().max())
    for a, b in zip(ys, ref32):
        assert float((a - b).abs().max()) <= 1e-5 * float(b.abs().max())
    # the stored entry is found and installed again
    AT.set_config(AT.FP64, {})
    AT.set_config(AT.FP32, {})
    e2, tuned = AT.load_or_tune(F, basis, ham, path=path)
    assert not tuned
    assert AT.get_config(AT.FP64) == entry["fp64"]["config"]
    assert AT.get_config(AT.FP32) == entry["fp32"]["config"]

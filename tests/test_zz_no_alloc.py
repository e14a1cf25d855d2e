"""hipotf_hvec_noalloc: the allocation-free variant of hvec_cuda (its four sigma work buffers in
the library's static device scratch array instead of device allocations).  It has no
counterpart in the reference; it serves callers that cannot allocate device memory.

Checked through the C ABI against hvec_cuda (FP64 sigma within 1e-12 relative; the alpha-beta
scatter's summation order is not fixed, so not bitwise) on two calls in a row, the second on a
smaller system, plus its two refusals: a basis beyond the scratch limit and an operator whose
device tables were never built (it must not build them).
"""
import numpy as np
import pytest

import binsim_amd.lib as L
from binsim_amd import otf
from tests.conftest import gpu_vec, host
from tests.helpers import integrals, rel_err, seeded_vector

pytestmark = pytest.mark.gpu

MAX_DIM = (64 << 20) // 32


def _entry():
    return L.load().hipotf_hvec_noalloc


def _system(n, na, nb):
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    H = otf.HipOTF.from_integrals(n, h1, g2)
    return B, H, otf.HipOTF_Functions(B, ham=H)


def test_noalloc_sigma_matches_hvec_twice(torch_cuda):
    f = _entry()
    for n, na, nb in ((12, 6, 6), (8, 4, 3)):  # second call: a smaller input
        print(f"noalloc sigma n={n} na={na} nb={nb}", flush=True)
        B, H, F = _system(n, na, nb)
        assert B.dim <= MAX_DIM
        x = gpu_vec(torch_cuda, seeded_vector(B.dim, 31 + n))
        ref = torch_cuda.empty_like(x)
        F.hvec(x, ref)  # the allocating entry point (also builds the device tables)
        y = torch_cuda.full_like(x, np.nan)
        assert f(B.ptr, H.ptr, x.data_ptr(), y.data_ptr()) == 0, L.last_error()
        torch_cuda.cuda.synchronize()
        assert rel_err(host(y), host(ref)) < 1e-12
        # repeated on the same inputs: the scratch keeps the previous contents, nothing
        # depends on them
        y2 = torch_cuda.full_like(x, np.nan)
        assert f(B.ptr, H.ptr, x.data_ptr(), y2.data_ptr()) == 0, L.last_error()
        torch_cuda.cuda.synchronize()
        assert rel_err(host(y2), host(ref)) < 1e-12


def test_noalloc_refuses_what_it_cannot_do_without_allocating(torch_cuda):
    f = _entry()
    # device tables never built for this basis: refused, not built
    B, H, F = _system(6, 3, 3)
    x = gpu_vec(torch_cuda, seeded_vector(B.dim, 5))
    y = torch_cuda.empty_like(x)
    assert f(B.ptr, H.ptr, x.data_ptr(), y.data_ptr()) == 1
    assert "hipotf_prepare" in L.last_error()
    # beyond the scratch limit: refused before any work
    Bb = otf.HipBasisManager(otf.BasisManager(13, 6, 6))
    assert Bb.dim > MAX_DIM
    assert f(Bb.ptr, H.ptr, 0, 0) == 1
    assert "exceed" in L.last_error()

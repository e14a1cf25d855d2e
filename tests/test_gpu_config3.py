"""The 8-root mixed block Davidson (k = 8) at 18 orbitals / 7+7 electrons (1,012,766,976
determinants, 8.1 GB FP64 vectors), bounded to three iterations.

What fits 288 GB is the point here: 16 + 16 FP64 subspace vectors (259 GB), the diagonal, the
residuals formed in free subspace slots and the sigma passes sized to the memory left
(hipotf_davidson_f64 checks all of it before allocating).  The run must complete its three
iterations (status "maxiter"), with 8 finite Ritz values in ascending order, the lowest below the
closed-shell HF energy (the first diagonal element), and one FP32 sigma per initial vector and
correction.  The full solve takes about an hour at this size.
"""
import numpy as np
import pytest

from binsim_amd import davidson as D
from binsim_amd import otf
from binsim_amd.integrals import for_norb, one_body_layout, reference_layout

pytestmark = pytest.mark.gpu


def test_block8_three_iterations_at_18_orbitals(torch_cuda):
    torch = torch_cuda
    print("18 orbitals 7+7, 8 roots, block 8, maxspace 16, 3 iterations", flush=True)
    n, k = 18, 7
    h, eri = for_norb(n)
    B = otf.HipBasisManager(otf.BasisManager(n, k, k))
    H = otf.HipOTF.from_integrals(n, one_body_layout(h), reference_layout(eri))
    torch.cuda.empty_cache()
    e, _, info = D.solve(B, H, nroots=8, mixed=True, tol=1e-6, maxiter=3, block=8, maxspace=16,
                         verbose=True)
    print(info, e, flush=True)
    assert info["iterations"] == 3 and info["converged"] == 0
    assert info["sigma32"] == 8 + 3 * 8 and info["sigma64"] == 0
    assert np.all(np.isfinite(e)) and np.all(np.diff(e) >= -1e-12)
    occ = range(k)
    ehf = 2.0 * sum(h[i, i] for i in occ) + sum(2.0 * eri[i, i, j, j] - eri[i, j, j, i] for i in occ for j in occ)
    assert e[0] < ehf
    assert np.isfinite(info["max_residual"]) and info["max_residual"] > 0

"""The VQE loop (binsim_amd.vqe.run_vqe, jl/vqe.jl:163-209) on the device objective against the
same loop on the oracle's energy_objective (jl/vqe.jl:1-21): 6 orbitals 3+3 electrons, all 117
UCC parameters from theta = 0, both run to convergence.  Energies within 1e-9 (the two loops see
objective values that differ at rounding level, so their iterates may part slightly; the
minimum they reach does not); the returned state is U(x_opt)|HF> within 1e-10; the checkpoint
holds the last evaluated x."""
import numpy as np
import pytest

from binsim_amd import otf
from binsim_amd import vqe as V
from binsim_amd.ansatz import ucc_spin_orbitals
from oracle import oracle as O
from tests.conftest import host
from tests.helpers import integrals, oracle_system

pytestmark = pytest.mark.gpu


def test_run_vqe_matches_oracle_loop(torch_cuda, tmp_path):
    n, na, nb = 6, 3, 3
    print(f"run_vqe n={n} na={na} nb={nb}", flush=True)
    S = oracle_system(n, na, nb, pool=True)
    _, _, h1, g2 = integrals(n)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    sos = ucc_spin_orbitals(n, na, nb)
    F = otf.HipOTF_Functions(B, ham=otf.HipOTF.from_integrals(n, h1, g2, 0.0),
                             pool=otf.HipOTF.from_spin_orbitals(n, sos))
    idxs = np.arange(len(sos))
    x0 = np.zeros(len(idxs))
    ref = torch_cuda.zeros(B.dim, dtype=torch_cuda.float64, device="cuda")
    ref[0] = 1.0
    p = str(tmp_path / "x.bin")
    e_dev, psi, x_dev, info = V.run_vqe(F, idxs, x0, ref, save_path=p, gtol=1e-8, ftol=1e-15)
    assert info["success"], info

    def fg(x):
        e, g, _ = O.energy_objective(S, idxs, x, O.hf_vector(S), np.zeros(S.dim))
        return e, g

    e_cpu, x_cpu, info_cpu = V.minimize_fg(fg, x0, gtol=1e-8, ftol=1e-15)
    assert info_cpu["success"]
    e_hf = O.energy_objective(S, idxs, x0, O.hf_vector(S), np.zeros(S.dim))[0]
    assert e_dev < e_hf - 1e-3
    assert abs(e_dev - e_cpu) < 1e-9
    state = O.hf_vector(S)
    for i, t in zip(idxs, x_dev):
        S.expm(i, t, state)
    assert np.max(np.abs(host(psi) - state)) < 1e-10
    assert abs(float(host(psi) @ S.hvec(host(psi))) - e_dev) < 1e-9
    assert V.load_x(p).shape == x_dev.shape

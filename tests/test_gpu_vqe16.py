"""VQE adjoint gradient at 16 orbitals / 8+8 electrons (165,636,900 determinants, FP64 states):
the full 5,792-parameter UCC ansatz with theta ~ U[-0.02, 0.02] (seeded), the fused energy +
adjoint gradient (jl/vqe.jl:1-21) against central finite differences of the energy-only entry
point (h = 1e-6) on sampled parameters covering every generator kind, and the reverse sweep
returning the state to Hartree-Fock.

Tolerances: gradient vs central differences 1e-8 (the FD truncation error at h = 1e-6 is
~h^2 |E'''| < 1e-11; rounding ~1e-16 |E| / h ~ 1e-9); energy-only vs fused energy 1e-12;
state back at HF within 1e-12 element-wise.
"""
import numpy as np
import pytest

from binsim_amd import otf
from binsim_amd.ansatz import ucc_spin_orbitals
from binsim_amd.integrals import SEED_BASE, for_norb, one_body_layout, reference_layout
from binsim_amd.vqe import fd_gradient

pytestmark = pytest.mark.gpu


def test_vqe_16_orbitals_adjoint_gradient_matches_fd(torch_cuda):
    torch = torch_cuda
    n, na, nb = 16, 8, 8
    print("vqe 16 orbitals: building", flush=True)
    h, eri = for_norb(n)
    sos = ucc_spin_orbitals(n, na, nb)
    assert len(sos) == 5792
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    F = otf.HipOTF_Functions(B, ham=otf.HipOTF.from_integrals(n, one_body_layout(h), reference_layout(eri)),
                             pool=otf.HipOTF.from_spin_orbitals(n, sos))
    th = np.random.default_rng(SEED_BASE + 2000).uniform(-0.02, 0.02, len(sos))
    idx = np.arange(len(sos))
    hf = torch.zeros(B.dim, dtype=torch.float64, device="cuda")
    hf[0] = 1.0
    psi = hf.clone()
    lam = torch.empty_like(psi)
    e, g, d2, nrm = F.vqe_energy_grad(idx, th, psi, lam)
    print(f"vqe 16 orbitals: energy {e:.12f}", flush=True)
    # the reverse sweep uncomputes the state (no stored intermediates)
    assert float((psi - hf).abs().max()) < 1e-12 and abs(nrm - 1.0) < 1e-12
    # one position of each kind: beta single, beta-beta double, alpha single, alpha-beta
    # double, alpha-alpha double (the parameter order of jl/ansatz.jl:67-167)
    kinds = {}
    for i, so in enumerate(sos):
        k = (len(so) // 2, sum(o % 2 == 0 for o in so[:len(so) // 2]))
        kinds.setdefault(k, []).append(i)
    rng = np.random.default_rng(SEED_BASE + 3000)
    sample = sorted(int(rng.choice(v)) for v in kinds.values())
    assert len(sample) == 5
    work = [torch.empty_like(psi) for _ in range(4)]
    fd = fd_gradient(F, idx, th, sample, 1e-6, hf, work)
    dev = np.abs(fd - g[sample])
    print(f"vqe 16 orbitals: fd max dev {dev.max():.3e} at {sample}", flush=True)
    assert dev.max() < 1e-8
    # energy-only entry point agrees with the fused one
    psi.copy_(hf)
    e1, d21, nrm1 = F.vqe_energy(idx, th, psi, lam)
    assert abs(e1 - e) < 1e-12 and abs(d21 - d2) < 1e-9 and abs(nrm1 - 1.0) < 1e-12
    # the sweep-only form (no Hamiltonian) leaves the same state U(theta)|HF>
    psi2 = hf.clone()
    F.vqe_sweep(idx, th, psi2, lam)
    assert float((psi2 - psi).abs().max()) < 1e-14

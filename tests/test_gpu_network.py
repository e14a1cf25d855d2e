"""The generic Pauli-group network path (csrc/network.hip) against the oracle.

Networks that are not spin-free two-body Hamiltonians -- the Heisenberg and transverse-field
Ising chains of jl/hamiltonian.jl:438-484 on the 2n spin-orbital qubits, H + S^2 and H +
(N - n)^2 + S^2 + S_z (apply_constraint, jl/hamiltonian.jl:429-435), and a hopping term without
its Jordan-Wigner string -- go through build_networkdev_f64 as their Pauli groups
(hipotf_network_kind 3).  hvec_cuda and get_diags_elements_cuda on them must equal the oracle's
network sigma and diagonal (oracle/oracle.c, pinned against dense matrices in
tests/test_oracle_spin.py) within 1e-12; the integral-form entry points refuse the handle.
"""
import numpy as np
import pytest

import binsim_amd.lib as L
from binsim_amd import davidson as D
from binsim_amd import otf
from oracle import glue
from oracle import oracle as O
from tests.conftest import gpu_vec, host
from tests.helpers import integrals, rel_err, seeded_vector

pytestmark = pytest.mark.gpu


def _ham_terms(n):
    _, _, h1, g2 = integrals(n)
    return [(int(glue.unzip_even(int(ax)) | glue.unzip_odd(int(bx))),
             int(glue.unzip_even(int(az)) | glue.unzip_odd(int(bz))), float(c))
            for ax, bx, az, bz, c in zip(*glue.jw_hamiltonian_terms(h1, g2, n))]


def _terms(kind, n, na, nb):
    if kind == "heisenberg":
        return glue.heisenberg_terms(2 * n, 0.8, 1.1, -0.6, pbc=True)
    if kind == "ising":
        return glue.ising_terms(2 * n, 1.0, 0.4)
    if kind == "s2_penalty":
        return glue.apply_constraint(_ham_terms(n), n, (na, nb), (0.0, 0.5, 0.0))
    if kind == "constraint":
        return glue.apply_constraint(_ham_terms(n), n, (na, nb), (0.7, 0.5, -0.3))
    if kind == "no_jw_string":
        return glue.add_terms(_ham_terms(n), [((1 << 0) | (1 << 4), 0, 0.3)])
    raise ValueError(kind)


KINDS = ["heisenberg", "ising", "s2_penalty", "constraint", "no_jw_string"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n,na,nb", [(4, 2, 2), (6, 3, 3), (6, 2, 4), (8, 4, 3)])
def test_generic_network_hvec_and_diag(torch_cuda, kind, n, na, nb):
    print(f"generic network {kind} n={n} na={na} nb={nb}", flush=True)
    terms = _terms(kind, n, na, nb)
    net = O.Network(glue.compress_by_svd(glue.term_arrays(terms)))
    S = O.System(n, na, nb, ham=net)
    H = otf.HipOTF.from_network(n, net.flat)
    assert L.load().hipotf_network_kind(H.ptr) == 3
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    F = otf.HipOTF_Functions(B, ham=H)
    x = seeded_vector(S.dim, 21 + n)
    ref = S.hvec(x)
    y = torch_cuda.empty(S.dim, dtype=torch_cuda.float64, device="cuda")
    F.hvec(gpu_vec(torch_cuda, x), y)
    assert rel_err(host(y), ref) < 1e-12
    d = gpu_vec(torch_cuda, np.full(S.dim, 0.5))  # accumulates
    F.get_diags(d)
    dref = S.diags()
    assert np.max(np.abs(host(d) - 0.5 - dref)) < 1e-12 * max(1.0, np.max(np.abs(dref)))


def test_generic_network_refused_by_integral_entry_points(torch_cuda):
    n, na, nb = 4, 2, 2
    net = O.Network(glue.compress_by_svd(glue.term_arrays(_terms("heisenberg", n, na, nb))))
    H = otf.HipOTF.from_network(n, net.flat)
    B = otf.HipBasisManager(otf.BasisManager(n, na, nb))
    with pytest.raises(L.HipOTFError, match="group by group"):
        D.solve(B, H, nroots=1)

"""VQE checkpoint / resume (jl/vqe.jl:122-129 load_x, :170-174 x written on every objective call)
and the optimisation loop around the objective, on the CPU with the oracle's energy_objective
(oracle/oracle.py, jl/vqe.jl:1-21) as the objective: a run killed part-way resumes from the
checkpoint and reaches the same minimum as an uninterrupted run."""
import json
import os

import numpy as np
import pytest

from binsim_amd import vqe as V
from binsim_amd.ansatz import ucc_spin_orbitals
from oracle import oracle as O
from tests.helpers import oracle_system


def test_save_load_roundtrip(tmp_path):
    p = str(tmp_path / "x.bin")
    x = np.random.default_rng(1).uniform(-1, 1, 37)
    V.save_x(p, x, calls=5)
    assert np.array_equal(V.load_x(p), x)
    with open(p + ".json") as f:
        side = json.load(f)
    assert side == {"calls": 5, "n": 37, "dtype": "float64"}
    assert not os.path.exists(p + ".tmp")
    V.save_x(p, x[:10])  # overwritten whole
    assert np.array_equal(V.load_x(p), x[:10])
    with open(p + ".json", "w") as f:
        json.dump({"n": 11}, f)
    with pytest.raises(ValueError):
        V.load_x(p)
    with pytest.raises(FileNotFoundError):
        V.load_x(str(tmp_path / "missing.bin"))


class _Killed(Exception):
    pass


def _objective(S, idxs, seen, kill_at=None):
    def fg(x):
        seen.append(np.array(x))
        if kill_at is not None and len(seen) == kill_at:
            raise _Killed()
        e, g, _ = O.energy_objective(S, idxs, x, O.hf_vector(S), np.zeros(S.dim))
        return e, g
    return fg


def test_killed_run_resumes_from_checkpoint(tmp_path):
    n, na, nb = 4, 2, 2
    S = oracle_system(n, na, nb, pool=True)
    idxs = np.arange(len(ucc_spin_orbitals(n, na, nb)))
    x0 = np.zeros(len(idxs))
    e0 = O.energy_objective(S, idxs, x0, O.hf_vector(S), np.zeros(S.dim))[0]
    e_full, x_full, info = V.minimize_fg(_objective(S, idxs, []), x0, gtol=1e-9, ftol=1e-15)
    assert info["success"] and e_full < e0 - 1e-3
    # the exact ground state bounds it from below (the pool spans the 2+2 electron space here)
    h = np.array([S.hvec(np.eye(S.dim)[i]) for i in range(S.dim)])
    assert e_full >= np.linalg.eigvalsh(0.5 * (h + h.T))[0] - 1e-9
    p = str(tmp_path / "vqe_x.bin")
    seen = []
    with pytest.raises(_Killed):
        V.minimize_fg(_objective(S, idxs, seen, kill_at=6), x0, save_path=p, gtol=1e-9, ftol=1e-15)
    xk = V.load_x(p)
    assert np.array_equal(xk, seen[-1])  # saved before the call that was interrupted
    e_res, x_res, info = V.minimize_fg(_objective(S, idxs, []), xk, save_path=p, gtol=1e-9, ftol=1e-15)
    assert info["success"] and abs(e_res - e_full) < 1e-9
    assert V.load_x(p).shape == x_res.shape

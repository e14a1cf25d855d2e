"""Autotuner for the sigma kernels' launch choices and the block Davidson block size.

The reference fixes its tiles at compile time (``BATCH_SIZE1/2/3``, ``RANK3`` in
src/include/ham/otf.hpp:6-9; ``TILE_A/B``, ``BATCH_SIZE_SH*`` in src/include/cuda/common.cuh:7-12).
libhipotf compiles every variant in and picks one per call from a process-wide table per
precision (``hipotf_set_tuning``); this module measures the candidates on the caller's own
basis and Hamiltonian and keeps the choice in a JSON file keyed by device and system.

Tuned per precision (FP64 sigma of one vector; FP32 block sigma of ``k`` vectors):

* ``ab_kernel``: alpha-beta kernel family -- 1 source-lane (one beta source string per lane,
  destination alpha row in LDS), 2 register-descriptor (whole source row in LDS), 3 chunked,
  4 source-lane with FP32 vector pairs on the class-split form (the beta strings split by the
  top orbital; one class's FP64 destination elements of both vectors in LDS), 5 matrix-core
  (per destination alpha row, the dense product of the integral block with the source rows on
  the MFMA units, scattered into an FP64 LDS row; up to 16 orbitals, one vector per launch);
* ``ab_split``: lanes per source string in the source-lane kernel, i.e. waves per CU
  (1: 512-thread workgroups, 8 waves per CU; 2: 1024-thread workgroups, 16 waves per CU);
* ``ab_db``: one or two LDS copies of the replicated integral row (source-lane);
* ``ab_nvs``: vectors per source-lane launch in FP32 (1 or 2);
* ``rs_block``: destination alpha (beta) strings per workgroup of the same-spin row SpMM that
  share each gathered source row (4, 8, 16; 0 = one destination row per wave);
* ``rs_passes``: the same-spin row SpMM gathers its sources in 1, 2, 4 or 8 alpha-string
  ranges (source alpha block of N/passes strings per launch; the library's default, 0, takes
  the fewest that keep a column tile's gathered slice within an XCD's 4 MB L2);

and, for the mixed-precision solver, the Davidson block size ``k`` (Ritz pairs refined per
iteration) by the time to a probe tolerance.  Every candidate's sigma is compared against the
first candidate's on the same input; one that differs by more than summation-order noise is
rejected, so a choice never changes results.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import lib as L

FIELDS = ("ab_kernel", "ab_split", "ab_db", "ab_nvs", "rs_passes", "deterministic", "rs_block",
          "sweep_mode")
FP64, FP32 = 0, 1
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

_RANGES = {"ab_kernel": (0, 1, 2, 3, 4, 5), "ab_split": (1, 2), "ab_db": (1, 2), "ab_nvs": (1, 2),
           "rs_passes": (0, 1, 2, 4, 8), "deterministic": (0, 1), "rs_block": (0, 4, 8, 16),
           "sweep_mode": tuple(range(8))}


def validate(cfg: dict) -> dict:
    """Return ``cfg`` completed with defaults; raise ValueError on an unknown or bad field."""
    out = dict(zip(FIELDS, (0, 1, 1, 2, 0, 0, 8, 4)))
    for k, v in cfg.items():
        if k not in _RANGES:
            raise ValueError(f"unknown tuning field {k!r}")
        if int(v) not in _RANGES[k]:
            raise ValueError(f"tuning field {k}={v} not in {_RANGES[k]}")
        out[k] = int(v)
    return out


def get_config(precision: int) -> dict:
    """The library's current launch choices for ``precision`` (0 FP64, 1 FP32)."""
    v = np.zeros(len(FIELDS), dtype=np.int32)
    L.check(L.load().hipotf_get_tuning(precision, v.ctypes.data_as(L._i32p), len(FIELDS)) == 0,
            "hipotf_get_tuning")
    return dict(zip(FIELDS, v.tolist()))


def set_config(precision: int, cfg: dict) -> None:
    """Install ``cfg`` (completed with defaults) for ``precision`` in the library."""
    c = validate(cfg)
    v = np.array([c[f] for f in FIELDS], dtype=np.int32)
    L.check(L.load().hipotf_set_tuning(precision, v.ctypes.data_as(L._i32p), len(FIELDS)) == 0,
            "hipotf_set_tuning")


def system_key(basis) -> str:
    """``n{norb}_a{na}_b{nb}`` of a device basis (HipBasisManager) or host BasisManager."""
    hb = getattr(basis, "host", basis)
    na, nb = hb.nelec
    return f"n{hb.norb}_a{na}_b{nb}"


def device_key() -> str:
    import torch

    p = torch.cuda.get_device_properties(0)
    arch = getattr(p, "gcnArchName", "") or ""
    return f"{p.name}|{arch.split(':')[0]}|cu{p.multi_processor_count}"


def candidates(precision: int, k: int) -> list[dict]:
    """Alpha-beta candidates: the source-lane variants, then register and chunked kernels."""
    out = []
    nvs = (1, 2) if precision == FP32 and k >= 2 else (2,)
    for v in nvs:
        for s in (1, 2):
            for d in (1, 2):
                out.append(validate(dict(ab_kernel=1, ab_split=s, ab_db=d, ab_nvs=v)))
    if precision == FP32 and k >= 2:
        out.append(validate(dict(ab_kernel=4)))
    out.append(validate(dict(ab_kernel=5)))
    out.append(validate(dict(ab_kernel=2)))
    out.append(validate(dict(ab_kernel=3)))
    return out


def _time_sigma(F, xs, ys, precision, reps, warmup) -> float:
    """Mean milliseconds of one sigma call over ``reps`` (HIP events on the library's stream)."""
    import torch

    call = (lambda: F.hvec(xs[0], ys[0])) if (precision == FP64 and len(xs) == 1) else \
        (lambda: F.hvec_block(xs, ys, precision))
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    # the library launches on the null stream, which is torch's default current stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _max_rel(ys, ref) -> float:
    return max(float((a - b).abs().max() / b.abs().max().clamp_min(1e-300)) for a, b in zip(ys, ref))


def tune_sigma(F, basis, precision: int, k: int = 1, reps: int = 3, warmup: int = 1,
               seed: int = 9001, log=None) -> dict:
    """Time every candidate for ``precision`` on ``k`` seeded vectors and install the fastest.

    The alpha-beta candidates are timed first (same-spin part at its defaults), then every
    same-spin row block size and pass count with the best alpha-beta choice.

    Returns ``{"config": ..., "ms": ..., "timings": [(config, ms), ...], "rejected": [...]}``.
    The tolerance for agreeing with the first candidate is 1e-12 relative in FP64 and 1e-5 in
    FP32 (different accumulation orders of the same terms).
    """
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    xs = [torch.randn(basis.dim, dtype=torch.float64, device="cuda", generator=g) for _ in range(k)]
    for x in xs:
        x /= x.norm()
    ys = [torch.empty_like(x) for x in xs]
    tol = 1e-12 if precision == FP64 else 1e-5
    saved = get_config(precision)
    ref, timings, rejected = None, [], []
    try:
        for cfg in candidates(precision, k):
            c = dict(cfg, rs_passes=0, rs_block=8, deterministic=saved["deterministic"])
            set_config(precision, c)
            ms = _time_sigma(F, xs, ys, precision, reps, warmup)
            if ref is None:
                ref = [y.clone() for y in ys]
            elif _max_rel(ys, ref) > tol:
                rejected.append(c)
                continue
            timings.append((c, ms))
            if log:
                log(f"  prec {precision} k {k} {c}: {ms:.3f} ms")
        best = min(timings, key=lambda t: t[1])[0]
        for rb, p in [(rb, p) for rb in (0, 4, 8, 16) for p in (1, 2, 4, 8)]:
            c = dict(best, rs_block=rb, rs_passes=p)
            set_config(precision, c)
            ms = _time_sigma(F, xs, ys, precision, reps, warmup)
            if _max_rel(ys, ref) > tol:
                rejected.append(c)
                continue
            timings.append((c, ms))
            if log:
                log(f"  prec {precision} k {k} {c}: {ms:.3f} ms")
        best, best_ms = min(timings, key=lambda t: t[1])
    except BaseException:
        set_config(precision, saved)
        raise
    set_config(precision, best)
    del xs, ys, ref
    return {"config": best, "ms": best_ms, "timings": timings, "rejected": rejected}


def tune_block(basis, ham, nroots: int = 1, blocks=(1, 2, 4, 8), tol: float = 1e-4,
               maxiter: int = 200, log=None) -> dict:
    """Davidson block size: wall seconds of the mixed solver to ``tol`` for each block size.

    ``tol`` stays above the FP32 switch threshold's refinement regime so that the probe is
    dominated by the block phase the choice affects.  Block sizes below ``nroots`` are skipped.
    """
    from . import davidson as D

    times = []
    for b in blocks:
        if b < nroots:
            continue
        t = time.perf_counter()
        _, _, info = D.solve(basis, ham, nroots=nroots, mixed=True, tol=tol, maxiter=maxiter, block=b)
        wall = time.perf_counter() - t
        ok = bool(info["converged"])
        times.append((b, wall if ok else float("inf"), int(info["iterations"])))
        if log:
            log(f"  block {b}: {wall:.3f} s, {int(info['iterations'])} iterations, converged {ok}")
    best = min(times, key=lambda t: t[1])
    return {"block": best[0], "seconds": best[1], "timings": times}


def load_file(path: str = DEFAULT_PATH) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    if not isinstance(data, dict):
        raise ValueError(f"{path}: expected a JSON object")
    return data


def save_file(data: dict, path: str = DEFAULT_PATH) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)


def lookup(basis, path: str = DEFAULT_PATH) -> dict | None:
    """The stored entry for this device and system, or None.

    The device part of the key is ``name|arch|cu<count>``; the marketing name can differ
    between driver installs of the same part, so an entry stored under another name with the
    same architecture and CU count is accepted.
    """
    data = load_file(path)
    key, sysk = device_key(), system_key(basis)
    if sysk in data.get(key, {}):
        return data[key][sysk]
    tail = key.split("|", 1)[-1]
    for k, v in data.items():
        if k.split("|", 1)[-1] == tail and sysk in v:
            return v[sysk]
    return None


def apply(entry: dict) -> None:
    """Install a stored entry's FP64 and FP32 launch choices."""
    set_config(FP64, entry["fp64"]["config"])
    set_config(FP32, entry["fp32"]["config"])


def tune(F, basis, ham=None, k: int = 4, nroots: int = 1, block_probe: bool = True,
         path: str | None = DEFAULT_PATH, reps: int = 3, log=None) -> dict:
    """Tune FP64 (one vector) and FP32 (``k`` vectors) sigma, optionally the Davidson block
    size, install the result and store it under (device, system) in ``path`` (None: no file)."""
    entry = {"fp64": tune_sigma(F, basis, FP64, 1, reps=reps, log=log),
             "fp32": tune_sigma(F, basis, FP32, k, reps=reps, log=log)}
    if block_probe and ham is not None:
        entry["davidson"] = tune_block(basis, ham, nroots=nroots, log=log)
    entry["fp32"]["k"] = k
    entry["when"] = time.strftime("%Y-%m-%dT%H:%M:%S")
    if path:
        data = load_file(path)
        data.setdefault(device_key(), {})[system_key(basis)] = _jsonable(entry)
        save_file(data, path)
    return entry


def load_or_tune(F, basis, ham=None, path: str = DEFAULT_PATH, **kw) -> tuple[dict, bool]:
    """Apply the stored entry if there is one, else tune (and store).  Returns (entry, tuned)."""
    e = lookup(basis, path)
    if e is not None:
        apply(e)
        return e, False
    return tune(F, basis, ham, path=path, **kw), True


def _jsonable(o):
    if isinstance(o, dict):
        return {k: _jsonable(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_jsonable(v) for v in o]
    if isinstance(o, float) and not np.isfinite(o):
        return None
    return o


def main(argv=None):
    """``python -m binsim_amd.autotune --norb 16 --na 8 --nb 8``: tune the benchmark system built
    from the seeded integrals of that orbital count and store the result."""
    import argparse
    import sys

    from . import otf
    from .integrals import for_norb, one_body_layout, reference_layout

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--norb", type=int, default=16)
    ap.add_argument("--na", type=int, default=8)
    ap.add_argument("--nb", type=int, default=8)
    ap.add_argument("--k", type=int, default=4, help="vectors per FP32 block sigma")
    ap.add_argument("--nroots", type=int, default=1)
    ap.add_argument("--no-block-probe", action="store_true")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--path", default=DEFAULT_PATH)
    a = ap.parse_args(argv)
    import torch

    # torch's HIP runtime is initialised before the library's first device call, as the
    # bench and tests do
    assert torch.cuda.is_available(), "no GPU visible"

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    h, eri = for_norb(a.norb)
    basis = otf.HipBasisManager(otf.BasisManager(a.norb, a.na, a.nb))
    ham = otf.HipOTF.from_integrals(a.norb, one_body_layout(h), reference_layout(eri))
    F = otf.HipOTF_Functions(basis, ham)
    log(f"tuning {system_key(basis)} on {device_key()}")
    e = tune(F, basis, ham, k=a.k, nroots=a.nroots, block_probe=not a.no_block_probe, path=a.path,
             reps=a.reps, log=log)
    print(json.dumps({"system": system_key(basis), "fp64": e["fp64"]["config"], "fp64_ms": e["fp64"]["ms"],
                      "fp32": e["fp32"]["config"], "fp32_ms": e["fp32"]["ms"], "k": a.k,
                      "davidson_block": e.get("davidson", {}).get("block")}))


if __name__ == "__main__":
    main()

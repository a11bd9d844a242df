"""Launch-shape autotuner of the mixed-precision multi-case engine.

The stage kernel's launch shape (one workgroup per block or a persistent grid of so many per CU,
XCD-aware block order,
non-temporal loads for read-once arrays) changes speed only: every shape gives bitwise the same
results. The tuner admits the workload's cases into a throwaway engine, times a few iterations per
candidate shape and keeps the fastest. Results are cached per device and workload class in a JSON
file (E3D_TUNE_CACHE, default ~/.cache/e3d_mi355x/tuning.json).
"""
from __future__ import annotations

import json
import math
import os
import time
from pathlib import Path

from .solver import Batch

# wg_per_cu 0: one workgroup per (case, block) pair; else a persistent grid of that many per CU
CANDIDATES = tuple((w, x, nt) for w in (0, 2, 3, 4, 5, 6, 8) for x in (8, 1) for nt in (True, False))


def cache_path() -> Path:
    return Path(os.environ.get("E3D_TUNE_CACHE", Path.home() / ".cache" / "e3d_mi355x" / "tuning.json"))


def workload_key(n_cells_total: int, schemes: set[int], device: str = "gfx950") -> str:
    """Workload class: device, total resident cells (power-of-two bucket), flux schemes present."""
    bucket = 1 << max(0, int(math.log2(max(1, n_cells_total))))
    return f"{device}/cells>={bucket}/schemes={''.join(str(s) for s in sorted(schemes))}"


def load_cache() -> dict:
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def store_cache(key: str, entry: dict) -> None:
    p = cache_path()
    data = load_cache()
    data[key] = entry
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(data, indent=1, sort_keys=True))
    except OSError:
        pass


def lookup(key: str):
    e = load_cache().get(key)
    return (e["wg_per_cu"], e["xcds"], e["nontemporal"]) if e else None


def autotune(cases, iters: int = 5, warmup: int = 2, candidates=CANDIDATES, device: int = 0,
             use_cache: bool = True) -> dict:
    """cases: list of (Mesh, Params). Returns {"best": (wg_per_cu, xcds, nontemporal), "table": [...],
    "key": ...}; the best shape is cached under the workload key."""
    n_total = sum(m.info()["n_elem"] for m, _ in cases)
    key = workload_key(n_total, {int(p.scheme) for _, p in cases})
    if use_cache:
        hit = lookup(key)
        if hit is not None:
            return {"best": tuple(hit), "table": [], "key": key, "cached": True}
    table = []
    with Batch(device=device, max_cases=max(1, len(cases))) as b:
        for m, p in cases:
            b.add(m, p, drop=0.0, max_iter=1 << 30, hist_cap=1)
        for w, x, nt in candidates:
            b.set_tuning(w, x, nt)
            b.run(warmup)
            b.wait()
            t0 = time.perf_counter()
            b.run(iters)
            b.wait()
            table.append({"wg_per_cu": w, "xcds": x, "nontemporal": nt,
                          "ms_per_iteration": (time.perf_counter() - t0) * 1e3 / iters})
    best = min(table, key=lambda r: r["ms_per_iteration"])
    store_cache(key, dict(best, n_cells=n_total))
    return {"best": (best["wg_per_cu"], best["xcds"], best["nontemporal"]), "table": table, "key": key,
            "cached": False}

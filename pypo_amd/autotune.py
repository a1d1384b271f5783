"""Autotuner for the sweep launches.

Times every candidate launch shape on the actual sweep, keeps the fastest,
and caches the choice with the whole timing table in a JSON file keyed by
the device and a signature of the sweep (instance count, pair counts).
Dimensions: targets per lane (TPL: the far field's target block is 256 TPL
targets per workgroup, its matrix-core kernel runs 4 TPL M-blocks of 16
targets per wave; TPL 4 is the 16-M-block kernel at one wave per SIMD),
256-source tiles per work item, and workgroups per CU in each launch.  Each
candidate is timed as the median of `reps` individually synchronised
replays after one warm run.  Cache path: $PYPO_AMD_TUNE_CACHE or
~/.cache/pypo_amd/autotune.json.
"""

import hashlib
import json
import os
import time

CANDIDATES = [
    {"tpl": tpl, "chunk_tiles": ch, "wgs_per_cu": w}
    for tpl in (1, 2)
    for ch in (64, 160, 4096)
    for w in (4, 16, 64)
] + [{"tpl": 4, "chunk_tiles": ch, "wgs_per_cu": w} for ch in (160, 4096) for w in (2, 64)]


def cache_path():
    return os.environ.get("PYPO_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pypo_amd", "autotune.json"))


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return "unknown"


def signature(sweep):
    info = sweep.info()
    key = f"{_device_name()}|{sweep.n_inst}|{info['pairs_jm']:.6e}|{info['pairs_ff']:.6e}"
    return hashlib.sha1(key.encode()).hexdigest()[:16], key


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass   # a read-only home is not an error: the tuning is just not remembered


def _time_run(sweep, reps=3):
    """Median, min and max seconds of `reps` individually synchronised runs."""
    import torch
    sweep.run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        sweep.run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts[len(ts) // 2], ts[0], ts[-1]


def describe(params):
    """The launch shape of a candidate, with the quantities it fixes."""
    d = dict(params)
    d["m_blocks"] = 4 * int(params["tpl"])
    d["targets_per_workgroup"] = 256 * int(params["tpl"])
    d["lds_tile_sources"] = 256
    return d


def tune(sweep, candidates=None, use_cache=True, use_graph=True, log=None):
    """Pick the fastest launch shape for `sweep`, apply it (with graph capture
    when use_graph) and return (params, seconds per run, table)."""
    sig, key = signature(sweep)
    db = _load() if use_cache else {}
    if use_cache and sig in db:
        best = db[sig]["params"]
        sweep.plan(use_graph=use_graph, **best)
        return best, db[sig]["seconds"], db[sig].get("table", [])
    table = []
    for c in candidates or CANDIDATES:
        sweep.plan(use_graph=False, **c)
        dt, lo, hi = _time_run(sweep)
        table.append({"params": c, "seconds": dt, "min_s": lo, "max_s": hi})
        if log:
            log(f"autotune {c}: {dt * 1e3:.2f} ms (range {lo * 1e3:.2f}..{hi * 1e3:.2f})")
    best = min(table, key=lambda r: r["seconds"])
    sweep.plan(use_graph=use_graph, **best["params"])
    if use_cache:
        db[sig] = {"key": key, "params": best["params"], "seconds": best["seconds"], "table": table}
        _store(db)
    return best["params"], best["seconds"], table

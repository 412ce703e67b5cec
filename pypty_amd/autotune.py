"""Launch-parameter autotuner for the pass engine.

For a problem shape the tuner times one full pass per candidate windows-per-launch
(``compute_batch``) after a warm-up pass, keeps the fastest, and remembers the choice per
(device name, window, slices, object modes, probe modes, positions) for the life of the
process, and across processes when ``PYPTY_AMD_TUNE_CACHE`` names a JSON file.  The
candidates only change how the positions are grouped into launches, so any choice gives
the same loss and gradients up to the order of the FP64 loss partial sums.
"""
from __future__ import annotations

import json
import os

import torch

_cache: dict = {}
_loaded = False


def _cache_path():
    return os.environ.get("PYPTY_AMD_TUNE_CACHE")


def _load():
    global _loaded
    if _loaded:
        return
    _loaded = True
    path = _cache_path()
    if path and os.path.exists(path):
        with open(path) as fh:
            _cache.update(json.load(fh))


def _save():
    path = _cache_path()
    if path:
        with open(path, "w") as fh:
            json.dump(_cache, fh, indent=1, sort_keys=True)


def shape_key(device, n, slices, obj_modes, probe_modes, positions) -> str:
    name = torch.cuda.get_device_name(device)
    return f"{name}|n{n}|s{slices}|k{obj_modes}|m{probe_modes}|p{positions}"


def choose_batch(key: str, candidates, make_pass, reps: int = 2, tie: float = 0.02):
    """Return the fastest candidate for ``key``; a larger candidate within ``tie`` (relative) of
    the fastest time wins, so that timing noise between near-equal candidates does not change
    the choice from run to run (fewer, larger launches).  ``make_pass(batch)`` returns a callable
    that launches one full pass on the current stream; the timing is taken with HIP events."""
    _load()
    if key in _cache:
        return int(_cache[key])
    times = {}
    for b in candidates:
        run = make_pass(int(b))
        run()  # warm-up: workspace and tables
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            run()
        stop.record()
        stop.synchronize()
        times[int(b)] = start.elapsed_time(stop) / reps
    fastest = min(times.values())
    best = max(b for b, ms in times.items() if ms <= fastest * (1.0 + tie))
    _cache[key] = best
    _save()
    return best

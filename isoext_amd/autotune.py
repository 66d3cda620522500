"""Tile-shape selection for the marching-cubes emit kernel.

The library compiles a handful of emit tile configurations (TX x TY cell
rows x NZC 64-cell chunks per workgroup; NZC is also the z-segment length of
the count/scan pass).  ``mc_config(shape)`` returns the configuration to use
for a grid shape: a measured choice from the tuning table if one exists for
the shape's size class, else a static rule.  ``tune(shape)`` times every
configuration on the current GPU with HIP events and records the fastest in
``autotune.json`` next to this file.
"""

from __future__ import annotations

import json
import threading
from pathlib import Path

_HERE = Path(__file__).resolve().parent
TABLE = _HERE / "autotune.json"
_lock = threading.Lock()
_table = None
_by_shape: dict = {}   # shape -> configuration (cleared when the table changes)

# (TX, TY, NZC) of each compiled configuration, in library order
# (configuration 6 is the read-once pipeline: k_stream regions + 16 x 16 face tiles)
CONFIGS = [(16, 16, 4), (8, 8, 2), (16, 16, 2), (16, 16, 1), (16, 16, 8), (8, 8, 1), (16, 16, 1)]


def size_class(shape) -> str:
    X, Y, Z = shape
    n = X * Y * Z
    zc = "z64" if Z <= 65 else ("z128" if Z <= 129 else "zbig")
    if n <= 1 << 18:
        nc = "tiny"
    elif n <= 1 << 21:
        nc = "small"
    elif n <= 1 << 24:
        nc = "medium"
    elif n <= 1 << 27:
        nc = "large"
    else:
        nc = "huge"
    return f"{nc}-{zc}"


def _load():
    global _table
    if _table is None:
        try:
            _table = json.loads(TABLE.read_text())
        except (OSError, ValueError):
            _table = {}
    return _table


def mc_config(shape) -> int:
    key = tuple(int(s) for s in shape)
    cfg = _by_shape.get(key)
    if cfg is not None:
        return cfg
    with _lock:
        t = _load()
        hit = t.get(size_class(key))
    if hit is not None:
        cfg = int(hit)
    else:
        from . import _lib
        cfg = int(_lib.lib().isoext_mc_auto_config(_lib.i64x3(key)))
    if len(_by_shape) > 4096:
        _by_shape.clear()
    _by_shape[key] = cfg
    return cfg


def tune(shape, values=None, level=0.0, reps=5, write=True):
    """Time each configuration on `values` (or a synthetic sphere field)."""
    import torch

    from .mc import mc_dense_raw

    if values is None:
        X, Y, Z = shape
        axes = [torch.linspace(-1, 1, n, device="cuda") for n in shape]
        gx, gy, gz = torch.meshgrid(*axes, indexing="ij")
        values = (gx * gx + gy * gy + gz * gz).sqrt() - 0.6
    results = {}
    for cfg in range(len(CONFIGS)):
        mc_dense_raw(values, shape, (-1, -1, -1), (1, 1, 1), level, 0, cfg)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        times = []
        for _ in range(reps):
            t0.record()
            mc_dense_raw(values, shape, (-1, -1, -1), (1, 1, 1), level, 0, cfg)
            t1.record()
            t1.synchronize()
            times.append(t0.elapsed_time(t1))
        results[cfg] = sorted(times)[len(times) // 2]
    best = min(results, key=results.get)
    if write:
        with _lock:
            t = _load()
            t[size_class(shape)] = best
            TABLE.write_text(json.dumps(t, indent=1, sort_keys=True) + "\n")
            _by_shape.clear()
    return best, results

"""Launch-shape autotuner for the PDHG kernels.

Picks rows-per-wavefront (rows of S a wave walks per launch), tile width
(64 * LPs-per-lane consecutive LPs per wave) and the hipGraph chunk length by
timing the engine's own iterations with hipEvents on the loaded problem.
Results are cached in a JSON file keyed by problem shape, batch, precision and
device name.  The reference has no autotuner (its launch shape is fixed at 4
rows per 128-thread block, src/gpu_pdhg_workspace.py:313,316, and chunk 64,
:67).
"""

from __future__ import annotations

import json
import os
from typing import Optional

import torch

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "lpe", "autotune.json")
SHAPES = ((1, 1), (2, 1), (4, 1), (8, 1), (1, 2), (2, 2), (4, 2), (1, 4), (2, 4))
CHUNKS = (16, 32, 64)


def _key(engine) -> str:
    name = torch.cuda.get_device_name(engine.device).replace(" ", "_")
    return f"{engine.m}x{engine.n}x{engine.A.nnz}_B{engine.batch}_{engine.precision}_{engine.method}_{name}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(engine, *, cache_path: Optional[str] = DEFAULT_CACHE, iterations: int = 128,
             shapes=SHAPES, chunks=CHUNKS, force: bool = False) -> dict:
    """Choose and apply (rows_per_wave, lps_per_lane, chunk) for ``engine``.

    The engine must have vectors loaded; the iterate is overwritten, so call
    ``engine.reset`` afterwards.  Returns the chosen configuration and the
    timings (ms per iteration) of every candidate.
    """
    key = _key(engine)
    cache = _load(cache_path) if cache_path else {}
    if not force and key in cache:
        best = cache[key]
        engine.set_launch(best["rows_per_wave"], best["lps_per_lane"])
        return dict(best, cached=True)
    engine.reset()
    engine.run(8)                                   # warm up code objects
    timings = {}
    for rpw, lpl in shapes:
        engine.set_launch(rpw, lpl)
        engine.reset()
        engine.time_iterations(16, chunk=16)        # graph capture outside the timing
        ms = engine.time_iterations(iterations, chunk=64) / iterations
        timings[f"{rpw}x{lpl}"] = ms
    best_shape = min(timings, key=timings.get)
    rpw, lpl = (int(v) for v in best_shape.split("x"))
    engine.set_launch(rpw, lpl)
    chunk_ms = {}
    for k in chunks:
        engine.reset()
        engine.time_iterations(k, chunk=k)
        chunk_ms[k] = engine.time_iterations(max(iterations, k) // k * k, chunk=k) / (max(iterations, k) // k * k)
    fastest = min(chunk_ms.values())
    # checks cost a few iterations each: prefer the longest chunk within 3% of the fastest
    chunk = max(k for k, v in chunk_ms.items() if v <= 1.03 * fastest)
    engine.reset()
    best = dict(rows_per_wave=rpw, lps_per_lane=lpl, chunk=chunk, ms_per_iteration=timings[best_shape],
                timings=timings, chunk_ms={str(k): v for k, v in chunk_ms.items()})
    if cache_path:
        cache = _load(cache_path)
        cache[key] = best
        try:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
            tmp = cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(cache, f, indent=1, sort_keys=True)
            os.replace(tmp, cache_path)
        except OSError:
            pass
    return dict(best, cached=False)

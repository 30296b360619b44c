"""Launch-shape autotuner for the Euler stage kernels.

For the mesh and numerics at hand it times every kernel variant (face sweep
+ gather, fused per cell, fused per cell tile; for batches also the tile
kernel as one persistent launch per stage) and workgroup size (64 to 512
threads; for tiles, pairs of workgroup size and tile capacity) for a few
steps, and keeps the fastest.  Choices
are cached in a JSON file (FVK_TUNE_CACHE, default
~/.cache/fvk/autotune.json) under a key made of the device name, the mesh
size and largest cell degree, and the numerics that change the kernels, so
a later run on the same kind of case reuses them.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from typing import Dict, Iterable, Optional, Sequence, Tuple

import numpy as np

from . import abi

VARIANTS = (1, 2, 3)
BLOCKS = (64, 128, 256, 512)
# variant 3 (fused tiles): workgroup size and tile capacity.  A tile of
# capacity T holds about 7/8 T cells, its owned and ring-1 cells about
# 1.2 times that and its faces about 1.6 times, so these pairs cover one
# pass per phase (staging, gradients, faces) up to two passes.
TILE_SHAPES = ((128, 112), (192, 160), (256, 192), (256, 200), (256, 208), (256, 216), (256, 224), (256, 240), (256, 256),
               (320, 224), (320, 256), (384, 256), (384, 320), (512, 400), (512, 432), (512, 448), (512, 464),
               (512, 480))


def _cache_path() -> str:
    return os.environ.get("FVK_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "fvk", "autotune.json"))


def _device_name() -> str:
    try:
        with open("/sys/class/kfd/kfd/topology/nodes/1/name") as f:
            return f.read().strip() or "gpu"
    except OSError:
        return "gpu"


def signature(n_cells: int, n_faces: int, max_degree: int, params: abi.ParamsC, batch: bool = False) -> str:
    key = (_device_name(), int(n_cells), int(n_faces), int(max_degree), params.equation, params.flux_scheme,
           params.order, params.limiter, params.rk_stages, params.local_dt, batch)
    return hashlib.sha1(repr(key).encode()).hexdigest()[:16]


def _load() -> Dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db: Dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def _time_solver(make, steps: int, warmup: int) -> float:
    s = make()
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    return (time.perf_counter() - t0) / steps


def _time_solver_device(make, steps: int, warmup: int) -> float:
    """Device seconds per step: one event pair on the solver's stream around
    the launches of one step() call (no host jitter in it)."""
    s = make()
    s.step(warmup)
    s.synchronize()
    s.set_kernel_timing(2)
    s.step(steps)
    tot_ms, _ = s.kernel_times()["stages"]
    s.set_kernel_timing(0)
    return tot_ms / steps / 1e3


def tune(mesh, bcs, params: abi.ParamsC, ic, steps: int = 20, warmup: int = 3,
         variants: Iterable[int] = VARIANTS, blocks: Iterable[int] = BLOCKS, use_cache: bool = True) -> Dict:
    """Fastest launch shape (variant, workgroup size, tile cells) for one
    case, by wall time per step over every shape, then by device time over
    three times the steps for the three fastest; sets it on `params` too."""
    from .solver import FVKSolver
    deg = int(np.diff(mesh.cell_face_ptr).max())
    key = signature(mesh.n_cells, mesh.n_faces, deg, params)
    db = _load() if use_cache else {}
    if key in db:
        best = db[key]
    else:
        from .solver import FVKError
        table, shapes = {}, {}
        if params.equation != abi.EQ_EULER:
            variants = [v for v in variants if v != 2]  # the per-cell fused variant is Euler-only
        for v in variants:
            for b, tc in (TILE_SHAPES if v == 3 else [(b, 0) for b in blocks]):
                p = abi.ParamsC.from_buffer_copy(params)
                p.kernel_variant, p.block_size, p.tile_cells = v, b, tc
                name = f"v{v}_b{b}" + (f"_t{tc}" if v == 3 else "")
                try:
                    table[name] = _time_solver(lambda: FVKSolver(mesh, bcs, p, ic), steps, warmup) * 1e3
                    shapes[name] = (v, b, tc)
                except FVKError:
                    continue  # not built for this case (e.g. a tile too large for LDS)
        # the closest shapes once more, over three times the steps and by
        # device time: shapes within the host's jitter of each other in the
        # first pass are told apart by what the device spends on them
        refined = {}
        for cand in sorted(table, key=table.get)[:3]:
            p = abi.ParamsC.from_buffer_copy(params)
            p.kernel_variant, p.block_size, p.tile_cells = shapes[cand]
            refined[cand] = _time_solver_device(lambda: FVKSolver(mesh, bcs, p, ic), 3 * steps, warmup) * 1e3
        name = min(refined, key=refined.get)
        v, b, tc = shapes[name]
        best = {"variant": v, "block": b, "tile_cells": tc, "ms_per_step": table[name],
                "table_ms_per_step": {k: round(x, 4) for k, x in table.items()},
                "refined_device_ms_per_step": {k: round(x, 4) for k, x in refined.items()}}
        if use_cache:
            db[key] = best
            _store(db)
    params.kernel_variant, params.block_size = best["variant"], best["block"]
    params.tile_cells = best.get("tile_cells", 0)
    return best


BATCH_TILE_SHAPES = ((128, 112), (256, 192), (256, 216), (256, 240), (512, 448))


def tune_batch(meshes: Sequence, bcs: Sequence, params: abi.ParamsC, ics: Sequence, steps: int = 10,
               warmup: int = 2, blocks: Iterable[int] = (128, 256, 512), use_cache: bool = True,
               progress=None) -> Dict:
    """Fastest launch shape for a batch: the fused tile kernel over the
    active cases' tiles (variant 3, a workgroup per tile, SSP-RK stages
    only) and its persistent form (variant 4, the workgroups that fit on the
    device pulling tiles from a queue) at several tile shapes, and the face
    sweep + gather (variant 1) at several workgroup sizes; sets it on
    `params` too.  Variant-1 workgroups larger than the smallest case are
    skipped.  progress(name, ms_per_step), if given, is called after each
    shape is timed."""
    from .batch import FVKBatch
    n = sum(m.n_cells for m in meshes)
    nf = sum(m.n_faces for m in meshes)
    deg = max(int(np.diff(m.cell_face_ptr).max()) for m in meshes)
    key = signature(n, nf, deg, params, batch=True)
    db = _load() if use_cache else {}
    if key in db:
        best = db[key]
    else:
        smallest = min(m.n_cells for m in meshes)
        table, shape = {}, {}
        for b in blocks:
            if b > smallest:
                continue
            p = abi.ParamsC.from_buffer_copy(params)
            p.kernel_variant, p.block_size, p.tile_cells = 1, b, 0
            table[f"v1_b{b}"] = _time_solver(lambda: FVKBatch(meshes, bcs, p, ics), steps, warmup) * 1e3
            shape[f"v1_b{b}"] = (1, b, 0)
            if progress:
                progress(f"v1_b{b}", table[f"v1_b{b}"])
        if params.rk_stages >= 2:
            for v in (3, 4):
                for b, tc in BATCH_TILE_SHAPES:
                    p = abi.ParamsC.from_buffer_copy(params)
                    p.kernel_variant, p.block_size, p.tile_cells = v, b, tc
                    try:
                        table[f"v{v}_b{b}_t{tc}"] = _time_solver(lambda: FVKBatch(meshes, bcs, p, ics), steps,
                                                                 warmup) * 1e3
                    except Exception:
                        continue  # not built for these meshes (LDS)
                    shape[f"v{v}_b{b}_t{tc}"] = (v, b, tc)
                    if progress:
                        progress(f"v{v}_b{b}_t{tc}", table[f"v{v}_b{b}_t{tc}"])
        name = min(table, key=table.get)
        v, b, tc = shape[name]
        best = {"variant": v, "block": b, "tile_cells": tc, "ms_per_step": table[name],
                "table_ms_per_step": {k: round(x, 4) for k, x in table.items()}}
        if use_cache:
            db[key] = best
            _store(db)
    params.kernel_variant, params.block_size = best["variant"], best["block"]
    params.tile_cells = best.get("tile_cells", 0)
    return best

"""Autotuner for the extraction's launch shapes (no counterpart in the reference).

The outputs never depend on the launch shape; only the time does.  The
tunables (mcx_set_tuning):

* columns_per_cu  -- count-kernel columns (one wave each) per CU for a single
  volume, at most: more columns give several rounds of waves per SIMD
  (balance where the surface is uneven) at the price of more slab-boundary
  planes read twice.  Smaller volumes take fewer (about 8192 cubes a column,
  at least 4 per CU), so at 256^3 and below the larger candidates coincide.
* emit_groups_per_cu -- persistent emit workgroups per CU.
* batch_slab -- planes per column of batched plans.

``tune_extract(vol, h)`` times every candidate pair on the caller's volume
with HIP events and keeps the fastest; results are cached per (device name,
volume shape, voxel kind) in memory and, when ``cache_path`` is given, in a
JSON file.
"""
import json
import os

import torch

from . import _capi, mesh

COLUMNS = (16, 32, 48, 56, 64, 96, 128)
EMIT_GROUPS = (4, 6, 8, 16, 32)
BATCH_SLABS = (16, 32, 64)

_cache = {}


def get():
    lib = _capi.load()
    return {"columns_per_cu": lib.mcx_get_tuning(_capi.TUNE_COLUMNS_PER_CU),
            "emit_groups_per_cu": lib.mcx_get_tuning(_capi.TUNE_EMIT_GROUPS_PER_CU),
            "batch_slab": lib.mcx_get_tuning(_capi.TUNE_BATCH_SLAB)}


def set(columns_per_cu=None, emit_groups_per_cu=None, batch_slab=None):  # noqa: A001
    lib = _capi.load()
    for key, val in ((_capi.TUNE_COLUMNS_PER_CU, columns_per_cu),
                     (_capi.TUNE_EMIT_GROUPS_PER_CU, emit_groups_per_cu),
                     (_capi.TUNE_BATCH_SLAB, batch_slab)):
        if val is not None:
            _capi.check(lib.mcx_set_tuning(key, int(val)), "mcx_set_tuning")


def _time(fn, reps):
    torch.cuda.synchronize()
    fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(reps)]
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in evs)
    return ms[len(ms) // 2]


def _key(vol):
    elem, X, Y, Z = mesh.volume_kind(vol)
    return (torch.cuda.get_device_name(vol.device), X, Y, Z, elem)


def tune_extract(vol, h, *, columns=COLUMNS, emit_groups=EMIT_GROUPS, reps=5, cache_path=None,
                 apply=True):
    """Fastest (columns_per_cu, emit_groups_per_cu) for extracting `vol`
    (vertex_voxel output for short2 volumes, positions for fp32); sets it when
    `apply`.  Returns {"columns_per_cu", "emit_groups_per_cu", "ms", "table"}."""
    key = _key(vol)
    skey = "|".join(str(k) for k in key)
    if cache_path and os.path.exists(cache_path) and skey not in _cache:
        try:
            with open(cache_path) as f:
                _cache.update(json.load(f))
        except (OSError, ValueError):
            pass
    if skey in _cache:
        best = _cache[skey]
        if apply:
            set(best["columns_per_cu"], best["emit_groups_per_cu"])
        return best
    ws = mesh.Workspace(vol.device)
    n = mesh.extract_into(vol, h, cap=0, workspace=ws)
    elem = mesh.volume_kind(vol)[0]
    out = {}
    if elem == _capi.ELEM_SHORT2:
        out["vv"] = torch.empty((max(n, 1), mesh.VV_WORDS), dtype=torch.int32, device=vol.device)
    else:
        out["pos"] = torch.empty((max(n, 1), 3), dtype=torch.float32, device=vol.device)
    d_count = torch.zeros(1, dtype=torch.int64, device=vol.device)
    saved = get()
    table = []
    try:
        for c in columns:
            for g in emit_groups:
                set(c, g)

                def run():
                    mesh.extract_into(vol, h, cap=n, d_count=d_count, workspace=ws, sync=False,
                                      **out)
                table.append((c, g, _time(run, reps)))
    finally:
        set(saved["columns_per_cu"], saved["emit_groups_per_cu"])
    c, g, ms = min(table, key=lambda r: r[2])
    best = {"columns_per_cu": c, "emit_groups_per_cu": g, "ms": ms,
            "table": [list(r) for r in table]}
    _cache[skey] = best
    if cache_path:
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(_cache, f, indent=1)
    if apply:
        set(c, g)
    return best


def tune_batch(make_plan, run, *, slabs=BATCH_SLABS, emit_groups=EMIT_GROUPS, reps=3,
               apply=True):
    """Fastest (batch_slab, emit_groups_per_cu) for a batch: make_plan() builds
    a plan (the slab applies at creation), run(plan) launches one batched
    extraction.  Returns {"batch_slab", "emit_groups_per_cu", "ms", "table"}."""
    saved = get()
    table = []
    try:
        for sl in slabs:
            set(batch_slab=sl)
            plan = make_plan()
            for g in emit_groups:
                set(emit_groups_per_cu=g)
                table.append((sl, g, _time(lambda: run(plan), reps)))
            del plan
    finally:
        set(saved["columns_per_cu"], saved["emit_groups_per_cu"], saved["batch_slab"])
    sl, g, ms = min(table, key=lambda r: r[2])
    if apply:
        set(batch_slab=sl, emit_groups_per_cu=g)
    return {"batch_slab": sl, "emit_groups_per_cu": g, "ms": ms, "table": [list(r) for r in table]}

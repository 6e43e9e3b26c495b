"""Launch-shape autotuner for the stencil passes.

A pass is a grid of 64 x tile_y workgroups, each thread walking tile_z planes
of one (x, y) column.  tile_y trades lateral halo re-reads (taller tiles
re-read fewer +-y rows through L2) against workgroups in flight; tile_z
trades z reuse in registers (longer walks re-read fewer k-1 / k+1 planes)
against the number of workgroups available to fill 256 CUs.  ``tune`` times
every candidate with HIP events on a zero-potential grid of the requested
size and returns the fastest; results are cached per (equation, precision,
schedule, nz, ny, nx) in a JSON file (``$WAP_AUTOTUNE_CACHE`` or
``~/.cache/wideanglepropagation_amd/autotune.json``).
"""
import json
import os

import torch

from .engine import FDTD3D

TILE_Y = (1, 2, 4, 8)
TILE_Z = (4, 8, 16, 32, 64)


def _cache_path():
    p = os.environ.get("WAP_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "wideanglepropagation_amd",
                        "autotune.json")


def _key(p):
    return f"{p.equation}/{p.precision}/{p.resolved_schedule()}/{p.nz}x{p.ny}x{p.nx}"


def load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(c):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def measure(params, tile_y, tile_z, steps=5, warmup=2):
    """Mean device ms per step of one launch shape (HIP events on the solver stream)."""
    p = params.with_(tile_y=tile_y, tile_z=tile_z)
    dt = torch.float64 if p.precision == "f64" else torch.float32
    V = torch.zeros(p.shape, dtype=dt, device="cuda")
    sim = FDTD3D(p, V=V)
    try:
        sim.step(warmup)
        _, tot = sim.step_profiled(steps)
    finally:
        sim.close()
    return tot / steps


def tune(params, tile_ys=TILE_Y, tile_zs=TILE_Z, steps=5, warmup=2, use_cache=True,
         verbose=False):
    """Return (tile_y, tile_z, ms_per_step, table) of the fastest candidate."""
    key = _key(params)
    if params.resolved_schedule() in ("kg_fused", "schr_fused") and tile_ys == TILE_Y:
        tile_ys = (4, 8, 16)   # the fused passes need >= 3 rows (one output row + halo)
    cache = load_cache() if use_cache else {}
    if use_cache and key in cache:
        e = cache[key]
        return e["tile_y"], e["tile_z"], e["ms"], e.get("table", [])
    table = []
    for ty in tile_ys:
        for tz in tile_zs:
            if tz > params.nz:
                continue
            ms = measure(params, ty, tz, steps, warmup)
            table.append((ty, tz, ms))
            if verbose:
                print(f"[autotune] {key} tile_y={ty} tile_z={tz}: {ms:.4f} ms/step", flush=True)
    ty, tz, ms = min(table, key=lambda r: r[2])
    if use_cache:
        cache = load_cache()
        cache[key] = {"tile_y": ty, "tile_z": tz, "ms": ms, "table": table}
        save_cache(cache)
    return ty, tz, ms, table


def tuned(params, **kw):
    """params with the tuned launch shape."""
    ty, tz, _, _ = tune(params, **kw)
    return params.with_(tile_y=ty, tile_z=tz)

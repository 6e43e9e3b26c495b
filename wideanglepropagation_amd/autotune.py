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

``tune_batch`` tunes the z chunk per size class, and the workgroup waves per
launch group, for grids that run together in one ``BatchedFDTD``: there the
batch fills the device whatever the chunk, so a single grid's best chunk (short,
to make enough workgroups on its own) is the wrong one (48 uneven grids: 82
Gvoxel-steps/s with single-grid choices against 105 with 32-plane chunks).  It
times the whole batch as it runs (its captured graph), first with one chunk for
every class, then the waves of each group, then class by class.
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
    elif params.resolved_schedule() == "schr_1pass":
        tile_ys = (16,)        # one fixed workgroup shape: only the z chunk is tuned
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


def measure_batch(param_list, steps=4, warmup=2):
    """Device ms per step of one BatchedFDTD over zero-potential grids of param_list, run
    the way a batch runs: its steps captured into one hipGraph (groups on forked streams),
    timed over a replay after a warm-up replay."""
    from .batched import BatchedFDTD
    sims = []
    try:
        for p in param_list:
            dt = torch.float64 if p.precision == "f64" else torch.float32
            sims.append(FDTD3D(p, V=torch.zeros(p.shape, dtype=dt, device="cuda")))
        b = BatchedFDTD(sims)
        try:
            b.capture(steps)
            for _ in range(max(1, -(-warmup // steps))):
                b.replay()
            b.synchronize()
            return b.replay(timed=True) / steps
        finally:
            b.close()
    finally:
        for s_ in sims:
            s_.close()


def _class_of(p):
    return (p.equation, p.precision, p.resolved_schedule(), p.shape)


def _group_of(p):
    return (p.equation, p.precision, p.resolved_schedule())


# workgroup waves of the schedules whose shape is fixed (tile_y does not apply): the
# one-pass Schrodinger step runs 14 row waves and a side wave
FIXED_WAVES = {"schr_1pass": 15}


def tune_batch(param_list, tile_zs=(16, 32, 64), tile_ys=(8, 16), tile_y=16, steps=4, warmup=2,
               verbose=False):
    """Launch shape for grids batched together: the z chunk per size class and the
    workgroup waves per launch group (the grids of one batched launch share them).
    Coordinate descent on the timed batch: one chunk for every class, then the waves of
    each group, then each class's chunk with the others fixed.  Returns
    ({class: (tile_y, tile_z)}, ms per batched step, table of (choice, ms)).  A class is
    (equation, precision, schedule, shape); a group (equation, precision, schedule)."""
    classes = sorted({_class_of(p) for p in param_list}, key=str)
    groups = sorted({_group_of(p) for p in param_list}, key=str)

    def label(choice, waves):
        return {"/".join((c[0], "x".join(map(str, c[3])))): (FIXED_WAVES.get(c[2], waves[c[:3]]), tz)
                for c, tz in choice.items()}

    def run(choice, waves):
        ps = [p.with_(tile_y=waves[_group_of(p)], tile_z=choice[_class_of(p)])
              for p in param_list]
        ms = measure_batch(ps, steps, warmup)
        table.append((label(choice, waves), ms))
        if verbose:
            print(f"[autotune] batch {table[-1][0]}: {ms:.3f} ms/step", flush=True)
        return ms

    table = []
    waves = {g: tile_y for g in groups}
    best, best_ms = None, float("inf")
    for tz in tile_zs:   # one chunk for every class
        choice = {c: tz for c in classes}
        ms = run(choice, waves)
        if ms < best_ms:
            best, best_ms = choice, ms
    for g in groups:     # the waves of each launch group (the one-pass Schrodinger shape is fixed)
        if g[2] == "schr_1pass":
            continue
        for ty in [t for t in tile_ys if t != waves[g]]:
            trial = dict(waves)
            trial[g] = ty
            ms = run(best, trial)
            if ms < best_ms:
                waves, best_ms = trial, ms
    for c in classes:    # then class by class
        for tz in [t for t in tile_zs if t != best[c]]:
            choice = dict(best)
            choice[c] = tz
            ms = run(choice, waves)
            if ms < best_ms:
                best, best_ms = choice, ms
    return {c: (waves[c[:3]], tz) for c, tz in best.items()}, best_ms, table

"""Autotuner of the tiled path's launch shapes.

The tiled update groups particles by tiles of cells (``Params.tile``) and
re-sorts them every ``Params.resort_interval`` steps.  The best choice depends
on the grid, the particle count and the moment layout: larger tiles mean
fewer halo cells per particle but fewer, longer workgroups; a shorter re-sort
interval keeps particles inside their tile's window at the cost of more
re-sorting updates.  ``tune`` times candidate shapes on the GPU (whole graph
replayed periods, wall clock) and ``tuned`` returns parameters with the
winner of a size class filled in, from a JSON table (the one shipped beside
this module, or one given by the caller), tuning on a miss when allowed.

Size class: precision, grid, particle count rounded to a power of two, the
number of moment components (6 or 9), and whether the strain, density and
quasi-2D options are on.

Ensembles (``tune_ensemble`` / ``tuned_ensemble``): inside a batched ensemble
the members' tiles run in one persistent launch with plenty of parallelism
across members, so the best shape of a size class there is not the one it
has alone (larger tiles pay their per-tile setup over more particles).  The
search then runs on the whole ensemble: the shape of each grid class (size
class without the particle count: the members of a grid share the kernels'
shapes) in turn with the others fixed, then the shared re-sort interval; the
entry is keyed by the ensemble's size classes and member count.

Usage::

    python -m stochasticsampling_amd.autotune --grid 64 64 64 16 16 --n 16777216 --out table.json
"""
import argparse
import ctypes as C
import json
import math
import os
import time

from .params import BIN_SORTED, FLAG_QUASI2D, FP32, FP64, make_params

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
TILES = [(4, 4, 4), (2, 4, 4), (4, 4, 2), (4, 4, 8), (4, 8, 4), (2, 2, 4), (8, 4, 4)]
# ensembles also try larger tiles (fewer per member)
ENSEMBLE_TILES = TILES + [(8, 8, 4), (8, 4, 8), (4, 8, 8), (8, 8, 8)]
RESORTS = [8, 16, 32]
LDS_LIMIT = 163840  # tile window + moment kernels of the tiled update (gfx950: 160 KiB per workgroup)


def size_class(p):
    """Key of the size class of parameters ``p``."""
    g = p.grid_size
    nm = 9 if p.stress_magnetic != 0.0 else 6
    n2 = int(round(math.log2(max(1, p.number_of_particles))))
    return "|".join([
        "fp32" if p.precision == FP32 else "fp64",
        f"{g.x}x{g.y}x{g.z}x{g.phi}x{g.theta}",
        f"n2^{n2}",
        f"nm{nm}",
        f"e{int(p.shape != 0.0)}r{int(p.volume_exclusion != 0.0)}q{int(bool(p.flags & FLAG_QUASI2D))}",
    ])


def _window(t, n):
    """Tile extent, halo and window of one axis as the library sets them."""
    ta = max(1, min(t, n))
    if n <= ta + 1:
        ta = n
    h = 0 if ta == n else 1
    return ta, ta + 2 * h, (n + ta - 1) // ta


def _tile_lds(ws, nm, na, naxes):
    """LDS bytes of the tiled update without staged cell records (the library's
    tile_lds_bytes): window accumulators, moment-kernel rows, window-slot tables."""
    return (ws * (nm + 1) * 8 + 15) // 16 * 16 + ((nm + 1) & ~1) * na * 8 + ((naxes + 56) * 4 + 15) // 16 * 16


def candidates(p, tiles=TILES, min_tiles=256):
    """Tile shapes valid for ``p`` (window within the LDS budget, at least
    ``min_tiles`` tiles), without duplicates after clamping to the grid."""
    g = p.grid_size
    dims = (g.x, g.y, g.z)
    nm = 9 if p.stress_magnetic != 0.0 else 6
    na = g.phi * g.theta
    out, seen = [], set()
    for t in tiles:
        ax = [_window(t[a], dims[a]) for a in range(3)]
        eff = tuple(a[0] for a in ax)
        ws = ax[0][1] * ax[1][1] * ax[2][1]
        ntiles = ax[0][2] * ax[1][2] * ax[2][2]
        if eff in seen or _tile_lds(ws, nm, na, sum(dims)) > LDS_LIMIT or ntiles < min_tiles:
            continue
        seen.add(eff)
        out.append(eff)
    return out or [(4, 4, 4)]


def _with(p, tile, resort):
    q = type(p).from_buffer_copy(p)
    q.tile = (C.c_int32 * 3)(*tile)
    q.resort_interval = resort
    q.binning = BIN_SORTED
    return q


def measure(p, parts, periods=2):
    """Wall-clock ms per step of ``p`` over ``periods`` whole re-sort periods
    (replayed graphs), after one warm-up period."""
    from .simulation import Simulation

    sim = Simulation(p)
    try:
        sim.init(parts)
        k = max(1, p.resort_interval)
        sim.step(k)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.step(periods * k)
        sim.synchronize()
        return (time.perf_counter() - t0) * 1e3 / (periods * k)
    finally:
        sim.close()


def tune(p, parts=None, tiles=None, resorts=RESORTS, measure_fn=None, periods=2):
    """Coordinate search: the tile shapes at a re-sort interval of 16, then
    the re-sort intervals for the best tile.  Returns the table entry
    {"tile", "resort_interval", "ms_per_step", "trials"}."""
    if measure_fn is None:
        if parts is None:
            from .simulation import initial_condition

            parts = initial_condition(p, amplitude=0.05, seed=p.seed)

        def measure_fn(q):
            return measure(q, parts, periods)

    trials = []
    best = None
    for t in candidates(p, tiles or TILES):
        ms = measure_fn(_with(p, t, 16))
        trials.append({"tile": list(t), "resort_interval": 16, "ms_per_step": ms})
        if best is None or ms < best[2]:
            best = (t, 16, ms)
    for r in resorts:
        if r == 16:
            continue
        ms = measure_fn(_with(p, best[0], r))
        trials.append({"tile": list(best[0]), "resort_interval": r, "ms_per_step": ms})
        if ms < best[2]:
            best = (best[0], r, ms)
    return {"tile": list(best[0]), "resort_interval": best[1], "ms_per_step": best[2], "trials": trials}


def load_table(path=None):
    path = path or TABLE
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_table(table, path=None):
    path = path or TABLE
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)


def tuned(p, table=None, path=None, tune_missing=False, **tune_kw):
    """``p`` with the tuned tile shape and re-sort interval of its size class.
    On a miss: tune (and store into ``path``) when ``tune_missing``, else
    return ``p`` unchanged (the library's defaults apply).  Parameters that
    set a tile shape or interval themselves are returned unchanged."""
    if any(p.tile) or p.resort_interval:
        return p
    table = load_table(path) if table is None else table
    key = size_class(p)
    entry = table.get(key)
    if entry is None and tune_missing:
        entry = tune(p, **tune_kw)
        table[key] = entry
        save_table(table, path)
    if entry is None:
        return p
    return _with(p, tuple(entry["tile"]), entry["resort_interval"])


def grid_class(p):
    """Size class without the particle count."""
    parts = size_class(p).split("|")
    return "|".join(parts[:2] + parts[3:])


def ensemble_key(plist):
    """Key of an ensemble: its members' size classes (sorted, without
    repeats) and the member count."""
    classes = sorted({size_class(p) for p in plist})
    return "ensemble|" + "+".join(classes) + f"|m{len(plist)}"


def measure_ensemble(plist, parts, periods=2):
    """Wall-clock ms per ensemble step over ``periods`` re-sort periods after
    one warm-up period (members share ``plist[0].resort_interval``)."""
    from .simulation import Ensemble

    ens = Ensemble(plist)
    try:
        for m, x in zip(ens.members, parts):
            m.init(x)
        k = max(1, plist[0].resort_interval)
        ens.step(k)
        t0 = time.perf_counter()
        ens.step(periods * k)  # returns after the steps have run
        return (time.perf_counter() - t0) * 1e3 / (periods * k)
    finally:
        ens.close()


def tune_ensemble(plist, parts=None, tiles=ENSEMBLE_TILES, resorts=RESORTS, measure_fn=None, periods=3,
                  start=None, min_tiles=16):
    """Coordinate search on the whole ensemble: starting from ``start``
    ({grid class: tile}, default 4x4x4), each grid class's tile shape in
    turn with the others fixed (at a re-sort interval of 16), then the shared
    re-sort interval.  Returns the table entry {"tiles": {grid class: tile},
    "resort_interval", "ms_per_step", "trials"}."""
    if measure_fn is None:
        if parts is None:
            from .simulation import initial_condition

            parts = [initial_condition(p, amplitude=0.05, seed=p.seed) for p in plist]

        def measure_fn(qs):
            return measure_ensemble(qs, parts, periods)

    classes = sorted({grid_class(p) for p in plist})
    first = {}
    for p in plist:
        first.setdefault(grid_class(p), p)
    if start is None:
        start = {c: (4, 4, 4) for c in classes}
    cur = {c: tuple(start[c]) for c in classes}

    def build(tiles_of, resort):
        return [_with(p, tiles_of[grid_class(p)], resort) for p in plist]

    trials = []
    best_ms = measure_fn(build(cur, 16))
    trials.append({"tiles": {c: list(t) for c, t in cur.items()}, "resort_interval": 16, "ms_per_step": best_ms})
    for c in classes:
        for t in candidates(first[c], tiles, min_tiles):
            if t == cur[c]:
                continue
            trial = dict(cur)
            trial[c] = t
            ms = measure_fn(build(trial, 16))
            trials.append({"tiles": {k: list(v) for k, v in trial.items()}, "resort_interval": 16, "ms_per_step": ms})
            if ms < best_ms:
                best_ms, cur = ms, trial
    best_r = 16
    for r in resorts:
        if r == 16:
            continue
        ms = measure_fn(build(cur, r))
        trials.append({"tiles": {k: list(v) for k, v in cur.items()}, "resort_interval": r, "ms_per_step": ms})
        if ms < best_ms:
            best_ms, best_r = ms, r
    return {"tiles": {c: list(t) for c, t in cur.items()}, "resort_interval": best_r, "ms_per_step": best_ms,
            "trials": trials}


def tuned_ensemble(plist, table=None, path=None, tune_missing=False, **tune_kw):
    """Members of an ensemble with the tile shapes and the shared re-sort
    interval of its entry (``ensemble_key``); on a miss, tune when
    ``tune_missing``, else each member as ``tuned`` gives it, sharing the
    first member's re-sort interval."""
    table = load_table(path) if table is None else table
    key = ensemble_key(plist)
    entry = table.get(key)
    if entry is None and tune_missing:
        entry = tune_ensemble(plist, **tune_kw)
        table[key] = entry
        save_table(table, path)
    if entry is None:
        out = [type(p).from_buffer_copy(tuned(p, table)) for p in plist]
        for q in out:
            q.resort_interval = out[0].resort_interval
        return out
    return [_with(p, tuple(entry["tiles"][grid_class(p)]), entry["resort_interval"]) for p in plist]


def main():
    ap = argparse.ArgumentParser(description="tune the tiled path's tile shape and re-sort interval")
    ap.add_argument("--grid", type=int, nargs=5, default=[64, 64, 64, 16, 16])
    ap.add_argument("--box", type=float, nargs=3, default=None)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--out", default=TABLE)
    args = ap.parse_args()
    box = args.box or [0.78 * args.grid[0], 0.78 * args.grid[1], 0.78 * args.grid[2]]
    p = make_params(box=tuple(box), grid=tuple(args.grid), n=args.n, seed=2,
                    precision=FP64 if args.fp64 else FP32)
    table = load_table(args.out)
    entry = tune(p)
    table[size_class(p)] = entry
    save_table(table, args.out)
    print(json.dumps({size_class(p): entry}))


if __name__ == "__main__":
    main()

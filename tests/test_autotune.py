"""Host logic of the autotuner (stochasticsampling_amd/autotune.py): size
classes, candidate filtering, the coordinate search and the table; one GPU
test tunes a small system for real."""
import json

import pytest

import stochasticsampling_amd as ss
from stochasticsampling_amd import autotune as at


def _p(**kw):
    base = dict(box=(50.0, 50.0, 50.0), grid=(64, 64, 64, 16, 16), n=1 << 24, precision=ss.FP32)
    base.update(kw)
    return ss.make_params(**base)


def test_size_class_separates_what_changes_the_kernels():
    a = at.size_class(_p())
    assert a == "fp32|64x64x64x16x16|n2^24|nm6|e0r0q0"
    assert at.size_class(_p(n=(1 << 24) + 1000)) == a  # same power-of-two bucket
    assert at.size_class(_p(n=1 << 22)) != a
    assert at.size_class(_p(stress_magnetic=0.5)) != a
    assert at.size_class(_p(shape=0.1)) != a
    assert at.size_class(_p(precision=ss.FP64)) != a


def test_candidates_respect_lds_budget_and_grid():
    c = at.candidates(_p())
    assert (4, 4, 4) in c and len(c) == len(set(c))
    # 16^3 x 8 x 8 fits every window; an 8-wide grid clamps 8-tiles to the axis
    small = at.candidates(_p(grid=(8, 8, 8, 4, 4), n=4096))
    for t in small:
        assert all(1 <= v <= 8 for v in t)
    # 24 x 24 angle bins leave no room for 4^3 windows of 9 components
    huge = at.candidates(_p(grid=(64, 64, 64, 24, 24), stress_magnetic=1.0))
    for t in huge:
        ws = 1
        for v in t:
            ws *= v + 2
        assert ws * 10 * 8 + 9 * 576 * 8 <= at.LDS_LIMIT or huge == [(4, 4, 4)]


def test_tune_coordinate_search_with_model_timings():
    cost = {(4, 4, 4): 1.0, (2, 4, 4): 0.9, (4, 4, 2): 1.1, (4, 4, 8): 1.2, (4, 8, 4): 1.3, (2, 2, 4): 1.4,
            (8, 4, 4): 1.5}
    rfac = {8: 1.05, 16: 1.0, 32: 0.97}
    seen = []

    def fake(q):
        t = tuple(q.tile)
        seen.append((t, q.resort_interval))
        assert q.binning == ss.BIN_SORTED
        return cost[t] * rfac[q.resort_interval]

    e = at.tune(_p(), measure_fn=fake)
    assert e["tile"] == [2, 4, 4] and e["resort_interval"] == 32
    assert abs(e["ms_per_step"] - 0.9 * 0.97) < 1e-12
    # tiles at interval 16 first, then the other intervals for the winner
    assert all(r == 16 for _, r in seen[: len(at.candidates(_p()))])
    assert seen[-2:] == [((2, 4, 4), 8), ((2, 4, 4), 32)]


def test_tuned_uses_table_and_respects_explicit_choices(tmp_path):
    path = str(tmp_path / "t.json")
    p = _p()
    at.save_table({at.size_class(p): {"tile": [4, 4, 8], "resort_interval": 32, "ms_per_step": 0.4}}, path)
    q = at.tuned(p, path=path)
    assert tuple(q.tile) == (4, 4, 8) and q.resort_interval == 32
    assert tuple(p.tile) == (0, 0, 0)  # the input is not modified
    explicit = _p(tile=(2, 2, 2))
    assert at.tuned(explicit, path=path) is explicit
    miss = _p(n=1 << 20)
    assert at.tuned(miss, path=path) is miss
    stored = at.tuned(miss, path=path, tune_missing=True, measure_fn=lambda q: float(sum(q.tile)))
    assert tuple(stored.tile) == (2, 2, 4)
    with open(path) as f:
        assert at.size_class(miss) in json.load(f)


def test_shipped_table_entries_are_valid():
    for key, e in at.load_table().items():
        assert e["resort_interval"] >= 1
        if key.startswith("ensemble|"):
            assert key.rsplit("|m", 1)[1].isdigit()
            for c, t in e["tiles"].items():
                assert c.count("|") == 3 and len(t) == 3 and all(v >= 1 for v in t)
            continue
        assert len(e["tile"]) == 3 and all(v >= 1 for v in e["tile"])
        assert key.count("|") == 4


def test_shipped_ensemble_entry_covers_the_benchmark_ensemble():
    import bench

    plist = bench.ensemble_configs()
    assert at.ensemble_key(plist) in at.load_table()
    out = at.tuned_ensemble(plist)
    assert len({q.resort_interval for q in out}) == 1 and all(any(q.tile) for q in out)


@pytest.mark.gpu
def test_tune_small_system_on_gpu():
    p = ss.make_params(box=(12.5, 12.5, 12.5), grid=(16, 16, 16, 8, 8), n=1 << 15, precision=ss.FP32, seed=3)
    tiles = [(4, 4, 4), (2, 2, 2), (2, 2, 4)]
    e = at.tune(p, tiles=tiles, resorts=[8, 16])
    cands = at.candidates(p, tiles)  # 4^3 tiles leave fewer than 256 tiles on 16^3
    assert cands == [(2, 2, 2), (2, 2, 4)]
    assert tuple(e["tile"]) in cands and e["resort_interval"] in (8, 16)
    assert len(e["trials"]) == len(cands) + 1 and all(t["ms_per_step"] > 0 for t in e["trials"])
    sim = ss.Simulation(at.tuned(p, table={at.size_class(p): e}))
    assert list(sim.get_info()["tile"]) == e["tile"]


def _ens():
    return [_p(grid=(32, 32, 32, 8, 8), n=1 << 18), _p(grid=(64, 64, 64, 16, 16), n=1 << 20),
            _p(grid=(32, 32, 32, 8, 8), n=(1 << 18) + 5)]


def test_ensemble_key_is_the_set_of_classes_and_the_member_count():
    plist = _ens()
    assert at.grid_class(plist[0]) == "fp32|32x32x32x8x8|nm6|e0r0q0"
    assert at.grid_class(_p(n=1 << 10)) == at.grid_class(_p())
    k = at.ensemble_key(plist)
    assert k == at.ensemble_key(list(reversed(plist)))
    assert k.startswith("ensemble|") and k.endswith("|m3")
    assert k.count("fp32|") == 2  # two distinct classes
    assert at.ensemble_key(plist[:2]) != k


def test_tune_ensemble_coordinate_search_with_model_timings():
    """Each class's shape in turn with the others fixed, then the interval;
    the model favours 4x4x8 for the 32^3 class and 8x8x4 for the 64^3 one."""
    plist = _ens()
    c32, c64 = at.grid_class(plist[0]), at.grid_class(plist[1])
    calls = []

    def fake(qs):
        assert len(qs) == 3 and all(q.binning == ss.BIN_SORTED for q in qs)
        assert len({q.resort_interval for q in qs}) == 1
        t32 = {tuple(q.tile) for q in qs if at.grid_class(q) == c32}
        assert len(t32) == 1  # members of one grid share the shape
        t64 = tuple(next(q.tile for q in qs if at.grid_class(q) == c64))
        calls.append(1)
        cost = {(4, 4, 8): 1.0}.get(t32.pop(), 1.5) + {(8, 8, 4): 1.0}.get(t64, 2.0)
        return cost * {8: 1.1, 16: 1.0, 32: 0.9}[qs[0].resort_interval]

    e = at.tune_ensemble(plist, measure_fn=fake, start={c32: (4, 4, 4), c64: (4, 4, 4)})
    assert e["tiles"] == {c32: [4, 4, 8], c64: [8, 8, 4]}
    assert e["resort_interval"] == 32 and abs(e["ms_per_step"] - 1.8) < 1e-12
    assert len(calls) == len(e["trials"])
    # applying the entry
    key = at.ensemble_key(plist)
    out = at.tuned_ensemble(plist, table={key: e})
    assert [tuple(q.tile) for q in out] == [(4, 4, 8), (8, 8, 4), (4, 4, 8)]
    assert {q.resort_interval for q in out} == {32}
    # a miss without tuning: the standalone entries, one shared interval
    miss = at.tuned_ensemble(plist, table={})
    assert len({q.resort_interval for q in miss}) == 1

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
        assert len(e["tile"]) == 3 and all(v >= 1 for v in e["tile"])
        assert e["resort_interval"] >= 1
        assert key.count("|") == 4


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

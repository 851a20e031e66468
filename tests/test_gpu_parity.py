"""GPU parity of libss against the FP64 CPU oracle (through the C ABI).

Bars: integer outputs (bin indices, counts) bit-exact; FP64 fields and states
within 1e-10 relative of the oracle; FP32 fields within 1e-5 relative of the
FP64 oracle (the stated target for u and grad u); FP32 states as stated per
test.
"""
import math

import numpy as np
import pytest

import stochasticsampling_amd as ss
from helpers import isotropic_particles, rel_err, small_params, periodic_diff

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_FIELD_TOL = 1e-5


# ---------------------------------------------------------------------------
# FFT plan (FFTPlan, crates/fftw3/src/fft.rs)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(16, 16, 16), (11, 11, 11), (8, 12, 10), (64, 64, 64), (32, 48, 6)])
@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
@pytest.mark.parametrize("direction", [-1, 1])
def test_fft_plan_matches_numpy(shape, precision, direction):
    rng = np.random.default_rng(1)
    batch = 3
    x = rng.standard_normal((batch,) + shape) + 1j * rng.standard_normal((batch,) + shape)
    dt = np.complex128 if precision == ss.FP64 else np.complex64
    a = np.ascontiguousarray(x.astype(dt))
    plan = ss.FFTPlan(shape, direction, batch=batch, precision=precision)
    plan.reexecute3d(a)
    axes = (1, 2, 3)
    ref = np.fft.fftn(x, axes=axes) if direction == -1 else np.fft.ifftn(x, axes=axes) * np.prod(shape)
    tol = 1e-13 if precision == ss.FP64 else 2e-6
    assert rel_err(a, ref) < tol


@pytest.mark.parametrize("shape", [(16, 16, 16), (64, 64, 64), (32, 48, 6), (8, 12, 10), (96, 16, 7)])
@pytest.mark.parametrize("direction", [-1, 1])
def test_fft_plan_dense_dft_variant(shape, direction):
    """Dense DFT products on the matrix cores (three-part bf16 splits, six
    products, FP32 accumulation) for axes of at most 64 points, radix stages
    for longer axes: about 24 significant bits per product, the radix FFT's
    FP32 bound of 2e-6 of the largest output."""
    rng = np.random.default_rng(2)
    batch = 2
    x = rng.standard_normal((batch,) + shape) + 1j * rng.standard_normal((batch,) + shape)
    a = np.ascontiguousarray(x.astype(np.complex64))
    plan = ss.FFTPlan(shape, direction, batch=batch, precision=ss.FP32, variant=ss.FFT_DFT)
    plan.reexecute3d(a)
    axes = (1, 2, 3)
    ref = np.fft.fftn(x, axes=axes) if direction == -1 else np.fft.ifftn(x, axes=axes) * np.prod(shape)
    assert rel_err(a, ref) < 2e-6


def test_fft_plan_dense_dft_variant_is_fp32_only():
    with pytest.raises(ss.SSError):
        ss.FFTPlan((16, 16, 16), -1, precision=ss.FP64, variant=ss.FFT_DFT)


# ---------------------------------------------------------------------------
# Binning rule (Distribution::coord_to_grid, distribution_test.rs:119-184)
# ---------------------------------------------------------------------------
def test_coord_to_grid_reference_table(oracle):
    p = ss.make_params(box=(1.0, 1.0, 1.0), grid=(50, 50, 1, 10, 1))
    inputs = [
        [0.0, 0.0, 0.0, 0.0, 0.0],
        [1.0, 0.0, 0.0, 0.0, 0.0],
        [0.0, 1.0, 0.0, 0.0, 0.0],
        [0.0, 0.0, 0.0, 0.5, 0.0],
        [0.0, 0.0, 0.0, 7.0, 0.0],
        [0.0, 0.0, 0.0, -1.0, 0.0],
        [0.96, 0.0, 0.0, 0.0, 0.0],
        [0.5, 0.5, 0.0, -1.0, 0.0],
        [0.5, 0.5, 0.0, 0.0, 0.0],
        [0.0, 0.0, 0.0, 2.0 * math.pi, 0.0],
        [0.51000000000000005, 0.5, 0.0, 6.283185307179584, 0.0],
    ]
    expect = [
        [0, 0, 0, 0, 0],
        [0, 0, 0, 0, 0],
        [0, 0, 0, 0, 0],
        [0, 0, 0, 0, 0],
        [0, 0, 0, 1, 0],
        [0, 0, 0, 8, 0],
        [48, 0, 0, 0, 0],
        [25, 25, 0, 8, 0],
        [25, 25, 0, 0, 0],
        [0, 0, 0, 0, 0],
        [25, 25, 0, 9, 0],
    ]
    sanitized = np.array([oracle.particle_new(i, (1.0, 1.0, 1.0)) for i in inputs])
    got = ss.debug_coord_to_grid(p, sanitized)
    np.testing.assert_array_equal(got, np.array(expect, dtype=np.int32))


def test_coord_to_grid_fp32_matches_oracle_rule(oracle):
    p = ss.make_params(box=(50.0, 50.0, 50.0), grid=(64, 64, 64, 16, 16), precision=ss.FP32)
    parts = isotropic_particles(20000, (50.0, 50.0, 50.0), 3)
    # plant values on and next to bin edges
    g = 50.0 / 64
    parts[:200, 0] = np.arange(200) % 64 * g
    parts[200:400, 0] = np.nextafter((np.arange(200) % 64 + 1) * g, 0)
    parts32 = parts.astype(np.float32).astype(np.float64)
    got = ss.debug_coord_to_grid(p, parts32)
    exp = np.array([oracle.coord_to_grid_f32(p, q) for q in parts32.astype(np.float32)])
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("precision", [ss.FP32, ss.FP64])
def test_coord_to_grid_on_bin_edges_all_axes(oracle, precision):
    """The binning multiplies by the rounded reciprocal of the bin width and
    divides only next to an integer quotient; on every axis, at each bin edge
    k * g (rounded to the working precision), its neighbours one ulp away and
    random values, the index equals floor(fl(q / g)) of the reference rule."""
    grid = (64, 48, 40, 16, 12)
    box = (50.0, 37.0, 29.5)
    p = ss.make_params(box=box, grid=grid, precision=precision)
    dt = np.float32 if precision == ss.FP32 else np.float64
    widths = [box[0] / grid[0], box[1] / grid[1], box[2] / grid[2], 2 * np.pi / grid[3], np.pi / grid[4]]
    tops = [box[0], box[1], box[2], 2 * np.pi, np.pi]
    rng = np.random.default_rng(21)
    cols = []
    for a in range(5):
        g = dt(widths[a])
        k = np.arange(grid[a] + 1, dtype=dt)
        edge = (k * g).astype(dt)
        vals = np.concatenate([edge, np.nextafter(edge, dt(-1)), np.nextafter(edge, dt(1e9)),
                               (rng.random(3000) * tops[a]).astype(dt)])
        vals = vals[(vals >= 0) & (vals <= dt(tops[a]))]
        cols.append(vals)
    m = max(len(c) for c in cols)
    parts = np.zeros((m, 5), dtype=dt)
    for a in range(5):
        parts[:, a] = np.resize(cols[a], m)
    got = ss.debug_coord_to_grid(p, parts.astype(np.float64))
    if precision == ss.FP32:
        exp = np.array([oracle.coord_to_grid_f32(p, q) for q in parts])
    else:
        exp = np.array([oracle.coord_to_grid(p, q) for q in parts])
    np.testing.assert_array_equal(got, exp)


# ---------------------------------------------------------------------------
# Particle update (modifiers_test.rs:39-97, mod_test.rs) and random cases
# ---------------------------------------------------------------------------
def _fields(u=None, W=None, E=None, rho=0.0, B=None, GB=None):
    f = np.zeros(34)
    if u is not None:
        f[0:3] = u
    if W is not None:
        f[3:12] = np.asarray(W).ravel()
    if E is not None:
        f[12:21] = np.asarray(E).ravel()
    f[21] = rho
    if B is not None:
        f[22:25] = B
    if GB is not None:
        f[25:34] = np.asarray(GB).ravel()
    return f


def test_update_matches_oracle_random_fp64(oracle):
    rng = np.random.default_rng(5)
    n = 4096
    p = ss.make_params(
        box=(10.0, 10.0, 10.0),
        timestep=1e-2,
        shape=0.7,
        magnetic_drag=0.3,
        magnetic_reorientation=1.3,
        magnetic_dipole_dipole=0.4,
        volume_exclusion=0.2,
    )
    parts = isotropic_particles(n, (10.0, 10.0, 10.0), 11)
    fields = rng.standard_normal((n, 34))
    W = rng.standard_normal((n, 3, 3))
    W = 0.5 * (W - W.transpose(0, 2, 1))
    E = rng.standard_normal((n, 3, 3))
    E = 0.5 * (E + E.transpose(0, 2, 1))
    fields[:, 3:12] = W.reshape(n, 9)
    fields[:, 12:21] = E.reshape(n, 9)
    fields[:, 21] = np.abs(fields[:, 21])
    rv = rng.standard_normal((n, 5))
    rv[:, 3] = rng.random(n) * 2 * np.pi
    rv[:, 4] = np.abs(rv[:, 4]) * 0.1
    got = ss.debug_update(p, parts, fields, rv)
    exp = np.array([oracle.update_one(p, parts[i], fields[i], rv[i]) for i in range(n)])
    assert np.max(periodic_diff(got[:, :3], exp[:, :3], 10.0)) < 1e-12
    assert np.max(periodic_diff(got[:, 3], exp[:, 3], 2 * np.pi)) < 1e-10
    assert np.max(np.abs(got[:, 4] - exp[:, 4])) < 1e-10


def test_update_modifier_cases_fp64(oracle):
    pi = math.pi
    bs = (10.0, 10.0, 10.0)
    p0 = [0.0, 0.0, 0.0, 0.0, 0.0]
    cases = []  # (params kwargs, particle, fields, rv, expected particle)
    zero_rv = np.zeros(5)
    cases.append(({}, p0, _fields(), zero_rv, [0.0, 0.0, 1.0, 0.0, 0.0]))  # self_propulsion
    cases.append(
        ({"magnetic_reorientation": 0.1}, [0, 0, 0, 0, pi / 2], _fields(u=[-1.0, 0.0, 0.0]), zero_rv,
         [0.0, 0.0, 0.0, 0.09966865249116204, pi / 2]))
    results = []
    for kw, part, f, rv, exp in cases:
        prm = ss.make_params(box=bs, timestep=1.0, diffusion_translational=0.0, **kw)
        part = oracle.particle_new(part, bs)
        got = ss.debug_update(prm, part[None, :], f[None, :], rv[None, :])[0]
        want = oracle.particle_new(exp, bs)
        results.append((got, want))
    for got, want in results:
        assert np.max(np.abs(got - want)) < 1e-15, (got, want)


# ---------------------------------------------------------------------------
# Spectral solves (ff_test fixture configuration; magnetic solve)
# ---------------------------------------------------------------------------
def _single_particle_dist(oracle, p, part):
    cnt = oracle.histogram(p, np.asarray(part)[None, :])
    return oracle.sample_from(p, cnt, 1)


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_flow_solve_single_particle_fixture(oracle, precision):
    p = ss.make_params(box=(11.0, 11.0, 11.0), grid=(11, 11, 11, 11, 11), n=1, stress_active=1.0,
                       precision=precision)
    part = oracle.particle_new([0, 0, 0, math.pi / 2, math.pi / 2], (11.0, 11.0, 11.0))
    dist = _single_particle_dist(oracle, p, part)
    u, g = ss.mean_flow_field(p, dist)
    ue, ge = oracle.mean_flow_field(p, 0.0, dist)
    tol = FP64_TOL if precision == ss.FP64 else FP32_FIELD_TOL
    assert rel_err(u, ue) < tol
    assert rel_err(g, ge) < tol


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
@pytest.mark.parametrize("magnetic_stress", [0.0, 0.8])
def test_flow_solve_random_distribution(oracle, precision, magnetic_stress):
    p = ss.make_params(box=(10.0, 12.0, 9.0), grid=(16, 12, 10, 6, 5), stress_active=-1.3,
                       stress_magnetic=magnetic_stress, shape=0.4, hydro_screening=0.25,
                       precision=precision)
    rng = np.random.default_rng(2)
    dist = rng.random(p.dist_shape) * 3.0
    u, g = ss.mean_flow_field(p, dist)
    ue, ge = oracle.mean_flow_field(p, 0.25, dist)
    tol = FP64_TOL if precision == ss.FP64 else FP32_FIELD_TOL
    assert rel_err(u, ue) < tol
    assert rel_err(g, ge) < tol


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_magnetic_solve_random_distribution(oracle, precision):
    p = ss.make_params(box=(10.0, 10.0, 10.0), grid=(12, 12, 12, 8, 6), precision=precision)
    rng = np.random.default_rng(3)
    dist = rng.random(p.dist_shape)
    b, g = ss.mean_magnetic_field(p, dist)
    be, ge = oracle.mean_magnetic_field(p, dist)
    tol = FP64_TOL if precision == ss.FP64 else FP32_FIELD_TOL
    assert rel_err(b, be) < tol
    assert rel_err(g, ge) < tol


# ---------------------------------------------------------------------------
# Whole steps against the oracle
# ---------------------------------------------------------------------------
def _run_pair(oracle, p, parts, steps):
    sim = ss.Simulation(p)
    sim.init(parts)
    osim = oracle.simulation(p)
    osim.init(parts)
    if steps:
        sim.step(steps)
        osim.step(steps)
    return sim, osim


@pytest.mark.parametrize("binning", [ss.BIN_HISTOGRAM, ss.BIN_SORTED, ss.BIN_PLAIN_TILED])
def test_init_histogram_bit_exact_fp64(oracle, binning):
    p = small_params(binning=binning, flags=ss.FLAG_KEEP_DISTRIBUTION)
    parts = isotropic_particles(p.number_of_particles, (10.0, 10.0, 10.0), 4)
    sim, osim = _run_pair(oracle, p, parts, 0)
    np.testing.assert_array_equal(sim.get_counts(), osim.counts())
    np.testing.assert_array_equal(sim.get_distribution(), osim.distribution())
    np.testing.assert_array_equal(sim.get_particles(), osim.particles())


@pytest.mark.parametrize("binning", [ss.BIN_HISTOGRAM, ss.BIN_SORTED, ss.BIN_PLAIN_TILED])
def test_steps_fp64_match_oracle(oracle, binning):
    p = small_params(binning=binning, flags=ss.FLAG_KEEP_DISTRIBUTION, n=4096)
    parts = ss.initial_condition(p, amplitude=0.05, seed=1)
    sim, osim = _run_pair(oracle, p, parts, 5)
    assert sim.get_timestep() == osim.timestep == 5
    np.testing.assert_array_equal(sim.get_counts(), osim.counts())
    u, g = sim.get_flow_field(with_gradient=True)
    ue, ge = osim.flow_field()
    assert rel_err(u, ue) < FP64_TOL
    assert rel_err(g, ge) < FP64_TOL
    a, b = sim.get_particles(), osim.particles()
    assert np.max(periodic_diff(a[:, :3], b[:, :3], 10.0)) < 1e-9
    assert np.max(periodic_diff(a[:, 3], b[:, 3], 2 * np.pi)) < 1e-9
    assert np.max(np.abs(a[:, 4] - b[:, 4])) < 1e-9


def test_sorted_and_histogram_paths_identical_fp64():
    base = dict(n=4096, flags=ss.FLAG_KEEP_DISTRIBUTION)
    p1 = small_params(binning=ss.BIN_HISTOGRAM, **base)
    p2 = small_params(binning=ss.BIN_SORTED, **base)
    parts = ss.initial_condition(p1, amplitude=0.05, seed=2)
    out = []
    for p in (p1, p2):
        sim = ss.Simulation(p)
        sim.init(parts)
        sim.step(4)
        out.append((sim.get_particles(), sim.get_counts()))
    np.testing.assert_array_equal(out[0][1], out[1][1])
    # identical per-particle arithmetic; the moments differ only in summation
    # order, so the states agree to rounding
    assert np.max(np.abs(out[0][0] - out[1][0])) < 1e-12


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_plain_tiled_binning_exact_and_independent_of_sort_interval(precision):
    """The plain path with tile-ordered particles builds the same count grid
    as the atomic histogram path and contracts it with the same kernel, so
    counts, fields and states equal that path's bit for bit, whatever the
    sort interval and however many particles drifted out of their tile
    between sorts (fast translational diffusion here: about a quarter of a
    cell per step).  The tiled path bins the same counts."""
    kw = dict(grid=(16, 16, 16, 8, 8), n=1 << 14, precision=precision, diffusion_translational=5.0)
    parts = ss.initial_condition(small_params(**kw), amplitude=0.05, seed=6)
    runs = {}
    for name, binning, resort in [("pt1", ss.BIN_PLAIN_TILED, 1), ("pt5", ss.BIN_PLAIN_TILED, 5),
                                  ("pt16", ss.BIN_PLAIN_TILED, 16), ("tiled", ss.BIN_SORTED, 4),
                                  ("hist", ss.BIN_HISTOGRAM, 0)]:
        sim = ss.Simulation(small_params(binning=binning, resort_interval=resort, flags=ss.FLAG_KEEP_DISTRIBUTION,
                                         **kw))
        if binning == ss.BIN_PLAIN_TILED:
            assert sim.get_info()["plain_tiled"]
        sim.init(parts)
        sim.step(1)
        first = (sim.get_flow_field(), sim.get_counts())
        sim.step(20)
        runs[name] = (*first, sim.get_particles(), sim.get_counts(), sim.get_flow_field())
        sim.close()
    ref = runs["hist"]
    for name in ("pt1", "pt5", "pt16"):
        for a, b in zip(runs[name], ref):
            np.testing.assert_array_equal(a, b)
    # the first step bins the same particles on every path; the tiled path
    # sums fixed-point moments
    np.testing.assert_array_equal(runs["tiled"][1], ref[1])
    assert rel_err(runs["tiled"][0], ref[0]) < (1e-12 if precision == ss.FP64 else 1e-5)


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_tiled_path_independent_of_sort_interval_and_tiles(precision):
    """The tiled binning sums exact fixed-point integers, so neither the
    re-sort interval nor the tile shape changes a single bit of the result."""
    parts = ss.initial_condition(small_params(grid=(16, 16, 16, 8, 8), n=1 << 14), amplitude=0.05, seed=6)
    out = []
    for tile, resort in [((4, 4, 4), 1), ((4, 4, 4), 7), ((2, 2, 2), 16), ((16, 4, 2), 3), ((1, 1, 1), 5)]:
        p = small_params(grid=(16, 16, 16, 8, 8), n=1 << 14, precision=precision, binning=ss.BIN_SORTED,
                         tile=tile, resort_interval=resort, flags=ss.FLAG_KEEP_DISTRIBUTION)
        sim = ss.Simulation(p)
        sim.init(parts)
        sim.step(9)
        out.append((sim.get_particles(), sim.get_counts(), sim.get_flow_field()))
    for o in out[1:]:
        np.testing.assert_array_equal(o[0], out[0][0])
        np.testing.assert_array_equal(o[1], out[0][1])
        np.testing.assert_array_equal(o[2], out[0][2])


def test_tiled_quasi2d_and_odd_grid(oracle):
    for grid, box, flags in [((12, 10, 1, 6, 4), (10.0, 8.0, 1.0), ss.FLAG_QUASI2D), ((11, 9, 7, 5, 3),
                                                                                      (11.0, 9.0, 7.0), 0)]:
        kw = dict(grid=grid, box=box, n=3000, flags=flags)
        p = small_params(binning=ss.BIN_SORTED, **kw)
        p.flags |= ss.FLAG_KEEP_DISTRIBUTION
        parts = ss.initial_condition(small_params(**kw), amplitude=0.0, seed=8)
        sim = ss.Simulation(p)
        sim.init(parts)
        osim = oracle.simulation(small_params(**kw))
        osim.init(parts)
        sim.step(3)
        osim.step(3)
        np.testing.assert_array_equal(sim.get_counts(), osim.counts())
        a, b = sim.get_particles(), osim.particles()
        assert np.max(periodic_diff(a[:, 0], b[:, 0], box[0])) < 1e-9


@pytest.mark.parametrize("precision, binning", [(ss.FP64, ss.BIN_HISTOGRAM), (ss.FP32, ss.BIN_SORTED),
                                                (ss.FP64, ss.BIN_SORTED)])
def test_graph_replays_match_eager_steps(precision, binning):
    """Steps replayed from captured hipGraphs equal eagerly enqueued steps bit
    for bit, replay after replay, with host reads between the replays."""
    kw = dict(grid=(16, 16, 16, 8, 8), n=1 << 13, precision=precision, binning=binning, resort_interval=4)
    parts = ss.initial_condition(small_params(**kw), amplitude=0.05, seed=9)
    g = ss.Simulation(small_params(**kw))
    e = ss.Simulation(small_params(flags=ss.FLAG_NO_GRAPH, **kw))
    g.init(parts)
    e.init(parts)
    for chunk in [1, 3, 4, 8, 2, 16, 5, 9]:
        assert g.step(chunk) == e.step(chunk)
        np.testing.assert_array_equal(g.get_particles(), e.get_particles())
    np.testing.assert_array_equal(g.get_flow_field(), e.get_flow_field())


@pytest.mark.parametrize("precision", [ss.FP32, ss.FP64])
@pytest.mark.parametrize(
    "grid, extra",
    [
        ((16, 16, 16, 8, 8), {}),
        ((32, 32, 32, 4, 4), {}),
        ((48, 48, 48, 2, 2), dict(volume_exclusion=0.3)),
        ((64, 64, 64, 2, 2), dict(shape=0.5)),
        ((12, 20, 15, 6, 4), dict(shape=0.3, volume_exclusion=0.2)),
        ((16, 8, 10, 8, 8), dict(stress_magnetic=0.7)),
        ((12, 10, 1, 6, 4), dict(flags=ss.FLAG_QUASI2D)),
        ((16, 16, 16, 8, 8), dict(magnetic_reorientation=2.0, field_direction=(0.0, 0.0, 1.0),
                                  magnetic_dipole_dipole=0.5, magnetic_drag=0.1)),
        ((12, 10, 8, 6, 4), dict(magnetic_dipole_dipole=0.4, shape=0.3)),
        ((16, 8, 10, 6, 6), dict(magnetic_drag=0.2, volume_exclusion=0.2)),
    ],
)
def test_fused_spectral_stage_equals_axis_passes(precision, grid, extra):
    """The fused plane/pencil kernels of the spectral stage run the same radix
    stages, axis order and solve as convert + one pass per FFT axis + solve +
    pack, so particles and fields agree bit for bit -- also where all axes are
    16, 32, 48 or 64 long and the stages have their sizes fixed at compile
    time, and with the magnetic solve (b times mu_dd and grad b in the
    records, alone or with E and the density)."""
    extra = dict(extra)
    flags = extra.pop("flags", 0)
    box = tuple(0.8 * g for g in grid[:3])
    kw = dict(grid=grid, box=box, n=6000, precision=precision, binning=ss.BIN_SORTED, resort_interval=3, **extra)
    parts = ss.initial_condition(small_params(**kw), amplitude=0.1, seed=12)
    sims = [ss.Simulation(small_params(flags=flags | f, **kw))
            for f in (0, ss.FLAG_AXIS_PASSES)]
    assert sims[0].get_info()["fused_spectral"] and not sims[1].get_info()["fused_spectral"]
    assert not sims[0].get_info()["paired_spectral"]
    out = []
    for sim in sims:
        sim.init(parts)
        sim.step(7)
        u, g = sim.get_flow_field(with_gradient=True)
        out.append((sim.get_particles(), u, g))
        if "magnetic_dipole_dipole" in extra or "magnetic_drag" in extra:
            out[-1] += sim.get_magnetic_field(with_gradient=True)
        # a second read of the fields finds the same (the x transform of the
        # moments is completed once)
        np.testing.assert_array_equal(sim.get_flow_field(), u)
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision,tol", [(ss.FP64, 1e-13), (ss.FP32, 2e-6)])
@pytest.mark.parametrize("grid,extra", [
    ((16, 16, 16, 8, 8), {}),
    ((32, 32, 32, 4, 4), {"shape": 0.6}),
    ((12, 10, 15, 6, 6), {"shape": 0.6, "volume_exclusion": 0.3}),
    ((15, 9, 5, 4, 6), {"shape": -0.4}),
    ((8, 12, 16, 6, 4), {"hydro_screening": 0.5}),
])
def test_paired_spectral_stage_matches_unpaired(precision, tol, grid, extra):
    """Two real fields per complex transform, a pencil solved with its mirror
    and the Hermitian part of each output: the same fields as one transform
    per component up to rounding -- on even and odd grids (Nyquist planes
    or none), with the strain (12 outputs), the density and screening."""
    box = tuple(0.8 * g for g in grid[:3])
    kw = dict(grid=grid, box=box, n=6000, precision=precision, binning=ss.BIN_SORTED, resort_interval=3,
              flags=ss.FLAG_KEEP_DISTRIBUTION, **extra)
    parts = ss.initial_condition(small_params(**kw), amplitude=0.1, seed=13)
    b = ss.Simulation(small_params(**kw))
    kw["flags"] |= ss.FLAG_PAIRED_SPECTRAL
    a = ss.Simulation(small_params(**kw))
    assert a.get_info()["paired_spectral"] and not b.get_info()["paired_spectral"]
    for sim in (a, b):
        sim.init(parts)
        sim.step(1)
    ua, ga = a.get_flow_field(with_gradient=True)
    ub, gb = b.get_flow_field(with_gradient=True)
    assert rel_err(ua, ub) < tol and rel_err(ga, gb) < tol
    np.testing.assert_array_equal(a.get_counts(), b.get_counts())
    # one more step through the records the paired stage wrote
    for sim in (a, b):
        sim.step(1)
    pa, pb = a.get_particles(), b.get_particles()
    L = np.array(box)
    d = np.abs(pa[:, :3] - pb[:, :3])
    d = np.minimum(d, L - d)
    assert np.max(d) < 1e3 * tol * max(L)
    assert np.max(np.abs(pa[:, 4] - pb[:, 4])) < 1e3 * tol


@pytest.mark.parametrize("default_physics", [False, True])
@pytest.mark.parametrize("precision", [ss.FP32, ss.FP64])
def test_ensemble_members_equal_single_systems(precision, default_physics):
    """Members of a batched ensemble (different grids, boxes, particle counts,
    alpha and D_r) evolve bit for bit like the same systems stepped alone.
    The members' spectral stages run batched by axis-length class (16^3,
    32^3, 48^3 and 64^3 with compile-time sizes, the rest at run time; with
    the strain and the density, with screening; groups larger than a side
    stream's share split into chunks), except a member on the paired stage,
    which keeps its own launches.  With a member on the strain and the
    density the persistent update runs the full Langevin chain; with every
    member on the default physics, the variant with those terms compiled
    out."""
    specs = [((16, 16, 16, 8, 8), 12.5, 5000, -2.0, 0.1, {}), ((12, 12, 12, 6, 6), 9.4, 3000, 0.5, 1.0, {}),
             ((8, 8, 8, 4, 4), 6.2, 700, 1.7, 0.5, {}), ((20, 16, 12, 8, 8), 15.6, 9000, -0.3, 0.5, {}),
             ((16, 16, 16, 8, 8), 12.5, 1, 1.0, 0.1, {}),
             ((32, 32, 32, 4, 4), 25.0, 20000, 0.8, 0.5, dict(shape=0.5, volume_exclusion=0.2)),
             ((48, 48, 48, 2, 2), 37.4, 30000, -1.5, 1.0, {}), ((64, 64, 64, 2, 2), 50.0, 30000, 1.2, 0.1, {}),
             ((64, 64, 64, 2, 2), 50.0, 10000, -0.7, 0.5, dict(hydro_screening=0.3)),
             ((16, 12, 10, 6, 6), 12.5, 4000, -1.0, 0.5, dict(flags=ss.FLAG_PAIRED_SPECTRAL))]
    if default_physics:
        specs[5] = ((32, 32, 32, 4, 4), 25.0, 20000, 0.8, 0.5, {})
    plist, parts = [], []
    for k, (grid, L, n, alpha, dr, extra) in enumerate(specs):
        p = ss.make_params(box=(L, L * grid[1] / grid[0], L * grid[2] / grid[0]), grid=grid, n=n, seed=100 + k,
                           stress_active=alpha, diffusion_rotational=dr, precision=precision,
                           binning=ss.BIN_SORTED, resort_interval=4, **extra)
        plist.append(p)
        parts.append(ss.initial_condition(p, amplitude=0.05, seed=100 + k))
    ens = ss.Ensemble(plist)
    for m, x in zip(ens.members, parts):
        m.init(x)
    singles = []
    for p, x in zip(plist, parts):
        sim = ss.Simulation(p)
        sim.init(x)
        singles.append(sim)
    for chunk in [3, 4, 8, 1, 6]:
        t = ens.step(chunk)
        for sim in singles:
            assert sim.step(chunk) == t
        for m, sim in zip(ens.members, singles):
            np.testing.assert_array_equal(m.get_particles(), sim.get_particles())
    for m, sim in zip(ens.members, singles):
        np.testing.assert_array_equal(m.get_flow_field(), sim.get_flow_field())
        assert m.get_timestep() == sim.get_timestep()


def test_external_noise_equals_philox_fp64(oracle):
    p = small_params(n=1024)
    parts = ss.initial_condition(p, amplitude=0.05, seed=3)
    sim = ss.Simulation(p)
    sim.init(parts)
    sim.step(1)
    pe = small_params(n=1024, noise_mode=ss.NOISE_EXTERNAL)
    dr = math.sqrt(2 * pe.diffusion_rotational * pe.timestep)
    rv = np.array([oracle.noise(pe.seed, 0, i, dr) for i in range(1024)])
    sim2 = ss.Simulation(pe)
    sim2.init(parts)
    sim2.set_noise(rv)
    sim2.step(1)
    assert np.max(np.abs(sim.get_particles() - sim2.get_particles())) < 1e-12


def test_fp32_fields_within_target_of_fp64(oracle):
    p32 = small_params(precision=ss.FP32, grid=(16, 16, 16, 8, 8), n=1 << 15, flags=ss.FLAG_KEEP_DISTRIBUTION)
    p64 = small_params(precision=ss.FP64, grid=(16, 16, 16, 8, 8), n=1 << 15)
    parts = ss.initial_condition(p64, amplitude=0.05, seed=4)
    sim = ss.Simulation(p32)
    sim.init(parts)
    sim.step(1)
    osim = oracle.simulation(p64)
    osim.init(parts)
    osim.step(1)
    # the counts of the FP32 binning: every particle binned, and the same
    # bins as the FP64 oracle except where an FP32-rounded position crossed
    # a cell edge (a handful of the 2^15 particles)
    cnt = sim.get_counts()
    assert int(cnt.sum()) == p32.number_of_particles
    moved = int(np.abs(cnt.astype(np.int64) - osim.counts().astype(np.int64)).sum())
    assert moved <= 20, moved
    # the FP32 fields against the FP64 oracle solve of the same distribution
    u, g = sim.get_flow_field(with_gradient=True)
    ue, ge = oracle.mean_flow_field(p64, 0.0, sim.get_distribution())
    assert rel_err(u, ue) < FP32_FIELD_TOL
    assert rel_err(g, ge) < FP32_FIELD_TOL
    a, b = sim.get_particles(), osim.particles()
    assert np.median(periodic_diff(a[:, :3], b[:, :3], 10.0)) < 1e-6
    assert np.percentile(periodic_diff(a[:, :3], b[:, :3], 10.0), 99) < 1e-5


def test_fp32_dense_dft_fields_against_fp64(oracle):
    """The FP32 step with the dense-DFT FFT variant against the FP64 oracle:
    u and grad u within the 1e-5 FP32 target, as the radix variant."""
    kw = dict(grid=(16, 16, 16, 8, 8), n=1 << 15)
    p32 = small_params(precision=ss.FP32, fft_variant=ss.FFT_DFT, flags=ss.FLAG_KEEP_DISTRIBUTION, **kw)
    p64 = small_params(precision=ss.FP64, **kw)
    parts = ss.initial_condition(p64, amplitude=0.05, seed=4)
    sim = ss.Simulation(p32)
    assert not sim.get_info()["fused_spectral"]
    sim.init(parts)
    sim.step(1)
    osim = oracle.simulation(p64)
    osim.init(parts)
    osim.step(1)
    moved = int(np.abs(sim.get_counts().astype(np.int64) - osim.counts().astype(np.int64)).sum())
    assert moved <= 20, moved
    u, g = sim.get_flow_field(with_gradient=True)
    ue, ge = oracle.mean_flow_field(p64, 0.0, sim.get_distribution())
    assert rel_err(u, ue) < 1e-5
    assert rel_err(g, ge) < 1e-5


def test_snapshot_resume_is_exact():
    for prec in (ss.FP64, ss.FP32):
        p = small_params(precision=prec, n=2048)
        parts = ss.initial_condition(p, amplitude=0.05, seed=5)
        a = ss.Simulation(p)
        a.init(parts)
        a.step(3)
        snap = a.get_snapshot()
        a.step(2)
        b = ss.Simulation(p)
        b.resume(snap)
        assert b.get_timestep() == 3
        b.step(2)
        np.testing.assert_array_equal(a.get_particles(), b.get_particles())


def test_init_rejects_wrong_particle_count():
    p = small_params(n=100)
    sim = ss.Simulation(p)
    with pytest.raises(ss.SSError, match="same number of particles"):
        sim.init(np.zeros((99, 5)))


# ---------------------------------------------------------------------------
# Size-independent properties at the full 2^24-particle, 64^3 x 16 x 16 size
# ---------------------------------------------------------------------------
def test_full_size_fp32_invariants():
    p = ss.make_params(box=(50.0, 50.0, 50.0), grid=(64, 64, 64, 16, 16), n=1 << 24, seed=2,
                       precision=ss.FP32, flags=ss.FLAG_KEEP_DISTRIBUTION)
    parts = ss.initial_condition(p, amplitude=0.05, seed=2)
    sim = ss.Simulation(p)
    sim.init(parts)
    sim.step(2)
    cnt = sim.get_counts()
    assert int(cnt.sum(dtype=np.uint64)) == p.number_of_particles
    u = sim.get_flow_field()
    # k = 0 mode of u is zero and the field is divergence free
    assert abs(u.mean()) < 1e-6 * np.abs(u).max()
    uh = np.fft.fftn(u, axes=(1, 2, 3))
    k = [2 * np.pi * np.fft.fftfreq(64, d=50.0 / 64)] * 3
    # use the reference's k-sampling (Nyquist stored negative)
    kx = k[0].copy()
    kx[32] = -kx[32] if kx[32] > 0 else kx[32]
    KX, KY, KZ = np.meshgrid(kx, kx, kx, indexing="ij")
    div = KX * uh[0] + KY * uh[1] + KZ * uh[2]
    inner = (np.abs(KX) < 0.9 * np.pi / (50 / 64)) & (np.abs(KY) < 0.9 * np.pi / (50 / 64)) & (
        np.abs(KZ) < 0.9 * np.pi / (50 / 64))
    assert np.abs(div[inner]).max() < 1e-4 * np.abs(uh).max() * np.abs(kx).max()
    parts2 = sim.get_particles()
    assert np.all(parts2[:, :3] >= 0) and np.all(parts2[:, :3] <= 50.0)
    assert np.all(parts2[:, 3] >= 0) and np.all(parts2[:, 3] <= 2 * np.pi)
    assert np.all(parts2[:, 4] >= 0) and np.all(parts2[:, 4] <= np.pi)


# ---------------------------------------------------------------------------
# Reusable solver handles (SpectralSolver / MagneticSolver) and the
# run-time switch of the kept distribution
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_solver_handles_reused_match_oracle(oracle, precision):
    """One SpectralSolver and one MagneticSolver (spectral_solver.rs:33-68,
    magnetic_solver.rs:32-62) serve three distributions and screenings in a
    row: each result equals the one-shot solve bit for bit and the oracle
    within the precision's bar."""
    p = ss.make_params(box=(10.0, 12.0, 9.0), grid=(16, 12, 10, 6, 5), stress_active=-1.3, stress_magnetic=0.8,
                       shape=0.4, precision=precision)
    tol = FP64_TOL if precision == ss.FP64 else FP32_FIELD_TOL
    fs = ss.SpectralSolver(p)
    bs = ss.MagneticSolver(p)
    rng = np.random.default_rng(7)
    for screening in (0.0, 0.25, 1.5):
        dist = rng.random(p.dist_shape) * 3.0
        u, g = fs.mean_flow_field(screening, dist)
        ue, ge = oracle.mean_flow_field(p, screening, dist)
        assert rel_err(u, ue) < tol and rel_err(g, ge) < tol
        p.hydro_screening = screening
        u1, g1 = ss.mean_flow_field(p, dist)
        np.testing.assert_array_equal(u, u1)
        np.testing.assert_array_equal(g, g1)
        b, gb = bs.mean_magnetic_field(dist)
        be, gbe = oracle.mean_magnetic_field(p, dist)
        assert rel_err(b, be) < tol and rel_err(gb, gbe) < tol
    fs.close()
    bs.close()


def test_keep_distribution_switched_at_run_time(oracle):
    """The FP32 tiled path without SS_FLAG_KEEP_DISTRIBUTION has no count grid;
    ss_keep_distribution turns it on between steps: at once the counts of
    the current particles (the reference's binning rule in binary32), then
    after every step the same counts and particles as a run keeping the
    grid from its creation; turned off, the steps go on unchanged."""
    kw = dict(grid=(16, 16, 16, 8, 8), n=1 << 14, precision=ss.FP32, resort_interval=4)
    parts = ss.initial_condition(small_params(**kw), amplitude=0.05, seed=14)
    a = ss.Simulation(small_params(**kw))
    b = ss.Simulation(small_params(flags=ss.FLAG_KEEP_DISTRIBUTION, **kw))
    for sim in (a, b):
        sim.init(parts)
        sim.step(5)
    with pytest.raises(ss.SSError, match="ss_keep_distribution"):
        a.get_counts()
    a.keep_distribution(True)
    p = small_params(**kw)
    cur = a.get_particles().astype(np.float32)
    exp = np.zeros(p.dist_shape, dtype=np.uint32)
    for q in cur:
        exp[tuple(oracle.coord_to_grid_f32(p, q))] += 1
    np.testing.assert_array_equal(a.get_counts(), exp)
    for _ in range(3):
        a.step(1)
        b.step(1)
        np.testing.assert_array_equal(a.get_counts(), b.get_counts())
        np.testing.assert_array_equal(a.get_particles(), b.get_particles())
    a.keep_distribution(False)
    with pytest.raises(ss.SSError):
        a.get_counts()
    a.step(9)
    b.step(9)
    np.testing.assert_array_equal(a.get_particles(), b.get_particles())


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_stepping_in_chunks_equals_one_call(precision):
    """Runs of steps inside a re-sort period replay captured graphs once a
    caller repeats them (the first request of each phase, length and buffer
    state runs eagerly); the bookkeeping of a replay (the re-sort a run
    starts with, the next re-sort's offsets a run ends with) must match
    eager stepping: 45 steps in chunks of 3 and 5 with a re-sort every 8
    steps give the same particles and fields bit for bit as one call."""
    p = ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=1 << 14, seed=7, precision=precision,
                       binning=ss.BIN_SORTED, resort_interval=8)
    parts = ss.initial_condition(ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=1 << 14),
                                 amplitude=0.05, seed=7)
    a = ss.Simulation(p)
    a.init(parts)
    a.step(45)
    b = ss.Simulation(p)
    b.init(parts)
    chunks = [3] * 5 + [5] * 3 + [3] * 5
    assert sum(chunks) == 45
    for c in chunks:
        b.step(c)
    assert a.get_timestep() == b.get_timestep() == 45
    np.testing.assert_array_equal(a.get_particles(), b.get_particles())
    ua, ga = a.get_flow_field(with_gradient=True)
    ub, gb = b.get_flow_field(with_gradient=True)
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(ga, gb)


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_more_segment_shapes_than_kept(precision):
    """Four ways of splitting a re-sort period of 16 steps, each used for
    four periods in a row, make about thirty distinct runs (phase, length,
    buffer state) that are asked for twice and captured, against eight
    captured segments kept, so most captured segments are evicted again;
    256 steps in such chunks equal one call bit for bit."""
    p = ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=1 << 14, seed=5, precision=precision,
                       binning=ss.BIN_SORTED, resort_interval=16)
    parts = ss.initial_condition(ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=1 << 14),
                                 amplitude=0.05, seed=5)
    a = ss.Simulation(p)
    a.init(parts)
    a.step(256)
    b = ss.Simulation(p)
    b.init(parts)
    splits = ([1, 2, 3, 4, 6], [6, 4, 3, 2, 1], [5, 5, 5, 1], [2, 7, 7])
    chunks = [c for split in splits for _ in range(4) for c in split]
    assert sum(chunks) == 256
    for c in chunks:
        b.step(c)
    assert a.get_timestep() == b.get_timestep() == 256
    np.testing.assert_array_equal(a.get_particles(), b.get_particles())
    ua, ga = a.get_flow_field(with_gradient=True)
    ub, gb = b.get_flow_field(with_gradient=True)
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(ga, gb)


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_prepared_steps_equal_unprepared(precision):
    """ss_prepare_steps captures the graphs of the next steps without running
    anything: the state it leaves is the state it found (the timestep, the
    particles), and the prepared steps give the same particles and fields
    bit for bit as unprepared ones, from a pending re-sort and from the
    middle of a period, across a period boundary."""
    p = ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=1 << 14, seed=9, precision=precision,
                       binning=ss.BIN_SORTED, resort_interval=8)
    parts = ss.initial_condition(ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=1 << 14),
                                 amplitude=0.05, seed=9)
    runs = []
    for prepare in (False, True):
        sim = ss.Simulation(p)
        sim.init(parts)
        sim.step(3)
        for n in (6, 11):
            if prepare:
                before = sim.get_particles()
                sim.prepare_steps(n)
                assert sim.get_timestep() in (3, 9)
                np.testing.assert_array_equal(sim.get_particles(), before)
            sim.step(n)
        sim.resort()
        if prepare:
            sim.prepare_steps(20)
        sim.step(20)
        u, g = sim.get_flow_field(with_gradient=True)
        runs.append((sim.get_timestep(), sim.get_particles(), u, g))
    (ta, pa, ua, ga), (tb, pb, ub, gb) = runs
    assert ta == tb == 40
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(ga, gb)

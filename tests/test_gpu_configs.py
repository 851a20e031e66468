"""The workloads of the benchmark table on the GPU against the FP64 CPU oracle.

* 16^3 x 8 x 8, L = 10, 2^16 particles, FP64, 100 steps (the CPU-runnable case);
* 64^3 x 16 x 16, L = 50, FP64, at 2^18 particles against the oracle and at
  the full 2^24 particles through size-independent properties;
* the same geometry in FP32 (compensated positions, FP32 FFT) at 2^18
  particles, with stated bounds on fields and states;
* a batched ensemble over the 32^3 x 8 x 8, 48^3 x 12 x 12 and 64^3 x 16 x 16
  classes (box 0.78 x grid points, alpha and D_r spread) against the oracle;
* every optional term of the Langevin chain and of the stress in whole FP64
  steps, on the plain and the tiled path;
* a 32 x 32 angle grid whose kernels need more than 64 KiB of LDS.

Reference: Simulation::do_timestep, src/bin/simulation/mod.rs:251-374.
Bars: counts bit-exact, FP64 fields <= 1e-10 relative, FP64 states <= 1e-9;
FP32 fields <= 1e-5 relative of the FP64 solve of the same distribution.
"""
import math

import numpy as np
import pytest

import stochasticsampling_amd as ss
from helpers import periodic_diff, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_FIELD_TOL = 1e-5


def _params(box, grid, n, seed, **kw):
    return ss.make_params(box=box, grid=grid, n=n, seed=seed, **kw)


def _state_diffs(a, b, box):
    """Per-particle distances: positions (periodic), phi (periodic), theta."""
    d = periodic_diff(a[:, :3], b[:, :3], np.asarray(box)[None, :])
    dphi = periodic_diff(a[:, 3], b[:, 3], 2 * np.pi)
    dth = np.abs(a[:, 4] - b[:, 4])
    return d, dphi, dth


def _assert_fp64_states(a, b, box, tol=1e-9):
    d, dphi, dth = _state_diffs(a, b, box)
    # phi is ill-defined at the poles: compare the orientation vectors there
    na = np.stack([np.sin(a[:, 4]) * np.cos(a[:, 3]), np.sin(a[:, 4]) * np.sin(a[:, 3]), np.cos(a[:, 4])], 1)
    nb = np.stack([np.sin(b[:, 4]) * np.cos(b[:, 3]), np.sin(b[:, 4]) * np.sin(b[:, 3]), np.cos(b[:, 4])], 1)
    assert d.max() < tol, d.max()
    assert dth.max() < tol, dth.max()
    assert np.abs(na - nb).max() < tol, np.abs(na - nb).max()
    pole = np.sin(b[:, 4]) < 1e-6
    assert dphi[~pole].max() < tol * 10, dphi[~pole].max()


def _run_pair(oracle, p, parts, steps, oracle_params=None):
    sim = ss.Simulation(p)
    sim.init(parts)
    osim = oracle.simulation(oracle_params or p)
    osim.init(parts)
    sim.step(steps)
    osim.step(steps)
    return sim, osim


# ---------------------------------------------------------------------------
# 16^3 x 8 x 8, L = 10, 2^16 particles, FP64, 100 steps
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("binning", [ss.BIN_HISTOGRAM, ss.BIN_SORTED, ss.BIN_PLAIN_TILED])
def test_small_case_fp64_100_steps(oracle, binning):
    box = (10.0, 10.0, 10.0)
    p = _params(box, (16, 16, 16, 8, 8), 1 << 16, 1, binning=binning, flags=ss.FLAG_KEEP_DISTRIBUTION)
    parts = ss.initial_condition(p, amplitude=0.05, seed=1)
    sim, osim = _run_pair(oracle, p, parts, 100, _params(box, (16, 16, 16, 8, 8), 1 << 16, 1))
    assert sim.get_timestep() == osim.timestep == 100
    np.testing.assert_array_equal(sim.get_counts(), osim.counts())
    u, g = sim.get_flow_field(with_gradient=True)
    ue, ge = osim.flow_field()
    assert rel_err(u, ue) < FP64_TOL
    assert rel_err(g, ge) < FP64_TOL
    _assert_fp64_states(sim.get_particles(), osim.particles(), box)


# ---------------------------------------------------------------------------
# 64^3 x 16 x 16, L = 50
# ---------------------------------------------------------------------------
BIG_BOX = (50.0, 50.0, 50.0)
BIG_GRID = (64, 64, 64, 16, 16)


@pytest.mark.parametrize("binning", [ss.BIN_HISTOGRAM, ss.BIN_SORTED, ss.BIN_PLAIN_TILED])
def test_large_grid_fp64_matches_oracle(oracle, binning):
    """FP64 at 64^3 x 16 x 16, L = 50, 2^18 particles, 3 steps."""
    n = 1 << 18
    p = _params(BIG_BOX, BIG_GRID, n, 2, binning=binning, flags=ss.FLAG_KEEP_DISTRIBUTION)
    parts = ss.initial_condition(p, amplitude=0.05, seed=2)
    sim, osim = _run_pair(oracle, p, parts, 3, _params(BIG_BOX, BIG_GRID, n, 2))
    np.testing.assert_array_equal(sim.get_counts(), osim.counts())
    u, g = sim.get_flow_field(with_gradient=True)
    ue, ge = osim.flow_field()
    assert rel_err(u, ue) < FP64_TOL
    assert rel_err(g, ge) < FP64_TOL
    _assert_fp64_states(sim.get_particles(), osim.particles(), BIG_BOX)


def test_large_grid_fp32_matches_oracle(oracle):
    """FP32 (compensated positions, FP32 FFT) at 64^3 x 16 x 16, L = 50, 2^18
    particles, 3 steps against the FP64 oracle from the same initial state and
    Philox noise.  Bounds: u and grad u of every step within 1e-5 of the FP64
    solve of the FP32 run's own distribution; counts differing from the FP64
    run's in at most 0.1% of the particles (FP32-rounded positions next to
    cell edges); positions within 1e-6 L at the 99th percentile and 1e-4 L at
    most, orientation vectors within 1e-5 at the 99th percentile."""
    n = 1 << 18
    p32 = _params(BIG_BOX, BIG_GRID, n, 2, precision=ss.FP32, flags=ss.FLAG_KEEP_DISTRIBUTION)
    p64 = _params(BIG_BOX, BIG_GRID, n, 2)
    parts = ss.initial_condition(p64, amplitude=0.05, seed=2)
    sim = ss.Simulation(p32)
    sim.init(parts)
    osim = oracle.simulation(p64)
    osim.init(parts)
    for _ in range(3):
        sim.step(1)
        osim.step(1)
        u, g = sim.get_flow_field(with_gradient=True)
        ue, ge = oracle.mean_flow_field(p64, 0.0, sim.get_distribution())
        assert rel_err(u, ue) < FP32_FIELD_TOL
        assert rel_err(g, ge) < FP32_FIELD_TOL
        moved = int(np.abs(sim.get_counts().astype(np.int64) - osim.counts().astype(np.int64)).sum())
        assert moved <= n // 1000, moved
    a, b = sim.get_particles(), osim.particles()
    d, _, _ = _state_diffs(a, b, BIG_BOX)
    assert np.percentile(d, 99) < 1e-6 * 50.0 and d.max() < 1e-4 * 50.0
    na = np.stack([np.sin(a[:, 4]) * np.cos(a[:, 3]), np.sin(a[:, 4]) * np.sin(a[:, 3]), np.cos(a[:, 4])], 1)
    nb = np.stack([np.sin(b[:, 4]) * np.cos(b[:, 3]), np.sin(b[:, 4]) * np.sin(b[:, 3]), np.cos(b[:, 4])], 1)
    dn = np.abs(na - nb).max(axis=1)
    assert np.percentile(dn, 99) < 1e-5, np.percentile(dn, 99)


def test_full_size_fp64_paths_agree():
    """2^24 particles on 64^3 x 16 x 16, FP64: the plain histogram path, the
    tiled path and the plain path with tile-ordered particles bin the same
    counts and move the particles alike (the moments differ only in summation
    order; both plain paths contract the same count grid with the same kernel,
    so they agree bit for bit), and the size-independent
    properties hold: every particle binned, u of zero mean and divergence
    free in k space, states inside the box and the angle ranges."""
    n = 1 << 24
    p = _params(BIG_BOX, BIG_GRID, n, 2, flags=ss.FLAG_KEEP_DISTRIBUTION)
    parts = ss.initial_condition(p, amplitude=0.05, seed=2)
    res = []
    for binning in (ss.BIN_HISTOGRAM, ss.BIN_SORTED, ss.BIN_PLAIN_TILED):
        print(f"full size FP64 binning={binning}", flush=True)
        sim = ss.Simulation(_params(BIG_BOX, BIG_GRID, n, 2, binning=binning, flags=ss.FLAG_KEEP_DISTRIBUTION))
        sim.init(parts)
        sim.step(2)
        res.append((sim.get_counts(), sim.get_flow_field(), sim.get_particles()))
        sim.close()
    (c0, u0, x0), (c1, u1, x1), (c2, u2, x2) = res
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_array_equal(c2, c0)
    np.testing.assert_array_equal(u2, u0)
    np.testing.assert_array_equal(x2, x0)
    assert int(c0.sum(dtype=np.uint64)) == n
    # the tiled path sums kernel values rounded to multiples of 2^-37 (the
    # fixed-point scale that keeps 2^24 particle values below 2^62): about
    # 4e-12 of u against the FP64 contraction of the counts, inside the FP64 bar
    assert rel_err(u1, u0) < FP64_TOL
    d, _, dth = _state_diffs(x0, x1, BIG_BOX)
    assert d.max() < 1e-11 and dth.max() < 1e-11
    assert abs(u0.mean()) < 1e-12 * np.abs(u0).max()
    uh = np.fft.fftn(u0, axes=(1, 2, 3))
    k = 2 * np.pi * np.fft.fftfreq(64, d=50.0 / 64)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    inner = (np.abs(KX) < 0.9 * np.abs(k).max()) & (np.abs(KY) < 0.9 * np.abs(k).max()) & (
        np.abs(KZ) < 0.9 * np.abs(k).max())
    div = KX * uh[0] + KY * uh[1] + KZ * uh[2]
    assert np.abs(div[inner]).max() < 1e-10 * np.abs(uh).max() * np.abs(k).max()
    assert np.all(x0[:, :3] >= 0) and np.all(x0[:, :3] <= 50.0)
    assert np.all(x0[:, 3] >= 0) and np.all(x0[:, 3] <= 2 * np.pi)
    assert np.all(x0[:, 4] >= 0) and np.all(x0[:, 4] <= np.pi)


# ---------------------------------------------------------------------------
# Batched ensemble over the three grid classes
# ---------------------------------------------------------------------------
ENSEMBLE_SPECS = [  # (grid, particles, alpha, D_r, seed)
    ((32, 32, 32, 8, 8), 1 << 14, -2.0, 0.1, 100),
    ((32, 32, 32, 8, 8), 23170, 1.2, 1.0, 101),
    ((48, 48, 48, 12, 12), 1 << 15, -0.4, 0.5, 102),
    ((48, 48, 48, 12, 12), 46341, 2.0, 0.1, 103),
    ((64, 64, 64, 16, 16), 1 << 16, 0.8, 1.0, 104),
    ((64, 64, 64, 16, 16), 92682, -1.2, 0.5, 105),
]


def _ensemble_params(precision, flags=0):
    out = []
    for grid, n, alpha, dr, seed in ENSEMBLE_SPECS:
        L = 0.78 * grid[0]
        out.append(_params((L, L, L), grid, n, seed, stress_active=alpha, diffusion_rotational=dr,
                           precision=precision, binning=ss.BIN_SORTED, flags=flags))
    return out


def test_ensemble_grid_classes_fp64_match_oracle(oracle):
    """FP64 ensemble members (32^3 x 8 x 8, 48^3 x 12 x 12, 64^3 x 16 x 16;
    L = 0.78 n; alpha in [-2, 2]; D_r in {0.1, 0.5, 1.0}) against the oracle
    for 3 steps: fields <= 1e-10, states <= 1e-9."""
    plist = _ensemble_params(ss.FP64)
    ens = ss.Ensemble(plist)
    parts = [ss.initial_condition(p, amplitude=0.05, seed=p.seed) for p in plist]
    for m, x in zip(ens.members, parts):
        m.init(x)
    assert ens.step(3) == 3
    for p, m, x in zip(plist, ens.members, parts):
        osim = oracle.simulation(_params(p.box, p.dist_shape, p.number_of_particles,
                                         p.seed, stress_active=p.stress_active,
                                         diffusion_rotational=p.diffusion_rotational))
        osim.init(x)
        osim.step(3)
        u, g = m.get_flow_field(with_gradient=True)
        ue, ge = osim.flow_field()
        assert rel_err(u, ue) < FP64_TOL
        assert rel_err(g, ge) < FP64_TOL
        _assert_fp64_states(m.get_particles(), osim.particles(), p.box)


def test_ensemble_grid_classes_fp32_match_oracle(oracle):
    """FP32 ensemble members over the three grid classes: each member equals
    the same system stepped alone bit for bit, and that system's u and grad u
    of each of 2 steps are within 1e-5 of the FP64 oracle solve of its own
    distribution; states after 2 steps against the FP64 oracle run within
    1e-6 L (99th percentile)."""
    plist = _ensemble_params(ss.FP32)
    ens = ss.Ensemble(plist)
    parts = [ss.initial_condition(p, amplitude=0.05, seed=p.seed) for p in plist]
    for m, x in zip(ens.members, parts):
        m.init(x)
    singles = []
    for p, x in zip(_ensemble_params(ss.FP32, ss.FLAG_KEEP_DISTRIBUTION), parts):
        sim = ss.Simulation(p)
        sim.init(x)
        singles.append(sim)
    oracles = []
    for p, x in zip(plist, parts):
        o64 = _params(p.box, p.dist_shape, p.number_of_particles, p.seed,
                      stress_active=p.stress_active, diffusion_rotational=p.diffusion_rotational)
        osim = oracle.simulation(o64)
        osim.init(x)
        oracles.append((o64, osim))
    for step in range(2):
        ens.step(1)
        for m, sim, (o64, osim) in zip(ens.members, singles, oracles):
            sim.step(1)
            osim.step(1)
            np.testing.assert_array_equal(m.get_particles(), sim.get_particles())
            u, g = sim.get_flow_field(with_gradient=True)
            np.testing.assert_array_equal(m.get_flow_field(), u)
            ue, ge = oracle.mean_flow_field(o64, 0.0, sim.get_distribution())
            assert rel_err(u, ue) < FP32_FIELD_TOL, (o64.dist_shape, step)
            assert rel_err(g, ge) < FP32_FIELD_TOL, (o64.dist_shape, step)
    for p, sim, (o64, osim) in zip(plist, singles, oracles):
        d, _, _ = _state_diffs(sim.get_particles(), osim.particles(), p.box)
        L = p.box[0]
        assert np.percentile(d, 99) < 1e-6 * L and d.max() < 1e-4 * L


# ---------------------------------------------------------------------------
# Every optional term in whole FP64 steps
# ---------------------------------------------------------------------------
OPTIONAL = dict(shape=0.6, volume_exclusion=0.3, hydro_screening=0.2, stress_magnetic=0.7,
                magnetic_reorientation=1.5, magnetic_dipole_dipole=0.8, magnetic_drag=0.4)


@pytest.mark.parametrize("binning", [ss.BIN_HISTOGRAM, ss.BIN_SORTED, ss.BIN_PLAIN_TILED])
@pytest.mark.parametrize("variant", ["y_field", "z_field", "small_alpha", "symmetric_stress"])
def test_optional_physics_whole_steps_fp64(oracle, binning, variant):
    """The Langevin chain of mod.rs:323-363 with Jeffery strain, density-
    dependent diffusion, screening, magnetic stress, field torque along the
    reference's y axis or a given z axis, dipole-dipole torque and drag, and
    the stress of mod.rs:100-104 with all three kernels: counts bit-exact,
    u, grad u, B, grad B <= 1e-10 relative, all five state columns <= 1e-9.
    "small_alpha" runs active stress 1e-3 alone (the fixed-point moments of
    the tiled path must keep their resolution for small coefficients);
    "symmetric_stress" leaves the magnetic stress out, so that the tiled path
    solves the flow and the magnetic field in its fused spectral stage."""
    box = (10.0, 9.0, 8.0)
    grid = (12, 10, 8, 6, 6)
    n = 6000
    if variant == "small_alpha":
        kw = dict(stress_active=1e-3)
    else:
        kw = dict(OPTIONAL)
        if variant == "z_field":
            kw["field_direction"] = (0.0, 0.0, 1.0)
        if variant == "symmetric_stress":
            del kw["stress_magnetic"]
    p = _params(box, grid, n, 5, binning=binning, flags=ss.FLAG_KEEP_DISTRIBUTION, **kw)
    parts = ss.initial_condition(_params(box, grid, n, 5), amplitude=0.2, seed=5)
    sim, osim = _run_pair(oracle, p, parts, 4, _params(box, grid, n, 5, **kw))
    if variant == "symmetric_stress" and binning == ss.BIN_SORTED:
        assert sim.get_info()["fused_spectral"]
    np.testing.assert_array_equal(sim.get_counts(), osim.counts())
    u, g = sim.get_flow_field(with_gradient=True)
    ue, ge = osim.flow_field()
    assert rel_err(u, ue) < FP64_TOL
    assert rel_err(g, ge) < FP64_TOL
    if variant != "small_alpha":
        b, gb = sim.get_magnetic_field(with_gradient=True)
        be, gbe = osim.magnetic_field()
        assert rel_err(b, be) < FP64_TOL
        assert rel_err(gb, gbe) < FP64_TOL
    _assert_fp64_states(sim.get_particles(), osim.particles(), box)


# ---------------------------------------------------------------------------
# Angle grids whose LDS needs exceed 64 KiB per workgroup
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("binning", [ss.BIN_HISTOGRAM, ss.BIN_SORTED, ss.BIN_PLAIN_TILED])
def test_fine_angle_grid_beyond_64k_lds_fp64(oracle, binning):
    """32 x 32 angle bins with magnetic stress (nine moment components): the
    tiled update's window and kernels take 91 KiB of LDS, the contraction's
    kernels 72 KiB, so both run with the raised per-kernel limit (gfx950:
    160 KiB) instead of being refused or falling back.  Whole FP64 steps
    against the oracle: counts exact, fields <= 1e-10, states <= 1e-9."""
    box = (8.0, 8.0, 8.0)
    grid = (8, 8, 8, 32, 32)
    n = 4096
    kw = dict(stress_magnetic=0.7)
    p = _params(box, grid, n, 9, binning=binning, flags=ss.FLAG_KEEP_DISTRIBUTION, **kw)
    sim = ss.Simulation(p)
    info = sim.get_info()
    assert info["sorted"] == (binning == ss.BIN_SORTED)
    assert info["plain_tiled"] == (binning == ss.BIN_PLAIN_TILED)
    sim.close()
    parts = ss.initial_condition(_params(box, grid, n, 9), amplitude=0.2, seed=9)
    sim, osim = _run_pair(oracle, p, parts, 3, _params(box, grid, n, 9, **kw))
    np.testing.assert_array_equal(sim.get_counts(), osim.counts())
    u, g = sim.get_flow_field(with_gradient=True)
    ue, ge = osim.flow_field()
    assert rel_err(u, ue) < FP64_TOL
    assert rel_err(g, ge) < FP64_TOL
    _assert_fp64_states(sim.get_particles(), osim.particles(), box)

"""GPU parity of the reference's own random streams (SS_NOISE_REFERENCE and
the initial conditions) against the oracle, through the C ABI.

Bars: generator states after every step bit-exact; the normals and the axis
angle bit-exact except for normals from the ziggurat's tail (about 2.6e-4 of
them), whose value goes through log and may differ from the host's libm in
the last bits (at most 4 ulp); the rotate angle dr sqrt(-2 ln(1 - U)) within
4 ulp for the same reason.  Initial conditions: positions and phi bit-exact,
theta within 4 ulp (acos), the rotated orientations of the homogeneous state
within the conditioning of its polar-angle formula (a few ulp of exp and sinh
amplified by the cancellation in exp(kappa) - 2 U sinh(kappa)).
"""
import math

import numpy as np
import pytest

import stochasticsampling_amd as ss
from helpers import isotropic_particles, periodic_diff, small_params

pytestmark = pytest.mark.gpu


def ulps(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    ia, ib = a.view(np.int64), b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-(2**63)) - ia, ia)
    ib = np.where(ib < 0, np.int64(-(2**63)) - ib, ib)
    return np.abs(ia - ib)


def ulps32(a, b):
    ia = np.ascontiguousarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    ib = np.ascontiguousarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(2**31) - ia, ia)
    ib = np.where(ib < 0, -(2**31) - ib, ib)
    return np.abs(ia - ib)


def _check_noise(got, want, R, fp32=False):
    d = ulps32 if fp32 else ulps
    tail = np.abs(want[:, :3]) > R
    dn = d(got[:, :3], want[:, :3])
    assert np.all(dn[~tail] == 0), f"non-tail normals differ: {np.count_nonzero(dn[~tail])}"
    assert np.all(dn[tail] <= 4)
    assert np.all(d(got[:, 3], want[:, 3]) == 0)
    assert np.all(d(got[:, 4], want[:, 4]) <= 4)


def _oracle_states(oracle, seed, t):
    return [oracle.pcg32_seed_from_u64(seed) for _ in range(t)]


@pytest.mark.parametrize(
    "n,t,binning,precision,resort",
    [
        (1000, 3, ss.BIN_HISTOGRAM, ss.FP64, 0),
        (5, 8, ss.BIN_HISTOGRAM, ss.FP64, 0),
        (50_000, 1, ss.BIN_HISTOGRAM, ss.FP64, 0),
        (20_000, 7, ss.BIN_SORTED, ss.FP64, 0),
        (20_000, 7, ss.BIN_PLAIN_TILED, ss.FP64, 1),
        (20_000, 5, ss.BIN_PLAIN_TILED, ss.FP64, 3),
        (20_000, 7, ss.BIN_SORTED, ss.FP32, 1),
        (1 << 20, 16, ss.BIN_SORTED, ss.FP32, 2),
    ],
)
def test_reference_noise_matches_oracle(oracle, n, t, binning, precision, resort):
    """Every step's RandomVectors and generator states; on the tiled path
    the noise records follow the storage order, which re-sorts change (every
    step with resort 1)."""
    print(f"reference noise n={n} T={t} binning={binning} precision={precision} resort={resort}", flush=True)
    p = small_params(n=n, binning=binning, precision=precision, noise_mode=ss.NOISE_REFERENCE, noise_streams=t,
                     seed=2024, resort_interval=resort)
    dr = math.sqrt(2.0 * p.diffusion_rotational * p.timestep)
    R = oracle.zig_tables()[0][1]
    states = _oracle_states(oracle, p.seed, t)
    sim = ss.Simulation(p)
    sim.init(isotropic_particles(n, (10.0, 10.0, 10.0), 1))
    st0 = sim.get_rng_state()
    assert st0.shape == (t, 2)
    assert all((int(a), int(b)) == (s.state, s.inc) for (a, b), s in zip(st0, states))
    for step in range(4):
        sim.step(1)
        want = oracle.ref_noise(states, n, dr)
        got = sim.get_noise()
        _check_noise(got, want.astype(np.float32) if precision == ss.FP32 else want, R,
                     fp32=precision == ss.FP32)
        st = sim.get_rng_state()
        assert [(int(a), int(b)) for a, b in st] == [(s.state, s.inc) for s in states], f"step {step}"


def test_reference_noise_graph_replays(oracle):
    """Whole re-sort periods replayed from captured graphs draw the same
    stream as eager steps."""
    n, t = 30_000, 5
    base = dict(n=n, binning=ss.BIN_SORTED, precision=ss.FP64, noise_mode=ss.NOISE_REFERENCE, noise_streams=t,
                resort_interval=4, seed=9)
    parts = isotropic_particles(n, (10.0, 10.0, 10.0), 2)
    a = ss.Simulation(small_params(**base))
    b = ss.Simulation(small_params(flags=ss.FLAG_NO_GRAPH, **base))
    a.init(parts)
    b.init(parts)
    a.step(8)
    b.step(8)
    np.testing.assert_array_equal(a.get_rng_state(), b.get_rng_state())
    np.testing.assert_array_equal(a.get_noise(), b.get_noise())
    np.testing.assert_array_equal(a.get_particles(), b.get_particles())
    states = _oracle_states(oracle, 9, t)
    dr = math.sqrt(2.0 * a.params.diffusion_rotational * a.params.timestep)
    for _ in range(8):
        oracle.ref_noise(states, n, dr)
    assert [(int(x), int(y)) for x, y in a.get_rng_state()] == [(s.state, s.inc) for s in states]


def test_reference_noise_trajectory_fp64(oracle):
    p = small_params(n=4096, noise_mode=ss.NOISE_REFERENCE, noise_streams=4, flags=ss.FLAG_KEEP_DISTRIBUTION)
    parts = ss.initial_condition(p, amplitude=0.05, seed=1)
    sim = ss.Simulation(p)
    sim.init(parts)
    osim = oracle.simulation(p)
    osim.init(parts)
    sim.step(4)
    osim.step(4)
    np.testing.assert_array_equal(sim.get_counts(), osim.counts())
    a, b = sim.get_particles(), osim.particles()
    assert np.max(periodic_diff(a[:, :3], b[:, :3], 10.0)) < 1e-9
    assert np.max(periodic_diff(a[:, 3], b[:, 3], 2 * np.pi)) < 1e-9
    assert np.max(np.abs(a[:, 4] - b[:, 4])) < 1e-9


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_reference_noise_snapshot_resume(precision):
    """Snapshot with rng_state, resume in a fresh simulation: the next steps
    equal the uninterrupted run (mod.rs:199-217)."""
    p = small_params(n=10_000, precision=precision, noise_mode=ss.NOISE_REFERENCE, noise_streams=6)
    parts = isotropic_particles(10_000, (10.0, 10.0, 10.0), 5)
    a = ss.Simulation(p)
    a.init(parts)
    a.step(3)
    snap = a.get_snapshot()
    assert snap.rng_state.shape == (6, 2)
    a.step(2)
    b = ss.Simulation(p)
    b.resume(snap)
    b.step(2)
    np.testing.assert_array_equal(a.get_noise(), b.get_noise())
    np.testing.assert_array_equal(a.get_particles(), b.get_particles())
    np.testing.assert_array_equal(a.get_rng_state(), b.get_rng_state())


def test_reference_noise_validation():
    p = small_params(n=100, noise_mode=ss.NOISE_REFERENCE, noise_streams=0)
    with pytest.raises(ss.SSError, match="RAYON_NUM_THREADS"):
        ss.Simulation(p)
    q = small_params(n=100)
    s = ss.Simulation(q)
    assert s.get_rng_state().shape == (0, 2)
    with pytest.raises(ss.SSError):
        s.get_noise()


@pytest.mark.parametrize("n,seed", [(1, 0), (1000, 1), (100_003, 77)])
def test_create_isotropic_matches_oracle(oracle, n, seed):
    box = (3.0, 5.0, 7.5)
    got = ss.create_isotropic(n, box, seed)
    want = oracle.create_isotropic(n, box, seed)
    np.testing.assert_array_equal(got[:, :4], want[:, :4])
    assert np.all(ulps(got[:, 4], want[:, 4]) <= 4)


@pytest.mark.parametrize("kappa", [0.5, 3.0, 10.0])
def test_create_homogeneous_matches_oracle(oracle, kappa):
    n, box = 50_000, (2.0, 2.0, 2.0)
    got = ss.create_homogeneous(n, kappa, box, 3)
    want = oracle.create_homogeneous(n, kappa, box, 3)
    np.testing.assert_array_equal(got[:, :3], want[:, :3])

    def vec(p):
        return np.stack([np.sin(p[:, 4]) * np.cos(p[:, 3]), np.sin(p[:, 4]) * np.sin(p[:, 3]), np.cos(p[:, 4])], 1)

    # cos(theta0) = ln(exp(kappa) - 2 U sinh(kappa)) / kappa cancels: a last-bit
    # difference of exp or sinh between two libms moves theta0 by about
    # 2^-52 e^kappa / (kappa exp(kappa n_y) sin(theta0)), n_y = cos(theta0)
    ny = vec(want)[:, 1]
    bound = 8 * 2.0**-52 * math.exp(kappa) / (kappa * np.exp(kappa * ny) * np.sqrt(np.maximum(1 - ny * ny, 1e-300)))
    assert np.all(np.max(np.abs(vec(got) - vec(want)), axis=1) <= bound + 1e-14)
    with pytest.raises(ss.SSError, match="Alignment of zero"):
        ss.create_homogeneous(10, 0.0, box, 3)

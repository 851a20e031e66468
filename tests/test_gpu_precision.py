"""Precision of the FP32 path (compensated positions, FP32 arithmetic) against
FP64: the compensated position sum, and the noise both precisions draw.

Reference: translational step and finalize of the LangevinBuilder chain
(src/bin/simulation/mod.rs:346-363, langevin_builder/mod.rs:114-125).
"""
import numpy as np
import pytest

import stochasticsampling_amd as ss
from helpers import periodic_diff

pytestmark = pytest.mark.gpu

L = 10.0


def _unit(p):
    st = np.sin(p[:, 4])
    return np.stack([st * np.cos(p[:, 3]), st * np.sin(p[:, 3]), np.cos(p[:, 4])], 1)


def test_compensated_positions_deterministic_drift():
    """With D_t = D_r = alpha = 0 every particle moves by dt n per step along
    a fixed orientation.  After 1000 steps the FP32 positions (high + low
    parts) stay within 2e-5 of the FP64 path (measured: 3e-6 at most); an
    FP32 sum of the same increments without the low parts, emulated here,
    drifts by more than 2e-5 for most particles (measured: 1.8e-4 median), so
    the bound fails without the compensation."""
    kw = dict(box=(L, L, L), grid=(16, 16, 16, 8, 8), n=1 << 14, seed=3, diffusion_translational=0.0,
              diffusion_rotational=0.0, stress_active=0.0)
    parts = ss.initial_condition(ss.make_params(**kw), amplitude=0.0, seed=3)
    out = {}
    for prec, binning in ((ss.FP32, ss.BIN_SORTED), (ss.FP64, ss.BIN_HISTOGRAM)):
        sim = ss.Simulation(ss.make_params(precision=prec, binning=binning, **kw))
        sim.init(parts)
        sim.step(1000)
        out[prec] = sim.get_particles()
    a, b = out[ss.FP32], out[ss.FP64]
    d = periodic_diff(a[:, :3], b[:, :3], L)
    assert d.max() < 2e-5, d.max()
    assert np.abs(_unit(a) - _unit(b)).max() < 2e-5
    # the same increments summed in FP32 without the low parts
    n0 = _unit(parts).astype(np.float32)
    x = parts[:, :3].astype(np.float32)
    inc = (np.float32(1e-3) * n0).astype(np.float32)
    for _ in range(1000):
        x = (x + inc).astype(np.float32)
        x = np.where(x >= np.float32(L), x - np.float32(L), x).astype(np.float32)
        x = np.where(x < 0, x + np.float32(L), x).astype(np.float32)
    e = periodic_diff(x.astype(np.float64), b[:, :3], L).max(axis=1)
    assert np.median(e) > 2e-5, np.median(e)


def test_fp32_noise_equals_fp64_noise():
    """Both precisions draw the same 52-bit uniforms: with the flow off
    (alpha = 0) one step moves every particle by dt n plus the same
    translational increment and rotation, so FP32 and FP64 agree for every
    particle to FP32 rounding (no Gaussian tail is cut in either); 2^20
    particles put normals beyond 4.5 sigma into the comparison."""
    n = 1 << 20
    kw = dict(box=(L, L, L), grid=(16, 16, 16, 8, 8), n=n, seed=11, stress_active=0.0, diffusion_translational=0.3,
              diffusion_rotational=2.0)
    parts = ss.initial_condition(ss.make_params(**kw), amplitude=0.0, seed=11)
    out = {}
    for prec, binning in ((ss.FP32, ss.BIN_SORTED), (ss.FP64, ss.BIN_HISTOGRAM)):
        sim = ss.Simulation(ss.make_params(precision=prec, binning=binning, **kw))
        sim.init(parts)
        sim.step(1)
        out[prec] = sim.get_particles()
    a, b = out[ss.FP32], out[ss.FP64]
    d = periodic_diff(a[:, :3], b[:, :3], L).max(axis=1)
    amp = np.sqrt(2 * 1e-3 * 0.3)
    # largest translational increment of the step (in units of its amplitude)
    big = np.abs(b[:, :3] - parts[:, :3])
    big = np.minimum(big, L - big).max(axis=1) / amp
    assert big.max() > 4.5, big.max()
    assert d.max() < 2e-6, d.max()
    assert np.abs(_unit(a) - _unit(b)).max() < 2e-6


def _lo_unit(hi):
    """2^-16 of the ulp of binary32 hi (exponent clamped below at -10): the
    step of the 16-bit low position parts."""
    h = np.abs(hi.astype(np.float64))
    e = np.where(h > 0, np.frexp(h)[1] - 1, -10)
    return np.ldexp(1.0, np.maximum(e, -10) - 39)


def test_compensated_positions_round_trip_resolution():
    """An FP32 state holds each position as binary32 hi plus a 16-bit low part
    in steps u(hi) = 2^-16 ulp(hi): positions read back (hi + lo) lie within
    u(hi) / 2 of the FP64 values imported, at binade edges, next to zero and
    next to the box length too, and a snapshot's compensation words resume
    to the same state bit for bit."""
    rng = np.random.default_rng(11)
    n = 4096
    parts = np.empty((n, 5))
    parts[:, :3] = rng.uniform(0.0, L, (n, 3))
    edge = np.array([0.0, 1e-30, 1e-9, 2.0 ** -10, 0.25, 0.5 - 1e-12, 0.5, 1.0, 2.0, 4.0, 8.0 - 1e-9, 8.0,
                     L - 1e-7, np.nextafter(L, 0.0)])
    parts[: edge.size, 0] = edge
    parts[: edge.size, 1] = edge[::-1]
    parts[:, 3] = rng.uniform(0.0, 2 * np.pi, n)
    parts[:, 4] = rng.uniform(0.0, np.pi, n)
    p = ss.make_params(box=(L, L, L), grid=(8, 8, 8, 4, 4), n=n, seed=5, precision=ss.FP32, binning=ss.BIN_SORTED)
    sim = ss.Simulation(p)
    sim.init(parts)
    got = sim.get_particles()
    # ids map back to the caller's order, so rows correspond
    hi = parts[:, :3].astype(np.float32)
    err = np.abs(got[:, :3] - parts[:, :3])
    err = np.minimum(err, L - err)
    assert np.all(err <= _lo_unit(hi) / 2 + 1e-300), err.max()
    # snapshot -> resume -> snapshot is exact; and one step from each agrees
    sim.step(3)
    snap = sim.get_snapshot()
    sim2 = ss.Simulation(p)
    sim2.resume(snap)
    snap2 = sim2.get_snapshot()
    np.testing.assert_array_equal(snap2.particles, snap.particles)
    np.testing.assert_array_equal(snap2.compensation, snap.compensation)
    sim.step(2)
    sim2.step(2)
    np.testing.assert_array_equal(sim.get_particles(), sim2.get_particles())

"""The magnetic suspension workload and the statistics its GPU test compares:
the polarisation time series <n_z>(t) and per-seed marginal histograms, each
against the seed-to-seed standard error of the difference of the two means.

Workload: external field of strength 2 along z acting as an orientational
torque (modifiers.rs:103-118 with the field direction given), mean-field
dipolar torque (mu_dd = 0.5, modifiers.rs:173-186) and dipolar drag force
(0.1, modifiers.rs:59-69), active stress -1, D_t = 0.1, D_r = 0.5, dt = 1e-3.
"""
import math

import numpy as np

import stochasticsampling_amd as ss

WORKLOAD = dict(box=(10.0, 10.0, 10.0), grid=(12, 12, 12, 8, 8), timestep=1e-3, magnetic_reorientation=2.0,
                field_direction=(0.0, 0.0, 1.0), magnetic_dipole_dipole=0.5, magnetic_drag=0.1,
                diffusion_rotational=0.5, diffusion_translational=0.1, stress_active=-1.0)
MARGINALS = [(4, -1.0, 1.0, np.cos), (3, 0.0, 2 * math.pi, None), (0, 0.0, 10.0, None)]  # cos theta, phi, x


def params(n, seed, **kw):
    return ss.make_params(n=n, seed=seed, **{**WORKLOAD, **kw})


def nz_mean(parts):
    return float(np.cos(parts[:, 4]).mean())


def histograms(parts, bins=10):
    """Per-marginal bin fractions of one system's particles."""
    out = []
    for col, lo, hi, f in MARGINALS:
        v = f(parts[:, col]) if f else parts[:, col]
        h, _ = np.histogram(v, bins=bins, range=(lo, hi))
        out.append(h / len(v))
    return np.array(out)


def oracle_run(oracle, n, seed, steps, every, **kw):
    """<n_z> every `every` steps (including t = 0) and the final histograms."""
    p = params(n, seed, **kw)
    sim = oracle.simulation(p)
    sim.init(ss.initial_condition(params(n, seed), amplitude=0.05, seed=seed))
    series = [nz_mean(sim.particles())]
    for _ in range(steps // every):
        sim.step(every)
        series.append(nz_mean(sim.particles()))
    return np.array(series), histograms(sim.particles())


def gap(a, b):
    """max over entries of |mean_a - mean_b| / SE, with SE from the seed-to-seed
    variances of both samples (axis 0 = seeds)."""
    a, b = np.asarray(a), np.asarray(b)
    se = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
    d = np.abs(a.mean(axis=0) - b.mean(axis=0))
    return float(np.max(d / np.maximum(se, 1e-9)))


# acceptance: the polarisation series within 4 SE at every time, each
# marginal bin within 4.5 SE (maxima over 11 times and 30 bins of t-like
# statistics from few seeds; a 10% torque error reaches 23 and 14 SE)
SERIES_SE = 4.0
MARGINAL_SE = 4.5

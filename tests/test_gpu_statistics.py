"""FP64 diagnostics and the magnetic suspension's statistics over 32 seeds
(GPU FP32 with in-kernel Philox noise against the FP64 CPU oracle)."""
import math

import numpy as np
import pytest

import stochasticsampling_amd as ss
from helpers import small_params

pytestmark = pytest.mark.gpu


def _orientation(parts):
    st, ct = np.sin(parts[:, 4]), np.cos(parts[:, 4])
    return np.stack([st * np.cos(parts[:, 3]), st * np.sin(parts[:, 3]), ct], axis=1)


@pytest.mark.parametrize("precision", [ss.FP64, ss.FP32])
def test_diagnostics_match_particle_averages(oracle, precision):
    p = small_params(n=5000, precision=precision, stress_magnetic=0.7, shape=0.2)
    parts = ss.initial_condition(p, amplitude=0.3, seed=3)
    sim = ss.Simulation(p)
    sim.init(parts)
    sim.step(3)
    d = sim.get_diagnostics()
    got = sim.get_particles()
    n = _orientation(got)
    assert np.max(np.abs(d["mean_orientation"] - n.mean(axis=0))) < (1e-10 if precision == ss.FP64 else 1e-7)
    # bin-centre orientation and stress kernel at each particle's bin
    K = oracle.stress_kernel(p)
    O = oracle.orientation_kernel(p)
    if precision == ss.FP64:
        bins = np.array([oracle.coord_to_grid(p, q) for q in got])
    else:
        bins = np.array([oracle.coord_to_grid_f32(p, q) for q in got.astype(np.float32)])
    kb = K[:, :, bins[:, 3], bins[:, 4]]
    ob = O[:, bins[:, 3], bins[:, 4]]
    assert np.max(np.abs(d["mean_polarisation"] - ob.mean(axis=1))) < 1e-10
    assert np.max(np.abs(d["mean_stress"] - kb.mean(axis=2))) < 1e-10


def test_diagnostics_independent_of_order_and_sort():
    kw = dict(grid=(16, 16, 16, 8, 8), n=1 << 14, precision=ss.FP32)
    parts = ss.initial_condition(small_params(**kw), amplitude=0.2, seed=4)
    out = []
    for tile, resort in [((4, 4, 4), 16), ((2, 2, 2), 3)]:
        sim = ss.Simulation(small_params(tile=tile, resort_interval=resort, **kw))
        sim.init(parts)
        sim.step(7)
        d = sim.get_diagnostics()
        out.append(np.concatenate([d["mean_orientation"], d["mean_polarisation"], d["mean_stress"].ravel()]))
    np.testing.assert_array_equal(out[0], out[1])


def test_magnetic_suspension_statistics_32_seeds(oracle):
    """External field of strength 2 along z (orientational torque), mean-field
    dipolar interaction (mu_dd = 0.5) and dipolar drag (0.1): the polarisation
    time series and the orientational and positional marginals of the GPU
    FP32 path agree with the FP64 oracle within the seed-to-seed confidence
    interval."""
    kw = dict(box=(10.0, 10.0, 10.0), grid=(12, 12, 12, 8, 8), n=2048, timestep=1e-3,
              magnetic_reorientation=2.0, field_direction=(0.0, 0.0, 1.0), magnetic_dipole_dipole=0.5,
              magnetic_drag=0.1, diffusion_rotational=0.5, diffusion_translational=0.1, stress_active=-1.0)
    steps, every, seeds = 60, 10, 32
    pz_gpu = np.zeros((seeds, steps // every + 1))
    pz_cpu = np.zeros_like(pz_gpu)
    fin_gpu, fin_cpu = [], []
    for k, seed in enumerate(range(100, 100 + seeds)):
        p32 = ss.make_params(seed=seed, precision=ss.FP32, **kw)
        p64 = ss.make_params(seed=seed, precision=ss.FP64, **kw)
        parts = ss.initial_condition(p64, amplitude=0.05, seed=seed)
        sim = ss.Simulation(p32)
        sim.init(parts)
        osim = oracle.simulation(p64)
        osim.init(parts)
        pz_gpu[k, 0] = sim.get_diagnostics()["mean_orientation"][2]
        pz_cpu[k, 0] = _orientation(osim.particles())[:, 2].mean()
        for j in range(1, steps // every + 1):
            sim.step(every)
            osim.step(every)
            pz_gpu[k, j] = sim.get_diagnostics()["mean_orientation"][2]
            pz_cpu[k, j] = _orientation(osim.particles())[:, 2].mean()
        fin_gpu.append(sim.get_particles())
        fin_cpu.append(osim.particles())
    # the field aligns the particles: polarisation along z grows
    assert pz_cpu[:, -1].mean() > pz_cpu[:, 0].mean() + 0.02
    mg, mc = pz_gpu.mean(axis=0), pz_cpu.mean(axis=0)
    se = np.sqrt(pz_gpu.var(axis=0, ddof=1) / seeds + pz_cpu.var(axis=0, ddof=1) / seeds)
    assert np.all(np.abs(mg - mc) <= 3 * se + 1e-6), (mg, mc, se)
    # marginals (pooled over seeds): cos theta, phi, x in 10 bins each
    G = np.concatenate(fin_gpu)
    Cc = np.concatenate(fin_cpu)
    for col, lo, hi, f in [(4, -1, 1, np.cos), (3, 0, 2 * math.pi, None), (0, 0, 10.0, None)]:
        a = f(G[:, col]) if f else G[:, col]
        b = f(Cc[:, col]) if f else Cc[:, col]
        ha, _ = np.histogram(a, bins=10, range=(lo, hi))
        hb, _ = np.histogram(b, bins=10, range=(lo, hi))
        n = len(a)
        pa, pb = ha / n, hb / n
        sd = np.sqrt((pa * (1 - pa) + pb * (1 - pb)) / n)
        assert np.all(np.abs(pa - pb) <= 4 * sd + 1e-12), (col, pa, pb)

"""Writes the committed golden vectors from the FP64 oracle (run after the
oracle matches the reference's fixtures, see tests/test_oracle.py):

  flow_single_particle.npz  the ff_test configuration (one particle at the
                            origin, phi = theta = pi/2, alpha = 1, 11^3 x 11 x 11,
                            L = 11, screening 0): u and grad u
  steps_small.npz           3 steps of an 8^3 x 4 x 4 system of 256 particles
                            with Philox noise: initial state, counts of the
                            last binning, final particles, u

Usage: python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle_lib import Oracle  # noqa: E402
from stochasticsampling_amd.params import make_params  # noqa: E402


def main():
    o = Oracle()
    p = make_params(box=(11, 11, 11), grid=(11, 11, 11, 11, 11), n=1, stress_active=1.0)
    part = o.particle_new([0, 0, 0, math.pi / 2, math.pi / 2], (11, 11, 11))
    dist = o.sample_from(p, o.histogram(p, part[None, :]), 1)
    u, g = o.mean_flow_field(p, 0.0, dist)
    np.savez_compressed(os.path.join(HERE, "flow_single_particle.npz"), u=u, grad_u=g)

    box, grid, n, seed, dt, steps = (8.0, 8.0, 8.0), (8, 8, 8, 4, 4), 256, 3, 1e-2, 3
    p = make_params(box=box, grid=grid, n=n, seed=seed, timestep=dt)
    rng = np.random.default_rng(12345)
    init = np.column_stack(
        [rng.random(n) * 8, rng.random(n) * 8, rng.random(n) * 8, rng.random(n) * 2 * math.pi,
         np.arccos(1 - 2 * rng.random(n))]
    )
    sim = o.simulation(p)
    sim.init(init)
    sim.step(steps)
    u, _ = sim.flow_field()
    np.savez_compressed(
        os.path.join(HERE, "steps_small.npz"), box=np.array(box), grid=np.array(grid), n=n, seed=seed,
        dt=dt, steps=steps, initial=init, counts=sim.counts(), particles=sim.particles(), u=u,
    )


if __name__ == "__main__":
    main()

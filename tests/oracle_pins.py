"""Checks that pin oracle functions the reference's fixtures leave open.

Each ``pin_*`` function takes an Oracle (the built one, or a copy compiled
with a planted mistake by tests/test_oracle_mutations.py) and raises
AssertionError when the oracle disagrees with a value fixed independently of
its code: a known answer of the reference's own tests, or a closed form
derived by hand from the mathematics of the cited reference function.
"""
import math

import numpy as np

from stochasticsampling_amd.params import make_params

PI = math.pi
SQ2 = math.sqrt(2.0)


def _fields(u=None, E=None, rho=0.0):
    f = np.zeros(34)
    if u is not None:
        f[0:3] = u
    if E is not None:
        f[12:21] = np.asarray(E, dtype=float).ravel()
    f[21] = rho
    return f


def pin_translational_diffusion_known_answer(o):
    """modifiers_test.rs:58-61: translational_diffusion with random vector
    [1, 1, 0] and standard deviation 1 moves the particle at the origin to
    (1, 1, 0).  In the full chain (mod.rs:346-347) the standard deviation is
    sqrt(2 dt (D_t + V rho)) = 1 for dt = 1, D_t = 0.5, V = 0; the
    orientation is (0, 0), n = (0, 0, 1), and the flow u = -n cancels the
    self-propulsion exactly."""
    bs = (10.0, 10.0, 10.0)
    p = make_params(box=bs, timestep=1.0, diffusion_translational=0.5)
    part = o.particle_new([0.0, 0.0, 0.0, 0.0, 0.0], bs)
    got = o.update_one(p, part, _fields(u=[0.0, 0.0, -1.0]), np.array([1.0, 1.0, 0.0, 0.0, 0.0]))
    assert np.array_equal(got, o.particle_new([1.0, 1.0, 0.0, 0.0, 0.0], bs)), got


def pin_translational_diffusion_amplitude(o):
    """The amplitude of the translational noise is sqrt(2 dt (D_t + V rho))
    (src/bin/simulation/mod.rs:346): with dt = 0.01, D_t = 0.1, V = 0.2 and
    rho = 3 a unit random vector component moves the particle by
    sqrt(0.014)."""
    bs = (10.0, 10.0, 10.0)
    dt, dtr, vex, rho = 0.01, 0.1, 0.2, 3.0
    p = make_params(box=bs, timestep=dt, diffusion_translational=dtr, volume_exclusion=vex)
    amp = math.sqrt(0.014)
    part = o.particle_new([5.0, 5.0, 5.0, 0.0, 0.0], bs)
    rv = np.array([1.0, -2.0, 0.5, 0.0, 0.0])
    got = o.update_one(p, part, _fields(u=[0.0, 0.0, -1.0], rho=rho), rv)
    expect = np.array([5.0 + amp, 5.0 - 2.0 * amp, 5.0 + 0.5 * amp])
    assert np.max(np.abs(got[:3] - expect)) < 1e-15, (got, expect)
    assert got[3] == 0.0 and got[4] == 0.0


def pin_jeffrey_strain_closed_forms(o):
    """jeffrey_strain (modifiers.rs:147-169) adds dt * shape * (E n - n (n.E n))
    to the orientation vector.  Two cases with a closed-form answer:

    * n at phi = pi/4 in the xy-plane, E = a e_x e_x: the increment is
      h (1, -1, 0)/sqrt(2) with h = dt shape a / 2, so phi' = pi/4 - atan(h);
    * n at theta0 in the xz-plane, E = a (e_x e_z + e_z e_x): the increment is
      eps e_theta with eps = dt shape a cos(2 theta0), so
      theta' = theta0 + atan(eps).
    """
    bs = (10.0, 10.0, 10.0)
    dt, shape = 0.01, 0.7
    p = make_params(box=bs, timestep=dt, diffusion_translational=0.0, shape=shape)
    a = 0.3
    E = np.zeros((3, 3))
    E[0, 0] = a
    part = o.particle_new([1.0, 1.0, 1.0, PI / 4, PI / 2], bs)
    got = o.update_one(p, part, _fields(E=E), np.zeros(5))
    h = dt * shape * a / 2.0
    assert abs(got[3] - (PI / 4 - math.atan(h))) < 1e-14, (got[3], PI / 4 - math.atan(h))
    assert abs(got[4] - PI / 2) < 1e-14

    dt, shape, a, th0 = 0.02, 0.8, 0.5, 0.6
    p = make_params(box=bs, timestep=dt, diffusion_translational=0.0, shape=shape)
    E = np.zeros((3, 3))
    E[0, 2] = E[2, 0] = a
    part = o.particle_new([1.0, 1.0, 1.0, 0.0, th0], bs)
    got = o.update_one(p, part, _fields(E=E), np.zeros(5))
    eps = dt * shape * a * math.cos(2 * th0)
    assert abs(got[4] - (th0 + math.atan(eps))) < 1e-14, (got[4], th0 + math.atan(eps))
    assert abs(got[3]) < 1e-14


def pin_rods_stress_special_angles(o):
    """stress_magnetic_rods (stress.rs:144-159) through stress_kernel with
    shape 1 and the other coefficients 0, at bin centres where the entries
    are known in closed form:
      (phi, theta) = (pi/8, pi/4): [[(sqrt2+1)/8, 1/8, 0], [1/8, (sqrt2-1)/8, 0], [0, 0, -sqrt2/4]]
      (phi, theta) = (3pi/8, 3pi/4): [[-(sqrt2-1)/8, 1/8, 0], [1/8, -(sqrt2+1)/8, 0], [0, 0, sqrt2/4]]
      (phi, theta) = (pi/8, pi/2): off-diagonal xz = -cos(pi/8)/2, yz = -sin(pi/8)/2, rest 0."""
    p = make_params(box=(1, 1, 1), grid=(1, 1, 1, 8, 2), stress_active=0.0, stress_magnetic=0.0, shape=1.0)
    k = o.stress_kernel(p)  # [3][3][phi][theta]; phi centres (2i+1) pi/8, theta centres pi/4, 3pi/4
    exp = np.array([[(SQ2 + 1) / 8, 1 / 8, 0.0], [1 / 8, (SQ2 - 1) / 8, 0.0], [0.0, 0.0, -SQ2 / 4]])
    assert np.max(np.abs(k[:, :, 0, 0] - exp)) < 1e-15, k[:, :, 0, 0]
    exp = np.array([[-(SQ2 - 1) / 8, 1 / 8, 0.0], [1 / 8, -(SQ2 + 1) / 8, 0.0], [0.0, 0.0, SQ2 / 4]])
    assert np.max(np.abs(k[:, :, 1, 1] - exp)) < 1e-15, k[:, :, 1, 1]
    p = make_params(box=(1, 1, 1), grid=(1, 1, 1, 8, 1), stress_active=0.0, stress_magnetic=0.0, shape=1.0)
    k = o.stress_kernel(p)
    c, s = math.cos(PI / 8) / 2, math.sin(PI / 8) / 2
    exp = np.array([[0.0, 0.0, -c], [0.0, 0.0, -s], [-c, -s, 0.0]])
    assert np.max(np.abs(k[:, :, 0, 0] - exp)) < 1e-15, k[:, :, 0, 0]


def pin_compare_implementations(o):
    """spectral_solver_test.rs:151-190 (test_compare_implementations): the
    deprecated solve_flow_field and mean_flow_field(0) agree after rounding
    both to multiples of 2^-51 (equal_floats, test_helper.rs:13-40), for
    100000 particles of create_isotropic(seed 1) on 20^3 x 15 x 15, box 10,
    active stress 1.  The distribution is sample_from without the box-volume
    factor, as in the reference test."""
    p = make_params(box=(10, 10, 10), grid=(20, 20, 20, 15, 15), n=100000, stress_active=1.0)
    parts = o.create_isotropic(100000, (10.0, 10.0, 10.0), 1)
    dist = o.sample_from(p, o.histogram(p, parts), 100000) / 1000.0
    old = o.solve_flow_field(p, dist)
    new, _ = o.mean_flow_field(p, 0.0, dist)
    f = 2.0**51
    va, vb = np.round(old * f) / f, np.round(new * f) / f
    eps = np.finfo(np.float64).eps
    d = np.abs(va - vb)
    s = np.abs(va) + np.abs(vb)
    ok = ((va == 0) & (vb == 0)) | (((va == 0) ^ (vb == 0)) & (d < eps)) | (
        (va != 0) & (vb != 0) & (d / np.where(s > 0, s, 1.0) < eps))
    assert ok.all(), f"{int((~ok).sum())} entries differ"


PINS = [
    pin_translational_diffusion_known_answer,
    pin_translational_diffusion_amplitude,
    pin_jeffrey_strain_closed_forms,
    pin_rods_stress_special_angles,
    pin_compare_implementations,
]

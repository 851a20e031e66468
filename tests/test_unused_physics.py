"""DensityGradient (src/distribution/density_gradient.rs) and the trilinear
interpolation of src/mesh/interpolate.rs, which the reference keeps beside
the step without calling them.

Oracle pins: the reference's own tables (interpolate.rs:129-222:
trilinear_test, get_neighbouring_cell_values_test,
interpolate_vector_field_test, compared like the reference with
equal_floats or exactly), a linear field reproduced exactly by the
interpolation, and the closed-form gradient of a single cosine mode of the
density.  The GPU entry points against the oracle: interpolation bit-exact
(FP64, same operation order), the gradient within 1e-12 of the largest value
in FP64 and 1e-5 in FP32.
"""
import math

import numpy as np
import pytest

import stochasticsampling_amd as ss
from stochasticsampling_amd.params import FP32, FP64, make_params

EPS = np.finfo(np.float64).eps


def equal_floats(a, b, eps=EPS):
    """test_helper.rs:13-40."""
    if a == 0.0 and b == 0.0:
        return True
    d = abs(a - b)
    if a == 0.0 or b == 0.0:
        return d < eps
    return d / min(abs(a) + abs(b), np.finfo(np.float64).max) < eps


TRILINEAR = [((0.1, 0.4, 0.3), 0.2), ((0.4, 0.1, 0.9), -0.4), ((0.2, 0.1, 0.7), -0.4), ((0.5, 0.3, 0.3), 0.5),
             ((0.9, 0.9, 0.9), 0.9), ((0.0, 0.0, 0.0), 0.0), ((0.0, 1.0, 0.0), 1.0), ((0.0, 0.0, 1.0), -1.0)]


@pytest.mark.parametrize("rel,want", TRILINEAR)
def test_trilinear_reference_table(oracle, rel, want):
    """interpolate.rs:129-149."""
    a = np.array([[[0.0, -1.0], [1.0, 0.0]], [[1.0, 0.0], [2.0, 1.0]]])
    assert equal_floats(oracle.trilinear(rel, a), want)


NEIGHBOURS = [
    ((0.1, 0.2, 0.3), [[[999, 990], [909, 900]], [[99, 90], [9, 0]]]),
    ((19.9, 19.9, 19.9), [[[999, 990], [909, 900]], [[99, 90], [9, 0]]]),
    ((1.1, 1.1, 1.1), [[[0, 1], [10, 11]], [[100, 101], [110, 111]]]),
    ((0.1, 0.1, 1.1), [[[990, 991], [900, 901]], [[90, 91], [0, 1]]]),
    ((1.0, 1.0, 1.0), [[[0, 1], [10, 11]], [[100, 101], [110, 111]]]),
]


@pytest.mark.parametrize("pos,want", NEIGHBOURS)
def test_neighbouring_cells_reference_table(oracle, pos, want):
    """interpolate.rs:151-188 (grid width 2, values 0..999 on 10^3 cells)."""
    a = np.linspace(0.0, 999.0, 1000).reshape(10, 10, 10)
    np.testing.assert_array_equal(oracle.neighbour_values(a, pos, (2.0, 2.0, 2.0)), np.array(want, dtype=float))


@pytest.mark.parametrize("pos,want", [((1.0, 1.0, 1.0), (0.0, 1000.0, 2000.0)),
                                      ((1.0, 1.0, 3.0), (1.0, 1001.0, 2001.0)),
                                      ((1.5, 1.5, 1.5), (27.75, 1027.75, 2027.75))])
def test_interpolate_vector_field_reference_table(oracle, pos, want):
    """interpolate.rs:190-222."""
    a = np.linspace(0.0, 2999.0, 3000).reshape(3, 10, 10, 10)
    got = oracle.interpolate_vector_field(a, pos, (2.0, 2.0, 2.0))
    assert all(equal_floats(g, w) for g, w in zip(got, want))


def test_interpolation_reproduces_linear_fields_and_uses_z(oracle):
    """Inside the grid (no wrap), a field linear in x, y and z is reproduced
    exactly up to rounding -- which needs the z offset from position.z (the
    reference's position.y would fail the z component)."""
    n, g = 12, 0.5
    i = np.arange(n) + 0.5
    X, Y, Z = np.meshgrid(i * g, i * g, i * g, indexing="ij")
    f = np.stack([2 * X - Y + 0.5 * Z, X + 3 * Y, 7 * Z])
    rng = np.random.default_rng(3)
    for _ in range(50):
        q = rng.uniform(g, (n - 1) * g, size=3)
        got = oracle.interpolate_vector_field(f, q, (g, g, g))
        want = [2 * q[0] - q[1] + 0.5 * q[2], q[0] + 3 * q[1], 7 * q[2]]
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def _mode_dist(p, amp, m):
    """A distribution whose density is 1 + amp cos(2 pi m x / Lx) at the cell
    centres, spread evenly over the angles."""
    g = p.grid_size
    gx = p.box_size.x / g.x
    dthph = (2 * math.pi / g.phi) * (math.pi / g.theta)
    x = (np.arange(g.x) + 0.5) * gx
    rho = 1.0 + amp * np.cos(2 * math.pi * m * x / p.box_size.x)
    d = np.broadcast_to(rho[:, None, None, None, None] / (g.phi * g.theta * dthph),
                        (g.x, g.y, g.z, g.phi, g.theta)).copy()
    return d, x, rho


def test_density_gradient_single_mode(oracle):
    """Spectral derivative of a resolved cosine: -amp k sin(k x) along x, zero
    along y and z."""
    p = make_params(box=(8.0, 6.0, 5.0), grid=(16, 12, 10, 4, 6))
    d, x, _ = _mode_dist(p, 0.3, 2)
    g = oracle.density_gradient(p, d)
    k = 2 * math.pi * 2 / 8.0
    want = -0.3 * k * np.sin(k * x)
    np.testing.assert_allclose(g[0], np.broadcast_to(want[:, None, None], g[0].shape), atol=1e-12)
    assert np.max(np.abs(g[1])) < 1e-12 and np.max(np.abs(g[2])) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_gpu_density_gradient_matches_oracle(oracle, precision, tol):
    p = make_params(box=(8.0, 6.0, 5.0), grid=(16, 12, 10, 4, 6), precision=precision)
    rng = np.random.default_rng(5)
    d = rng.random((16, 12, 10, 4, 6))
    want = oracle.density_gradient(p, d)
    got = ss.density_gradient(p, d)
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))


@pytest.mark.gpu
def test_gpu_interpolation_matches_oracle(oracle):
    p = make_params(box=(8.0, 6.0, 5.0), grid=(16, 12, 10, 4, 4))
    rng = np.random.default_rng(6)
    f = rng.standard_normal((3, 16, 12, 10))
    q = rng.random((2000, 3)) * np.array([8.0, 6.0, 5.0])
    q[:3] = [[0.0, 0.0, 0.0], [8.0 - 1e-12, 6.0 - 1e-12, 5.0 - 1e-12], [0.25, 0.25, 0.25]]
    got = ss.interpolate_vector_field(p, f, q)
    gw = (0.5, 0.5, 0.5)
    want = np.array([oracle.interpolate_vector_field(f, qq, gw) for qq in q])
    np.testing.assert_array_equal(got, want)

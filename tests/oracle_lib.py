"""ctypes bindings of the FP64 CPU oracle (oracle/ss_oracle.c).

Test infrastructure: only the tests, the smoke check and the CPU-baseline leg
of the benchmark load this library.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from stochasticsampling_amd.params import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "libss_oracle.so")

_P = C.POINTER
_d = _P(C.c_double)


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def _dp(a):
    return a.ctypes.data_as(_d)


class OPcg32(C.Structure):
    _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]


class OMcg128(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


class Oracle:
    def __init__(self, so_path=None):
        if so_path is None:
            so_path = ORACLE_SO
            if not os.path.exists(so_path):
                build_oracle()
        L = C.CDLL(so_path)
        self.L = L
        pp = _P(Params)
        L.osim_create.argtypes = [pp]
        L.osim_create.restype = C.c_void_p
        L.osim_destroy.argtypes = [C.c_void_p]
        L.osim_init.argtypes = [C.c_void_p, _d, C.c_uint64]
        L.osim_set_noise.argtypes = [C.c_void_p, _d, C.c_uint64]
        L.osim_step.argtypes = [C.c_void_p, C.c_uint64]
        L.osim_get_timestep.argtypes = [C.c_void_p]
        L.osim_get_timestep.restype = C.c_uint64
        L.osim_get_particles.argtypes = [C.c_void_p, _d, C.c_uint64]
        L.osim_get_counts.argtypes = [C.c_void_p, _P(C.c_uint32)]
        L.osim_get_distribution.argtypes = [C.c_void_p, _d]
        L.osim_get_flow_field.argtypes = [C.c_void_p, _d, _d]
        L.osim_get_magnetic_field.argtypes = [C.c_void_p, _d, _d]
        L.osim_get_noise.argtypes = [C.c_void_p, _d]
        L.osim_set_threads.argtypes = [C.c_int]
        L.o_modulo.argtypes = [C.c_double, C.c_double]
        L.o_modulo.restype = C.c_double
        L.o_ang_pbc.argtypes = [C.c_double, C.c_double, _d]
        L.o_particle_new.argtypes = [_d, C.c_void_p, _d]
        L.o_orientation_from_vector.argtypes = [_d, _d]
        L.o_coord_to_grid.argtypes = [pp, _d, _P(C.c_int32)]
        L.o_coord_to_grid_f32.argtypes = [pp, _P(C.c_float), _P(C.c_int32)]
        L.o_k_sampling.argtypes = [C.c_int, C.c_double, _d]
        L.o_inverse_norm_squared.argtypes = [pp, _d]
        L.o_stress_kernel.argtypes = [pp, _d]
        L.o_orientation_kernel.argtypes = [pp, _d]
        L.o_linspace.argtypes = [C.c_double, C.c_double, C.c_int, _d]
        L.o_fft1d.argtypes = [C.c_int, _d, C.c_int]
        L.o_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _d, C.c_int]
        L.o_histogram.argtypes = [pp, _d, C.c_uint64, _P(C.c_uint32)]
        L.o_sample_from.argtypes = [pp, _P(C.c_uint32), C.c_uint64, _d]
        L.o_mean_flow_field.argtypes = [pp, C.c_double, _d, _d, _d]
        L.o_mean_magnetic_field.argtypes = [pp, _d, _d, _d]
        L.o_solve_flow_field.argtypes = [pp, _d, _d]
        L.o_update_one.argtypes = [pp, _d, _d, _d, _d]
        L.o_philox4x32_10.argtypes = [_P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint32)]
        L.o_noise.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, _d]
        u64p = _P(C.c_uint64)
        L.o_pcg32_new.argtypes = [C.c_uint64, C.c_uint64, _P(OPcg32)]
        L.o_pcg32_from_seed.argtypes = [_P(C.c_uint8), _P(OPcg32)]
        L.o_pcg32_seed_from_u64.argtypes = [C.c_uint64, _P(OPcg32)]
        L.o_pcg32_next_u32.argtypes = [_P(OPcg32)]
        L.o_pcg32_next_u32.restype = C.c_uint32
        L.o_pcg32_next_u64.argtypes = [_P(OPcg32)]
        L.o_pcg32_next_u64.restype = C.c_uint64
        L.o_mcg128_new.argtypes = [C.c_uint64, C.c_uint64, _P(OMcg128)]
        L.o_mcg128_seed_from_u64.argtypes = [C.c_uint64, _P(OMcg128)]
        L.o_mcg128_next_u64.argtypes = [_P(OMcg128)]
        L.o_mcg128_next_u64.restype = C.c_uint64
        for f in ("o_uniform01", "o_gen_f64", "o_open01"):
            getattr(L, f).argtypes = [C.c_uint64]
            getattr(L, f).restype = C.c_double
        L.o_zig_tables.argtypes = [_d, _d]
        L.o_standard_normal.argtypes = [_P(OPcg32)]
        L.o_standard_normal.restype = C.c_double
        L.o_ref_noise.argtypes = [_P(OPcg32), C.c_int, C.c_uint64, C.c_double, _d]
        L.o_create_isotropic.argtypes = [C.c_uint64, C.c_void_p, C.c_uint64, _d]
        L.o_create_homogeneous.argtypes = [C.c_uint64, C.c_double, C.c_void_p, C.c_uint64, _d]
        L.osim_get_rng_state.argtypes = [C.c_void_p, u64p, C.c_int]
        L.osim_set_rng_state.argtypes = [C.c_void_p, u64p, C.c_int]
        L.osim_set_timestep.argtypes = [C.c_void_p, C.c_uint64]
        L.o_density_gradient.argtypes = [pp, _d, _d]
        L.o_trilinear.argtypes = [_d, _d]
        L.o_trilinear.restype = C.c_double
        L.o_neighbour_values.argtypes = [_d, C.c_int, C.c_int, C.c_int, _d, _d, _d]
        L.o_interpolate_vector_field.argtypes = [_d, C.c_int, C.c_int, C.c_int, _d, _d, _d]

    # pure functions -------------------------------------------------------
    def modulo(self, f, m):
        return self.L.o_modulo(f, m)

    def ang_pbc(self, phi, theta):
        out = np.zeros(2)
        self.L.o_ang_pbc(phi, theta, _dp(out))
        return out

    def particle_new(self, p5, box):
        from stochasticsampling_amd.params import Box

        inp = np.ascontiguousarray(p5, dtype=np.float64)
        out = np.zeros(5)
        b = Box(*box)
        self.L.o_particle_new(_dp(inp), C.byref(b), _dp(out))
        return out

    def orientation_from_vector(self, v):
        inp = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(2)
        self.L.o_orientation_from_vector(_dp(inp), _dp(out))
        return out

    def coord_to_grid(self, params, p5):
        inp = np.ascontiguousarray(p5, dtype=np.float64)
        out = np.zeros(5, dtype=np.int32)
        self.L.o_coord_to_grid(C.byref(params), _dp(inp), out.ctypes.data_as(_P(C.c_int32)))
        return out

    def coord_to_grid_f32(self, params, p5):
        inp = np.ascontiguousarray(p5, dtype=np.float32)
        out = np.zeros(5, dtype=np.int32)
        self.L.o_coord_to_grid_f32(
            C.byref(params), inp.ctypes.data_as(_P(C.c_float)), out.ctypes.data_as(_P(C.c_int32))
        )
        return out

    def k_sampling(self, n, length):
        out = np.zeros(n)
        self.L.o_k_sampling(n, length, _dp(out))
        return out

    def inverse_norm_squared(self, params):
        out = np.zeros(params.ns)
        self.L.o_inverse_norm_squared(C.byref(params), _dp(out))
        return out.reshape(params.grid_shape)

    def stress_kernel(self, params):
        out = np.zeros(9 * params.na)
        self.L.o_stress_kernel(C.byref(params), _dp(out))
        g = params.grid_size
        return out.reshape(3, 3, g.phi, g.theta)

    def orientation_kernel(self, params):
        out = np.zeros(3 * params.na)
        self.L.o_orientation_kernel(C.byref(params), _dp(out))
        g = params.grid_size
        return out.reshape(3, g.phi, g.theta)

    def linspace(self, a, b, n):
        out = np.zeros(n)
        self.L.o_linspace(a, b, n, _dp(out))
        return out

    def fft1d(self, x, sign):
        a = np.array(x, dtype=np.complex128, copy=True)
        self.L.o_fft1d(len(a), a.ctypes.data_as(_d), sign)
        return a

    def fft3d(self, x, sign):
        a = np.array(x, dtype=np.complex128, copy=True, order="C")
        self.L.o_fft3d(a.shape[0], a.shape[1], a.shape[2], a.ctypes.data_as(_d), sign)
        return a

    def histogram(self, params, particles):
        parts = np.ascontiguousarray(particles, dtype=np.float64)
        counts = np.zeros(params.ns * params.na, dtype=np.uint32)
        self.L.o_histogram(
            C.byref(params), _dp(parts), len(parts), counts.ctypes.data_as(_P(C.c_uint32))
        )
        return counts.reshape(params.dist_shape)

    def sample_from(self, params, counts, n):
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        out = np.zeros(params.ns * params.na)
        self.L.o_sample_from(C.byref(params), c.ctypes.data_as(_P(C.c_uint32)), n, _dp(out))
        return out.reshape(params.dist_shape)

    def mean_flow_field(self, params, screening, dist):
        d = np.ascontiguousarray(dist, dtype=np.float64)
        u = np.zeros(3 * params.ns)
        g = np.zeros(9 * params.ns)
        rc = self.L.o_mean_flow_field(C.byref(params), screening, _dp(d), _dp(u), _dp(g))
        assert rc == 0
        return u.reshape((3,) + params.grid_shape), g.reshape((3, 3) + params.grid_shape)

    def solve_flow_field(self, params, dist):
        """The deprecated formulation (spectral_solver.rs:85-131): real u."""
        d = np.ascontiguousarray(dist, dtype=np.float64)
        u = np.zeros(3 * params.ns)
        assert self.L.o_solve_flow_field(C.byref(params), _dp(d), _dp(u)) == 0
        return u.reshape((3,) + params.grid_shape)

    def mean_magnetic_field(self, params, dist):
        d = np.ascontiguousarray(dist, dtype=np.float64)
        b = np.zeros(3 * params.ns)
        g = np.zeros(9 * params.ns)
        rc = self.L.o_mean_magnetic_field(C.byref(params), _dp(d), _dp(b), _dp(g))
        assert rc == 0
        return b.reshape((3,) + params.grid_shape), g.reshape((3, 3) + params.grid_shape)

    def update_one(self, params, p5, fields, rv):
        p = np.ascontiguousarray(p5, dtype=np.float64)
        f = np.ascontiguousarray(fields, dtype=np.float64)
        r = np.ascontiguousarray(rv, dtype=np.float64)
        out = np.zeros(5)
        self.L.o_update_one(C.byref(params), _dp(p), _dp(f), _dp(r), _dp(out))
        return out

    def philox(self, ctr, key):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        self.L.o_philox4x32_10(c, k, o)
        return list(o)

    def noise(self, seed, step, pid, dr):
        out = np.zeros(5)
        self.L.o_noise(seed, step, pid, dr, _dp(out))
        return out

    def set_threads(self, n):
        self.L.osim_set_threads(n)

    # physics the live step leaves unused -----------------------------------
    def density_gradient(self, params, dist):
        d = np.ascontiguousarray(dist, dtype=np.float64)
        out = np.zeros(3 * params.ns)
        self.L.o_density_gradient(C.byref(params), _dp(d), _dp(out))
        return out.reshape((3,) + params.grid_shape)

    def trilinear(self, rel, v):
        r = np.ascontiguousarray(rel, dtype=np.float64)
        a = np.ascontiguousarray(v, dtype=np.float64).ravel()
        return self.L.o_trilinear(_dp(r), _dp(a))

    def neighbour_values(self, f, pos, gw):
        f = np.ascontiguousarray(f, dtype=np.float64)
        p = np.ascontiguousarray(pos, dtype=np.float64)
        g = np.ascontiguousarray(gw, dtype=np.float64)
        out = np.zeros(8)
        self.L.o_neighbour_values(_dp(f), f.shape[0], f.shape[1], f.shape[2], _dp(p), _dp(g), _dp(out))
        return out.reshape(2, 2, 2)

    def interpolate_vector_field(self, field, pos, gw):
        f = np.ascontiguousarray(field, dtype=np.float64)
        p = np.ascontiguousarray(pos, dtype=np.float64)
        g = np.ascontiguousarray(gw, dtype=np.float64)
        out = np.zeros(3)
        self.L.o_interpolate_vector_field(_dp(f), f.shape[1], f.shape[2], f.shape[3], _dp(p), _dp(g), _dp(out))
        return out

    # the reference's random streams ----------------------------------------
    def pcg32_new(self, state, stream):
        r = OPcg32()
        self.L.o_pcg32_new(state, stream, C.byref(r))
        return r

    def pcg32_from_seed(self, seed16):
        r = OPcg32()
        b = (C.c_uint8 * 16)(*seed16)
        self.L.o_pcg32_from_seed(b, C.byref(r))
        return r

    def pcg32_seed_from_u64(self, s):
        r = OPcg32()
        self.L.o_pcg32_seed_from_u64(s, C.byref(r))
        return r

    def pcg32_u32(self, r, k=1):
        return [self.L.o_pcg32_next_u32(C.byref(r)) for _ in range(k)]

    def pcg32_u64(self, r, k=1):
        return [self.L.o_pcg32_next_u64(C.byref(r)) for _ in range(k)]

    def mcg128_new(self, state128):
        r = OMcg128()
        self.L.o_mcg128_new(state128 & (2**64 - 1), state128 >> 64, C.byref(r))
        return r

    def mcg128_seed_from_u64(self, s):
        r = OMcg128()
        self.L.o_mcg128_seed_from_u64(s, C.byref(r))
        return r

    def mcg128_u64(self, r, k=1):
        return [self.L.o_mcg128_next_u64(C.byref(r)) for _ in range(k)]

    def zig_tables(self):
        x = np.zeros(257)
        f = np.zeros(257)
        self.L.o_zig_tables(_dp(x), _dp(f))
        return x, f

    def normals(self, r, k):
        return np.array([self.L.o_standard_normal(C.byref(r)) for _ in range(k)])

    def ref_noise(self, states, n, dr):
        """One step of the reference noise from generators ``states`` (list
        of OPcg32, advanced in place): N x 5."""
        arr = (OPcg32 * len(states))(*states)
        out = np.zeros((n, 5))
        self.L.o_ref_noise(arr, len(states), n, dr, _dp(out))
        for k in range(len(states)):
            states[k].state, states[k].inc = arr[k].state, arr[k].inc
        return out

    def create_isotropic(self, n, box, seed):
        from stochasticsampling_amd.params import Box

        out = np.zeros((n, 5))
        b = Box(*box)
        self.L.o_create_isotropic(n, C.byref(b), seed, _dp(out))
        return out

    def create_homogeneous(self, n, kappa, box, seed):
        from stochasticsampling_amd.params import Box

        out = np.zeros((n, 5))
        b = Box(*box)
        assert self.L.o_create_homogeneous(n, kappa, C.byref(b), seed, _dp(out)) == 0
        return out

    def simulation(self, params):
        return OracleSim(self, params)


class OracleSim:
    """Simulation (src/bin/simulation/mod.rs:60-393) restated on the CPU."""

    def __init__(self, oracle, params):
        self.o = oracle
        self.L = oracle.L
        self.params = params
        self.h = self.L.osim_create(C.byref(params))
        if not self.h:
            raise MemoryError("osim_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.osim_destroy(self.h)
            self.h = None

    def init(self, particles):
        p = np.ascontiguousarray(particles, dtype=np.float64)
        assert self.L.osim_init(self.h, _dp(p), len(p)) == 0

    def set_noise(self, rv):
        r = np.ascontiguousarray(rv, dtype=np.float64)
        assert self.L.osim_set_noise(self.h, _dp(r), len(r)) == 0

    def step(self, n=1):
        assert self.L.osim_step(self.h, n) == 0
        return self.timestep

    @property
    def timestep(self):
        return self.L.osim_get_timestep(self.h)

    def particles(self):
        n = self.params.number_of_particles
        out = np.zeros((n, 5))
        assert self.L.osim_get_particles(self.h, _dp(out), n) == 0
        return out

    def counts(self):
        out = np.zeros(self.params.ns * self.params.na, dtype=np.uint32)
        self.L.osim_get_counts(self.h, out.ctypes.data_as(_P(C.c_uint32)))
        return out.reshape(self.params.dist_shape)

    def distribution(self):
        out = np.zeros(self.params.ns * self.params.na)
        self.L.osim_get_distribution(self.h, _dp(out))
        return out.reshape(self.params.dist_shape)

    def flow_field(self):
        u = np.zeros(3 * self.params.ns)
        g = np.zeros(9 * self.params.ns)
        self.L.osim_get_flow_field(self.h, _dp(u), _dp(g))
        return u.reshape((3,) + self.params.grid_shape), g.reshape((3, 3) + self.params.grid_shape)

    def magnetic_field(self):
        b = np.zeros(3 * self.params.ns)
        g = np.zeros(9 * self.params.ns)
        self.L.osim_get_magnetic_field(self.h, _dp(b), _dp(g))
        return b.reshape((3,) + self.params.grid_shape), g.reshape((3, 3) + self.params.grid_shape)

    def noise(self):
        out = np.zeros((self.params.number_of_particles, 5))
        self.L.osim_get_noise(self.h, _dp(out))
        return out

    def rng_state(self):
        t = max(1, int(self.params.noise_streams))
        buf = np.zeros((t, 2), dtype=np.uint64)
        k = self.L.osim_get_rng_state(self.h, buf.ctypes.data_as(_P(C.c_uint64)), t)
        assert k >= 0
        return buf[:k].copy()

    def set_rng_state(self, st):
        a = np.ascontiguousarray(st, dtype=np.uint64).reshape(-1, 2)
        assert self.L.osim_set_rng_state(self.h, a.ctypes.data_as(_P(C.c_uint64)), a.shape[0]) == 0

    def set_timestep(self, t):
        self.L.osim_set_timestep(self.h, t)

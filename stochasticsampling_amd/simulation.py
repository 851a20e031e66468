"""Python mirror of the reference's ``Simulation`` API
(src/bin/simulation/mod.rs:60-393), ``SpectralSolver`` / ``MagneticSolver``
solves (spectral_solver.rs:233-258, magnetic_solver.rs:139-169) and ``FFTPlan``
(crates/fftw3/src/fft.rs:33-244), over the C ABI of libss.

Method names, argument meaning and error behaviour follow the reference:
``Simulation(params)`` is ``Simulation::new``; ``init`` rejects a particle
count that differs from the parameters (the reference asserts); getters copy
out like the reference's clones.
"""
import ctypes as C

import numpy as np

from . import lib as _l
from .params import Box, Info, Params

_P = C.POINTER


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


class Snapshot:
    """Snapshot (src/bin/simulation/mod.rs:80-86): particles, the low-order
    position parts of the FP32 path, the timestep (with the seed, the whole
    state of the counter-based Philox noise) and, for the reference's own
    noise streams, ``rng_state``: (state, increment) per Pcg32 generator,
    shape (T, 2) uint64 (None or empty otherwise)."""

    def __init__(self, particles, compensation, timestep, rng_state=None):
        self.particles = particles
        self.compensation = compensation
        self.timestep = timestep
        self.rng_state = rng_state


class Simulation:
    def __init__(self, params: Params):
        self.L = _l.load()
        self.params = params
        h = C.c_void_p()
        _l.check(self.L.ss_create(C.byref(params), C.byref(h)))
        self._h = h
        self.n = params.number_of_particles

    # lifecycle -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self.L.ss_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # Simulation::init / resume / get_snapshot --------------------------------
    def init(self, particles):
        p = _f64(particles)
        if p.ndim != 2 or p.shape[1] != 5:
            raise ValueError("particles must be N x 5 [x, y, z, phi, theta]")
        _l.check(self.L.ss_init(self._h, _l.dptr(p), p.shape[0]))

    def resume(self, snapshot: Snapshot):
        p = _f64(snapshot.particles)
        comp = None if snapshot.compensation is None else _f64(snapshot.compensation)
        _l.check(
            self.L.ss_resume(
                self._h,
                _l.dptr(p),
                None if comp is None else _l.dptr(comp),
                p.shape[0],
                snapshot.timestep,
            )
        )
        rs = getattr(snapshot, "rng_state", None)
        if rs is not None and len(rs):
            self.set_rng_state(rs)

    def get_snapshot(self):
        p = np.zeros((self.n, 5))
        comp = np.zeros((self.n, 3))
        t = C.c_uint64()
        _l.check(self.L.ss_get_snapshot(self._h, _l.dptr(p), _l.dptr(comp), C.byref(t)))
        return Snapshot(p, comp, t.value, self.get_rng_state())

    def get_rng_state(self):
        """(T, 2) uint64 (state, increment) of the reference-noise generators;
        an empty array in the other noise modes."""
        t = max(0, int(self.params.noise_streams))
        buf = np.zeros((max(t, 1), 2), dtype=np.uint64)
        rc = self.L.ss_get_rng_state(self._h, buf.ctypes.data_as(_P(C.c_uint64)), t)
        _l.check(rc if rc < 0 else 0)
        return buf[:rc].copy()

    def set_rng_state(self, st):
        """Replace the first min(len(st), T) generators (Simulation::resume's
        zip, mod.rs:199-207)."""
        a = np.ascontiguousarray(st, dtype=np.uint64).reshape(-1, 2)
        _l.check(self.L.ss_set_rng_state(self._h, a.ctypes.data_as(_P(C.c_uint64)), a.shape[0]))

    # noise -------------------------------------------------------------------
    def set_noise(self, rv):
        r = _f64(rv)
        _l.check(self.L.ss_set_noise(self._h, _l.dptr(r), r.shape[0]))

    def get_noise(self):
        """RandomVectors of the last step (N x 5: x, y, z, axis_angle,
        rotate_angle), external or reference noise."""
        out = np.zeros((self.n, 5))
        _l.check(self.L.ss_get_noise(self._h, _l.dptr(out), self.n))
        return out

    # Simulation::do_timestep / Iterator ---------------------------------------
    def do_timestep(self):
        t = C.c_uint64()
        _l.check(self.L.ss_step(self._h, 1, C.byref(t)))
        return t.value

    def step(self, nsteps=1):
        t = C.c_uint64()
        _l.check(self.L.ss_step(self._h, nsteps, C.byref(t)))
        return t.value

    def __iter__(self):
        return self

    def __next__(self):
        return self.do_timestep()

    def synchronize(self):
        _l.check(self.L.ss_synchronize(self._h))

    def keep_distribution(self, on=True):
        """Keep (or stop keeping) the count grid of every step on the tiled
        path (ss_keep_distribution); get_distribution is valid at once."""
        _l.check(self.L.ss_keep_distribution(self._h, 1 if on else 0))

    def resort(self):
        """Start a new re-sort period at the next step (ss_resort)."""
        _l.check(self.L.ss_resort(self._h))

    def prepare_steps(self, nsteps):
        """Capture the graphs the next ``step(nsteps)`` replays, running
        nothing (ss_prepare_steps): the set-up of a run of steps."""
        _l.check(self.L.ss_prepare_steps(self._h, int(nsteps)))

    # getters ----------------------------------------------------------------
    def get_timestep(self):
        t = C.c_uint64()
        _l.check(self.L.ss_get_timestep(self._h, C.byref(t)))
        return t.value

    def get_particles(self):
        return self.get_particles_head(self.n)

    def get_particles_head(self, n):
        out = np.zeros((n, 5))
        _l.check(self.L.ss_get_particles(self._h, _l.dptr(out), n))
        return out

    def get_counts(self):
        p = self.params
        out = np.zeros(p.dist_shape, dtype=np.uint32)
        _l.check(self.L.ss_get_distribution(self._h, out.ctypes.data_as(_P(C.c_uint32)), None))
        return out

    def get_distribution(self):
        p = self.params
        out = np.zeros(p.dist_shape)
        _l.check(self.L.ss_get_distribution(self._h, None, _l.dptr(out)))
        return out

    def get_flow_field(self, with_gradient=False):
        p = self.params
        u = np.zeros((3,) + p.grid_shape)
        g = np.zeros((3, 3) + p.grid_shape) if with_gradient else None
        _l.check(self.L.ss_get_flow_field(self._h, _l.dptr(u), None if g is None else _l.dptr(g)))
        return (u, g) if with_gradient else u

    def get_magnetic_field(self, with_gradient=False):
        p = self.params
        b = np.zeros((3,) + p.grid_shape)
        g = np.zeros((3, 3) + p.grid_shape) if with_gradient else None
        _l.check(self.L.ss_get_magnetic_field(self._h, _l.dptr(b), None if g is None else _l.dptr(g)))
        return (b, g) if with_gradient else b

    def get_diagnostics(self):
        """FP64 global diagnostics: dict with 'mean_orientation' (3),
        'mean_polarisation' (3, bin centres) and 'mean_stress' (3x3)."""
        out = np.zeros(15)
        _l.check(self.L.ss_get_diagnostics(self._h, _l.dptr(out)))
        return {
            "mean_orientation": out[0:3],
            "mean_polarisation": out[3:6],
            "mean_stress": out[6:15].reshape(3, 3),
        }

    def get_info(self):
        """The execution plan chosen at creation (ss_get_info) as a dict."""
        i = Info()
        _l.check(self.L.ss_get_info(self._h, C.byref(i)))
        return {
            "sorted": bool(i.sorted),
            "fused_spectral": bool(i.fused_spectral),
            "paired_spectral": i.fused_spectral == 2,
            "tile": tuple(i.tile),
            "ntiles": i.ntiles,
            "resort_interval": i.resort_interval,
            "pencils": i.pencils,
            "graphs": bool(i.graphs),
            "update_groups": i.update_groups,
            "plain_tiled": bool(i.plain_tiled),
        }

    def profile_steps(self, nsteps, max_stages=16):
        """Runs nsteps steps timing every stage with HIP events on the
        simulation's stream; returns {stage: mean ms per step}."""
        ms = np.zeros(max_stages)
        stride = 32
        names = C.create_string_buffer(stride * max_stages)
        k = _l.check(self.L.ss_profile_steps(self._h, nsteps, _l.dptr(ms), names, stride, max_stages))
        raw = names.raw
        out = {}
        for i in range(k):
            nm = raw[i * stride : (i + 1) * stride].split(b"\0", 1)[0].decode()
            out[nm] = float(ms[i])
        return out


class _Member(Simulation):
    """Borrowed member handle of an Ensemble (not destroyed by Python)."""

    def __init__(self, ensemble, params, handle):
        self.L = _l.load()
        self.params = params
        self._h = handle
        self.n = params.number_of_particles
        self._ens = ensemble

    def close(self):
        self._h = None

    def do_timestep(self):
        raise RuntimeError("step an ensemble member through Ensemble.step")

    step = do_timestep


class Ensemble:
    """Independent systems stepped together (ss_ensemble_*): one persistent
    update launch over all members per step, the members' spectral solves as
    parallel graph branches.  ``members[j]`` has the Simulation getters."""

    def __init__(self, params_list):
        self.L = _l.load()
        self.params = list(params_list)
        arr = (Params * len(self.params))(*self.params)
        h = C.c_void_p()
        _l.check(self.L.ss_ensemble_create(arr, len(self.params), C.byref(h)))
        self._h = h
        self.members = []
        for j, p in enumerate(self.params):
            mh = C.c_void_p()
            _l.check(self.L.ss_ensemble_member(self._h, j, C.byref(mh)))
            self.members.append(_Member(self, p, mh))

    def step(self, nsteps=1):
        t = C.c_uint64()
        _l.check(self.L.ss_ensemble_step(self._h, nsteps, C.byref(t)))
        return t.value

    def profile_steps(self, nsteps, max_stages=8):
        """nsteps eager steps timed per stage (ss_ensemble_profile_steps):
        {"spectral", "update", "sort": mean ms per step}."""
        ms = np.zeros(max_stages)
        stride = 32
        names = C.create_string_buffer(stride * max_stages)
        k = _l.check(self.L.ss_ensemble_profile_steps(self._h, nsteps, _l.dptr(ms), names, stride, max_stages))
        raw = names.raw
        return {raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode(): float(ms[i]) for i in range(k)}

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            self.L.ss_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def initial_condition(params, amplitude=0.05, seed=None, n=None):
    """Uniform positions, isotropic orientations with a polar perturbation
    amplitude*sin(2 pi x / Lx) along z (N x 5, serde order)."""
    L = _l.load()
    n = params.number_of_particles if n is None else n
    seed = params.seed if seed is None else seed
    out = np.zeros((n, 5))
    _l.check(L.ss_initial_condition(C.byref(params), amplitude, seed, _l.dptr(out), n))
    return out


def create_isotropic(n, box, seed, device=0):
    """Particle::create_isotropic (src/particle.rs:224-255) from the
    reference's Pcg64Mcg stream, drawn on the GPU: N x 5 [x, y, z, phi, theta]."""
    L = _l.load()
    out = np.zeros((n, 5))
    b = Box(*box)
    _l.check(L.ss_create_isotropic(n, C.byref(b), seed, device, _l.dptr(out)))
    return out


def create_homogeneous(n, kappa, box, seed, device=0):
    """Particle::create_homogeneous (src/particle.rs:257-297): polar angle from
    sin(theta) exp(kappa cos theta) about the y axis, same stream."""
    L = _l.load()
    out = np.zeros((n, 5))
    b = Box(*box)
    _l.check(L.ss_create_homogeneous(n, kappa, C.byref(b), seed, device, _l.dptr(out)))
    return out


def density_gradient(params, dist):
    """DensityGradient::get_gradient + get_real_gradient
    (src/distribution/density_gradient.rs): gradient of the density of a
    normalised distribution, real parts, shape (3, nx, ny, nz)."""
    L = _l.load()
    d = _f64(dist)
    out = np.zeros((3,) + params.grid_shape)
    _l.check(L.ss_density_gradient(C.byref(params), _l.dptr(d), _l.dptr(out)))
    return out


def interpolate_vector_field(params, field, positions):
    """interpolate_vector_field (src/mesh/interpolate.rs:11-39): trilinear
    interpolation of a cell-centred periodic field (3, nx, ny, nz) at
    positions (n, 3); the z offset comes from position.z (the reference
    takes position.y)."""
    L = _l.load()
    f = _f64(field)
    q = _f64(positions).reshape(-1, 3)
    out = np.zeros((q.shape[0], 3))
    _l.check(L.ss_interpolate_vector_field(C.byref(params), _l.dptr(f), _l.dptr(q), q.shape[0], _l.dptr(out)))
    return out


def mean_flow_field(params, dist):
    """SpectralSolver::new + mean_flow_field (spectral_solver.rs:33-68,233-254)
    on a given distribution; returns (u [3,...], grad u [3,3,...]) real parts."""
    L = _l.load()
    d = _f64(dist)
    u = np.zeros((3,) + params.grid_shape)
    g = np.zeros((3, 3) + params.grid_shape)
    _l.check(L.ss_flow_solve(C.byref(params), _l.dptr(d), _l.dptr(u), _l.dptr(g)))
    return u, g


def mean_magnetic_field(params, dist):
    """MagneticSolver::mean_magnetic_field (magnetic_solver.rs:139-161)."""
    L = _l.load()
    d = _f64(dist)
    b = np.zeros((3,) + params.grid_shape)
    g = np.zeros((3, 3) + params.grid_shape)
    _l.check(L.ss_magnetic_solve(C.byref(params), _l.dptr(d), _l.dptr(b), _l.dptr(g)))
    return b, g


class _Solver:
    KIND = 0

    def __init__(self, params):
        self.L = _l.load()
        self.params = params
        h = C.c_void_p()
        _l.check(self.L.ss_solver_create(C.byref(params), self.KIND, C.byref(h)))
        self._h = h

    def _out(self, with_gradient):
        p = self.params
        a = np.zeros((3,) + p.grid_shape)
        g = np.zeros((3, 3) + p.grid_shape) if with_gradient else None
        return a, g

    def close(self):
        if getattr(self, "_h", None):
            self.L.ss_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class SpectralSolver(_Solver):
    """SpectralSolver (spectral_solver.rs:20-258): planned once for the
    params' grid, box, stress and precision, then mean_flow_field(screening,
    dist) -> (u, grad_u) real parts for any number of distributions."""

    KIND = 0

    def mean_flow_field(self, hydro_screening, dist, with_gradient=True):
        d = _f64(dist, self.params.dist_shape)
        u, g = self._out(with_gradient)
        _l.check(self.L.ss_solver_mean_flow_field(self._h, float(hydro_screening), _l.dptr(d), _l.dptr(u),
                                                  None if g is None else _l.dptr(g)))
        return (u, g) if with_gradient else u


class MagneticSolver(_Solver):
    """MagneticSolver (magnetic_solver.rs:20-169): planned once, then
    mean_magnetic_field(dist) -> (B, grad_B) real parts."""

    KIND = 1

    def mean_magnetic_field(self, dist, with_gradient=True):
        d = _f64(dist, self.params.dist_shape)
        b, g = self._out(with_gradient)
        _l.check(self.L.ss_solver_mean_magnetic_field(self._h, _l.dptr(d), _l.dptr(b),
                                                      None if g is None else _l.dptr(g)))
        return (b, g) if with_gradient else b


class FFTPlan:
    """FFTPlan (crates/fftw3/src/fft.rs:33-244): unnormalised in-place c2c 3-D,
    ``direction`` -1 forward / +1 backward, batched over leading arrays.
    ``variant``: FFT_AUTO / FFT_RADIX (LDS Stockham) or FFT_DFT (FP32 only:
    axes of 2..64 points as dense DFT products on the matrix cores, 3-pass
    split bf16 with FP32 accumulation)."""

    def __init__(self, shape, direction, batch=1, precision=0, device=0, variant=0):
        self.L = _l.load()
        self.shape = tuple(shape)
        self.batch = batch
        self.precision = precision
        h = C.c_void_p()
        nx, ny, nz = self.shape
        _l.check(
            self.L.ss_fft_plan_create(nx, ny, nz, direction, batch, precision, variant, device, C.byref(h))
        )
        self._h = h

    def reexecute3d(self, a):
        """Transforms a (complex, shape (batch,)+shape or shape) in place."""
        dt = np.complex128 if self.precision == 0 else np.complex64
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError(f"array must be C-contiguous {np.dtype(dt).name}")
        _l.check(self.L.ss_fft_execute_host(self._h, a.ctypes.data_as(C.c_void_p)))
        return a

    def reexecute3d_noalloc(self, a):
        """Same transform through ss_fft_execute_host_noalloc: staged in the
        library's per-GPU buffer instead of a device allocation per call."""
        dt = np.complex128 if self.precision == 0 else np.complex64
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError(f"array must be C-contiguous {np.dtype(dt).name}")
        _l.check(self.L.ss_fft_execute_host_noalloc(self._h, a.ctypes.data_as(C.c_void_p)))
        return a

    def close(self):
        if getattr(self, "_h", None):
            self.L.ss_fft_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def debug_update(params, particles, fields, rv):
    """The device particle update on explicit per-particle inputs
    (ss_debug_update); fields: n x 34, rv: n x 5."""
    L = _l.load()
    p = _f64(particles)
    f = _f64(fields)
    r = _f64(rv)
    out = np.zeros_like(p)
    _l.check(L.ss_debug_update(C.byref(params), p.shape[0], _l.dptr(p), _l.dptr(f), _l.dptr(r), _l.dptr(out)))
    return out


def debug_coord_to_grid(params, particles):
    """The device binning rule (ss_debug_coord_to_grid); returns n x 5 int32."""
    L = _l.load()
    p = _f64(particles)
    out = np.zeros(p.shape, dtype=np.int32)
    _l.check(L.ss_debug_coord_to_grid(C.byref(params), p.shape[0], _l.dptr(p), out.ctypes.data_as(_P(C.c_int32))))
    return out

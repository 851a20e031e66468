"""ctypes mirror of ``ss_params`` (include/ss_api.h) and the Python-side
parameter constructor.

The fields follow ``Parameters`` and ``SimulationSettings`` of the reference
(src/bin/simulation/settings/mod.rs:51-67, 108-118) plus the build's knobs.
"""
import ctypes as C

FP64 = 0
FP32 = 1
NOISE_PHILOX = 0
NOISE_EXTERNAL = 1
NOISE_REFERENCE = 2
BIN_AUTO = 0
BIN_HISTOGRAM = 1
BIN_SORTED = 2
BIN_PLAIN_TILED = 3
FFT_AUTO = 0
FFT_RADIX = 1
FFT_DFT = 2
FLAG_QUASI2D = 1
FLAG_MAGNETIC_ALWAYS = 2
FLAG_KEEP_DISTRIBUTION = 4
FLAG_NO_GRAPH = 8
FLAG_AXIS_PASSES = 16
FLAG_PAIRED_SPECTRAL = 32

SS_OK = 0
ERRORS = {
    -1: "SS_ERR_INVALID",
    -2: "SS_ERR_DEVICE",
    -3: "SS_ERR_STATE",
    -4: "SS_ERR_UNSUPPORTED",
    -5: "SS_ERR_ALLOC",
}


class Box(C.Structure):
    """BoxSize (lib.rs:35-41)."""

    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Grid(C.Structure):
    """GridSize (lib.rs:43-51)."""

    _fields_ = [
        ("x", C.c_int32),
        ("y", C.c_int32),
        ("z", C.c_int32),
        ("phi", C.c_int32),
        ("theta", C.c_int32),
    ]


class Params(C.Structure):
    _fields_ = [
        ("box_size", Box),
        ("grid_size", Grid),
        ("number_of_particles", C.c_uint64),
        ("timestep", C.c_double),
        ("seed", C.c_uint64),
        ("magnetic_drag", C.c_double),
        ("shape", C.c_double),
        ("hydro_screening", C.c_double),
        ("volume_exclusion", C.c_double),
        ("magnetic_reorientation", C.c_double),
        ("diffusion_translational", C.c_double),
        ("diffusion_rotational", C.c_double),
        ("stress_active", C.c_double),
        ("stress_magnetic", C.c_double),
        ("magnetic_dipole_dipole", C.c_double),
        ("field_direction", C.c_double * 3),
        ("precision", C.c_int32),
        ("noise_mode", C.c_int32),
        ("binning", C.c_int32),
        ("fft_variant", C.c_int32),
        ("flags", C.c_uint32),
        ("device", C.c_int32),
        ("tile", C.c_int32 * 3),
        ("resort_interval", C.c_int32),
        ("noise_streams", C.c_int32),
    ]

    @property
    def ns(self):
        g = self.grid_size
        return g.x * g.y * g.z

    @property
    def na(self):
        g = self.grid_size
        return g.phi * g.theta

    @property
    def grid_shape(self):
        g = self.grid_size
        return (g.x, g.y, g.z)

    @property
    def dist_shape(self):
        g = self.grid_size
        return (g.x, g.y, g.z, g.phi, g.theta)

    @property
    def box(self):
        b = self.box_size
        return (b.x, b.y, b.z)


class Info(C.Structure):
    """ss_info: the execution plan chosen at ss_create."""

    _fields_ = [
        ("sorted", C.c_int32),
        ("fused_spectral", C.c_int32),
        ("tile", C.c_int32 * 3),
        ("ntiles", C.c_int32),
        ("resort_interval", C.c_int32),
        ("pencils", C.c_int32),
        ("graphs", C.c_int32),
        ("update_groups", C.c_int32),
        ("plain_tiled", C.c_int32),
    ]


def make_params(
    box=(10.0, 10.0, 10.0),
    grid=(16, 16, 16, 8, 8),
    n=1 << 16,
    timestep=1e-3,
    seed=1,
    magnetic_drag=0.0,
    shape=0.0,
    hydro_screening=0.0,
    volume_exclusion=0.0,
    magnetic_reorientation=0.0,
    diffusion_translational=0.1,
    diffusion_rotational=0.5,
    stress_active=-1.0,
    stress_magnetic=0.0,
    magnetic_dipole_dipole=0.0,
    field_direction=(0.0, 0.0, 0.0),
    precision=FP64,
    noise_mode=NOISE_PHILOX,
    binning=BIN_AUTO,
    fft_variant=FFT_AUTO,
    flags=0,
    device=0,
    tile=(0, 0, 0),
    resort_interval=0,
    noise_streams=0,
):
    """Build a ``Params``; defaults are the small CPU-runnable configuration
    (L=10, 16^3 x 8 x 8, 2^16 particles, dt=1e-3, D_t=0.1, D_r=0.5, alpha=-1)."""
    p = Params()
    p.box_size = Box(*box)
    p.grid_size = Grid(*grid)
    p.number_of_particles = n
    p.timestep = timestep
    p.seed = seed
    p.magnetic_drag = magnetic_drag
    p.shape = shape
    p.hydro_screening = hydro_screening
    p.volume_exclusion = volume_exclusion
    p.magnetic_reorientation = magnetic_reorientation
    p.diffusion_translational = diffusion_translational
    p.diffusion_rotational = diffusion_rotational
    p.stress_active = stress_active
    p.stress_magnetic = stress_magnetic
    p.magnetic_dipole_dipole = magnetic_dipole_dipole
    p.field_direction = (C.c_double * 3)(*field_direction)
    p.precision = precision
    p.noise_mode = noise_mode
    p.binning = binning
    p.fft_variant = fft_variant
    p.flags = flags
    p.device = device
    p.tile = (C.c_int32 * 3)(*tile)
    p.resort_interval = resort_interval
    p.noise_streams = noise_streams
    return p

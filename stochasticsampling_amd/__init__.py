"""MI355X-native hot loop of stochasticsampling: Python mirror of the
reference's Simulation / SpectralSolver / FFTPlan API over libss (HIP)."""
from .params import (  # noqa: F401
    BIN_AUTO,
    BIN_HISTOGRAM,
    BIN_SORTED,
    FFT_AUTO,
    FFT_DFT,
    FFT_RADIX,
    FLAG_AXIS_PASSES,
    FLAG_KEEP_DISTRIBUTION,
    FLAG_MAGNETIC_ALWAYS,
    FLAG_NO_GRAPH,
    FLAG_QUASI2D,
    FLAG_PAIRED_SPECTRAL,
    FP32,
    FP64,
    NOISE_EXTERNAL,
    NOISE_PHILOX,
    NOISE_REFERENCE,
    Params,
    make_params,
)
from .lib import SSError, device_count, load  # noqa: F401
from .simulation import (  # noqa: F401
    Ensemble,
    FFTPlan,
    Simulation,
    Snapshot,
    create_homogeneous,
    create_isotropic,
    debug_coord_to_grid,
    debug_update,
    density_gradient,
    initial_condition,
    interpolate_vector_field,
    mean_flow_field,
    mean_magnetic_field,
)

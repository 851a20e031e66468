"""Parameter files of the reference (TOML), read into ``Settings`` and turned
into ``Params`` for the C ABI.

Mirrors src/bin/simulation/settings/mod.rs:24-205: the same sections and
keys, the same defaults (``#[serde(default)]``), unknown keys rejected at
every level (``#[serde(deny_unknown_fields)]``), the same sanity checks
(``check_settings``), and the SI-unit variant of settings/si.rs:24-199
(``read_parameter_file_si`` followed by ``into_settings``).  Errors raise
``SettingsError``.
"""
import math
from dataclasses import dataclass, field
from typing import Optional

try:
    import tomllib as _toml  # Python 3.11+
except ImportError:  # pragma: no cover - this image has tomli
    import tomli as _toml

from .params import FP32, FP64, make_params

INIT_DISTRIBUTIONS = ("Isotropic", "Homogeneous")
OUTPUT_FORMATS = ("CBOR", "Bincode", "MsgPack")
DEFAULT_IO_QUEUE_SIZE = 1
DEFAULT_OUTPUT_FORMAT = "MsgPack"
DEFAULT_INIT_TYPE = "Isotropic"
BOLTZMANN = 1.380_648_52e-23


class SettingsError(ValueError):
    """A parameter file that the reference would reject."""


def _take(table, path, spec):
    """Check ``table`` against ``spec`` {key: (kind, default)} where kind is
    'float', 'uint', 'bool', 'str', 'table' or a tuple of enum names, and
    default is REQUIRED or a value; returns a dict of converted values."""
    if not isinstance(table, dict):
        raise SettingsError(f"{path}: expected a table")
    unknown = sorted(set(table) - set(spec))
    if unknown:
        raise SettingsError(f"{path}: unknown field(s) {', '.join(unknown)}")
    out = {}
    for key, (kind, default) in spec.items():
        where = f"{path}.{key}" if path else key
        if key not in table:
            if default is REQUIRED:
                raise SettingsError(f"{where}: missing field")
            out[key] = default
            continue
        v = table[key]
        if kind == "float":
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise SettingsError(f"{where}: expected a number")
            out[key] = float(v)
        elif kind == "uint":
            if isinstance(v, bool) or not isinstance(v, int) or v < 0:
                raise SettingsError(f"{where}: expected a non-negative integer")
            out[key] = v
        elif kind == "bool":
            if not isinstance(v, bool):
                raise SettingsError(f"{where}: expected a boolean")
            out[key] = v
        elif kind == "str":
            if not isinstance(v, str):
                raise SettingsError(f"{where}: expected a string")
            out[key] = v
        elif kind == "table":
            if not isinstance(v, dict):
                raise SettingsError(f"{where}: expected a table")
            out[key] = v
        elif isinstance(kind, tuple):
            if v not in kind:
                raise SettingsError(f"{where}: unknown variant {v!r}, expected one of {', '.join(kind)}")
            out[key] = v
        else:  # pragma: no cover
            raise AssertionError(kind)
    return out


REQUIRED = object()


@dataclass
class BoxSize:
    x: float
    y: float
    z: float


@dataclass
class GridSize:
    x: int
    y: int
    z: int
    phi: int
    theta: int


@dataclass
class Output:
    """Output schedule (settings/mod.rs:71-89): every how many steps."""

    initial_condition: bool = True
    final_snapshot: bool = True
    distribution: Optional[int] = None
    flowfield: Optional[int] = None
    magneticfield: Optional[int] = None
    particles_head: Optional[int] = None
    particles: Optional[int] = None
    snapshot: Optional[int] = None


@dataclass
class SimulationSettings:
    number_of_particles: int
    number_of_timesteps: int
    timestep: float
    seed: int
    output_at_timestep: Output
    box_size: BoxSize
    grid_size: GridSize
    init_distribution: str = DEFAULT_INIT_TYPE


@dataclass
class DiffusionConstants:
    translational: float
    rotational: float


@dataclass
class StressPrefactors:
    active: float
    magnetic: float


@dataclass
class Parameters:
    magnetic_reorientation: float
    diffusion: DiffusionConstants
    stress: StressPrefactors
    magnetic_dipole_dipole: float = 0.0
    magnetic_drag: float = 0.0
    shape: float = 0.0
    hydro_screening: float = 0.0
    volume_exclusion: float = 0.0


@dataclass
class EnvironmentSettings:
    prefix: str
    version: str = ""
    init_file: Optional[str] = None
    io_queue_size: int = DEFAULT_IO_QUEUE_SIZE
    output_format: str = DEFAULT_OUTPUT_FORMAT


@dataclass
class Settings:
    simulation: SimulationSettings
    parameters: Parameters
    environment: EnvironmentSettings

    def set_version(self, version):
        self.environment.version = version

    def to_params(self, precision=FP32, **overrides):
        """``Params`` of this run (``Simulation::new``,
        src/bin/simulation/mod.rs:91-160); ``overrides`` go to make_params
        (binning, flags, tile, ...)."""
        s, q = self.simulation, self.parameters
        kw = dict(
            box=(s.box_size.x, s.box_size.y, s.box_size.z),
            grid=(s.grid_size.x, s.grid_size.y, s.grid_size.z, s.grid_size.phi, s.grid_size.theta),
            n=s.number_of_particles,
            timestep=s.timestep,
            seed=s.seed,
            magnetic_drag=q.magnetic_drag,
            shape=q.shape,
            hydro_screening=q.hydro_screening,
            volume_exclusion=q.volume_exclusion,
            magnetic_reorientation=q.magnetic_reorientation,
            diffusion_translational=q.diffusion.translational,
            diffusion_rotational=q.diffusion.rotational,
            stress_active=q.stress.active,
            stress_magnetic=q.stress.magnetic,
            magnetic_dipole_dipole=q.magnetic_dipole_dipole,
            precision=precision,
        )
        kw.update(overrides)
        return make_params(**kw)

    def to_serde(self):
        """The settings as the reference serialises them (field order of the
        Rust structs; enums as variant names), for output headers."""
        s, q, e = self.simulation, self.parameters, self.environment
        o = s.output_at_timestep
        return {
            "simulation": {
                "number_of_particles": s.number_of_particles,
                "number_of_timesteps": s.number_of_timesteps,
                "timestep": s.timestep,
                "init_distribution": ("enum", s.init_distribution, INIT_DISTRIBUTIONS),
                "seed": s.seed,
                "output_at_timestep": {
                    "initial_condition": o.initial_condition,
                    "final_snapshot": o.final_snapshot,
                    "distribution": ("option", o.distribution),
                    "flowfield": ("option", o.flowfield),
                    "magneticfield": ("option", o.magneticfield),
                    "particles_head": ("option", o.particles_head),
                    "particles": ("option", o.particles),
                    "snapshot": ("option", o.snapshot),
                },
                "box_size": {"x": s.box_size.x, "y": s.box_size.y, "z": s.box_size.z},
                "grid_size": {"x": s.grid_size.x, "y": s.grid_size.y, "z": s.grid_size.z,
                              "phi": s.grid_size.phi, "theta": s.grid_size.theta},
            },
            "parameters": {
                "magnetic_drag": q.magnetic_drag,
                "shape": q.shape,
                "hydro_screening": q.hydro_screening,
                "volume_exclusion": q.volume_exclusion,
                "magnetic_reorientation": q.magnetic_reorientation,
                "diffusion": {"translational": q.diffusion.translational, "rotational": q.diffusion.rotational},
                "stress": {"active": q.stress.active, "magnetic": q.stress.magnetic},
                "magnetic_dipole": {"magnetic_dipole_dipole": q.magnetic_dipole_dipole},
            },
            "environment": {
                "prefix": e.prefix,
                "version": e.version,
                "init_file": ("option", e.init_file),
                "io_queue_size": e.io_queue_size,
                "output_format": ("enum", e.output_format, OUTPUT_FORMATS),
            },
        }


def _simulation(t):
    s = _take(t, "simulation", {
        "number_of_particles": ("uint", REQUIRED),
        "number_of_timesteps": ("uint", REQUIRED),
        "timestep": ("float", REQUIRED),
        "init_distribution": (INIT_DISTRIBUTIONS, DEFAULT_INIT_TYPE),
        "seed": ("uint", REQUIRED),
        "output_at_timestep": ("table", REQUIRED),
        "box_size": ("table", REQUIRED),
        "grid_size": ("table", REQUIRED),
    })
    o = _take(s["output_at_timestep"], "simulation.output_at_timestep", {
        "initial_condition": ("bool", True),
        "final_snapshot": ("bool", True),
        "distribution": ("uint", None),
        "flowfield": ("uint", None),
        "magneticfield": ("uint", None),
        "particles_head": ("uint", None),
        "particles": ("uint", None),
        "snapshot": ("uint", None),
    })
    b = _take(s["box_size"], "simulation.box_size", {k: ("float", REQUIRED) for k in "xyz"})
    g = _take(s["grid_size"], "simulation.grid_size",
              {k: ("uint", REQUIRED) for k in ("x", "y", "z", "phi", "theta")})
    return SimulationSettings(
        number_of_particles=s["number_of_particles"], number_of_timesteps=s["number_of_timesteps"],
        timestep=s["timestep"], seed=s["seed"], output_at_timestep=Output(**o), box_size=BoxSize(**b),
        grid_size=GridSize(**g), init_distribution=s["init_distribution"])


def _environment(t):
    e = _take(t, "environment", {
        "prefix": ("str", REQUIRED),
        "init_file": ("str", None),
        "io_queue_size": ("uint", DEFAULT_IO_QUEUE_SIZE),
        "output_format": (OUTPUT_FORMATS, DEFAULT_OUTPUT_FORMAT),
    })
    # `version` is skip_deserializing in the reference: a key of that name is unknown
    return EnvironmentSettings(prefix=e["prefix"], init_file=e["init_file"], io_queue_size=e["io_queue_size"],
                               output_format=e["output_format"])


def check_settings(s):
    """settings/mod.rs:180-205."""
    b = s.simulation.box_size
    if b.x <= 0.0 or b.y <= 0.0 or b.z <= 0.0:
        raise SettingsError(f"Box size is invalid. Must be bigger than 0: {b}")
    o = s.simulation.output_at_timestep
    if o.particles_head is not None and s.simulation.number_of_particles < o.particles_head:
        raise SettingsError("Cannot output more particles than available. `particles_head` "
                            "must be smaller or equal to `number_of_particles`")


def _load(path, text):
    try:
        if text is not None:
            return _toml.loads(text)
        with open(path, "rb") as f:
            return _toml.load(f)
    except _toml.TOMLDecodeError as e:
        raise SettingsError(f"Unable to parse parameter file: {e}") from e
    except OSError as e:
        raise SettingsError(f"Unable to read parameter file: {e}") from e


def read_parameter_file(path=None, text=None):
    """``read_parameter_file`` (settings/mod.rs:163-178) of a file, or of the
    TOML ``text`` itself."""
    top = _take(_load(path, text), "", {
        "simulation": ("table", REQUIRED),
        "parameters": ("table", REQUIRED),
        "environment": ("table", REQUIRED),
    })
    p = _take(top["parameters"], "parameters", {
        "magnetic_drag": ("float", 0.0),
        "shape": ("float", 0.0),
        "hydro_screening": ("float", 0.0),
        "volume_exclusion": ("float", 0.0),
        "magnetic_reorientation": ("float", REQUIRED),
        "diffusion": ("table", REQUIRED),
        "stress": ("table", REQUIRED),
        "magnetic_dipole": ("table", REQUIRED),
    })
    d = _take(p["diffusion"], "parameters.diffusion", {"translational": ("float", REQUIRED),
                                                       "rotational": ("float", REQUIRED)})
    st = _take(p["stress"], "parameters.stress", {"active": ("float", REQUIRED), "magnetic": ("float", REQUIRED)})
    md = _take(p["magnetic_dipole"], "parameters.magnetic_dipole", {"magnetic_dipole_dipole": ("float", 0.0)})
    settings = Settings(
        simulation=_simulation(top["simulation"]),
        parameters=Parameters(
            magnetic_reorientation=p["magnetic_reorientation"], diffusion=DiffusionConstants(**d),
            stress=StressPrefactors(**st), magnetic_dipole_dipole=md["magnetic_dipole_dipole"],
            magnetic_drag=p["magnetic_drag"], shape=p["shape"], hydro_screening=p["hydro_screening"],
            volume_exclusion=p["volume_exclusion"]),
        environment=_environment(top["environment"]),
    )
    check_settings(settings)
    return settings


def read_parameter_file_si(path=None, text=None):
    """``si::read_parameter_file`` + ``into_settings`` (si.rs:87-199):
    parameters in SI units converted to simulation units."""
    top = _take(_load(path, text), "", {
        "simulation": ("table", REQUIRED),
        "parameters": ("table", REQUIRED),
        "environment": ("table", REQUIRED),
    })
    p = _take(top["parameters"], "parameters", {
        "hydro_screening": ("float", REQUIRED),
        "volume_exclusion": ("float", REQUIRED),
        "viscocity": ("float", REQUIRED),
        "temperature": ("float", REQUIRED),
        "volume_fraction": ("float", REQUIRED),
        "external_field": ("float", REQUIRED),
        "particle": ("table", REQUIRED),
    })
    pa = _take(p["particle"], "parameters.particle", {
        "radius": ("float", REQUIRED),
        "shape": ("float", 0.0),
        "self_propulsion_speed": ("float", REQUIRED),
        "force_dipole": ("float", REQUIRED),
        "magnetic_dipole_moment": ("float", REQUIRED),
        "persistance_time": ("float", REQUIRED),
    })
    sim = _simulation(top["simulation"])
    env = _environment(top["environment"])
    si = Settings(simulation=sim, parameters=None, environment=env)
    check_settings(si)
    return _into_settings(sim, env, p, pa)


def _into_settings(sim, env, p, pa):
    """si.rs:126-194, operation for operation in FP64."""
    pi = math.pi
    r = pa["radius"]
    volp = 4.0 / 3.0 * pi * (r * r * r)  # powi(3)
    number_density = p["volume_fraction"] / volp
    xc = number_density ** (-1.0 / 3.0)
    uc = pa["self_propulsion_speed"]
    tc = xc / uc
    stressf = number_density ** (2.0 / 3.0) / pa["self_propulsion_speed"] / p["viscocity"]
    stress = StressPrefactors(active=stressf * pa["force_dipole"],
                              magnetic=stressf * pa["magnetic_dipole_moment"] * p["external_field"])
    rotfriction = 8.0 * pi * p["viscocity"] * (r * r * r)
    transfriction = 6.0 * pi * p["viscocity"] * r
    m = pa["magnetic_dipole_moment"]
    rotdiff_brown = BOLTZMANN * p["temperature"] / rotfriction
    rotdiff_active = 1.0 / 2.0 / pa["persistance_time"]
    transdiff_brown = BOLTZMANN * p["temperature"] / transfriction
    diff = DiffusionConstants(
        translational=number_density ** (1.0 / 3.0) / uc * transdiff_brown,
        rotational=number_density ** (-1.0 / 3.0) / uc * (rotdiff_brown + rotdiff_active))
    alignment = pa["magnetic_dipole_moment"] * p["external_field"] / rotfriction / (rotdiff_brown + rotdiff_active)
    params = Parameters(
        magnetic_reorientation=alignment * diff.rotational,
        diffusion=diff,
        stress=stress,
        magnetic_dipole_dipole=number_density ** (2.0 / 3.0) / uc / rotfriction * 4.0e-7 * pi * (m * m),
        volume_exclusion=p["volume_exclusion"],
        shape=pa["shape"],
        hydro_screening=p["hydro_screening"],
        magnetic_drag=number_density / uc / transfriction * 4.0e-7 * pi * (m * m),
    )
    sim.box_size = BoxSize(sim.box_size.x / xc, sim.box_size.y / xc, sim.box_size.z / xc)
    sim.timestep = sim.timestep / tc
    return Settings(simulation=sim, parameters=params, environment=env)

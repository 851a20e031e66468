"""Host driver of one run over the C ABI: initial condition, output schedule,
output files and snapshots, after the reference's ``init_simulation``
(src/bin/init.rs:51-115) and ``run_simulation`` (src/bin/main.rs:139-289).

Behaviour kept from the reference:

* the loop runs timesteps ``start+1 ..= start+1+number_of_timesteps``, i.e.
  one step more than ``number_of_timesteps`` (main.rs:176-182);
* an entry is written when it holds a distribution, a flow field or
  particles (a magnetic field alone is not written, main.rs:256);
* ``particles_head`` limits the particles written; the initial entry holds
  the distribution and, when particle output is on, the particles;
* a snapshot every ``snapshot`` steps and after the last step when
  ``final_snapshot``;
* the isotropic and homogeneous initial conditions are the reference's own
  streams (Pcg64Mcg::seed_from_u64(seed), particle.rs:224-297), drawn on
  the GPU.

Noise: ``noise="philox"`` (default) is the counter-based stream keyed by
(seed, step, particle), whose snapshots carry no generator state;
``noise="reference"`` is the reference's own Pcg32 streams (one per rayon
thread, ``noise_streams`` or the environment's RAYON_NUM_THREADS like
mod.rs:113-116), whose generator states go into every snapshot.
"""
import os
import time

import numpy as np

from . import output as out
from .params import FP32, NOISE_REFERENCE
from .simulation import Simulation, Snapshot
from .simulation import create_homogeneous as _gpu_homogeneous
from .simulation import create_isotropic as _gpu_isotropic

VERSION = "0.1.0"


def create_isotropic(n, box, seed, device=0):
    """Particle::create_isotropic (particle.rs:224-255)."""
    return _gpu_isotropic(n, box, seed, device)


def create_homogeneous(n, kappa, box, seed, device=0):
    """Particle::create_homogeneous (particle.rs:257-297); the reference
    panics for kappa == 0 and for an angle that is NaN, this raises."""
    if kappa == 0.0:
        raise ValueError("Alignment of zero is the isotropic state. Please use that instead.")
    return _gpu_homogeneous(n, kappa, box, seed, device)


def output_base(out_dir, prefix, when=None):
    """OutputPath (src/bin/output/path.rs:15-62): ``out_dir/<id>/<id>`` with
    id = ``<prefix>-<YYYY-mm-dd_HHMMSS>_v<version>``; the directory is
    created."""
    stamp = time.strftime("%Y-%m-%d_%H%M%S", time.localtime(when))
    oid = f"{prefix}-{stamp}_v{VERSION.replace('.', '_')}"
    d = os.path.join(out_dir, oid)
    os.makedirs(d, exist_ok=False)
    return os.path.join(d, oid)


def save_toml(settings, path):
    """Settings::save_to_file (settings/mod.rs:213-224), in simulation units."""
    def val(v):
        if isinstance(v, bool):
            return "true" if v else "false"
        if isinstance(v, float):
            return repr(v)
        if isinstance(v, int):
            return str(v)
        return '"' + str(v).replace("\\", "\\\\").replace('"', '\\"') + '"'

    lines = []

    def table(name, d):
        scalars = [(k, v) for k, v in d.items() if not isinstance(v, dict)]
        subs = [(k, v) for k, v in d.items() if isinstance(v, dict)]
        lines.append(f"[{name}]")
        for k, v in scalars:
            if isinstance(v, tuple) and v[0] == "option":
                if v[1] is not None:
                    lines.append(f"{k} = {val(v[1])}")
            elif isinstance(v, tuple) and v[0] == "enum":
                lines.append(f"{k} = {val(v[1])}")
            else:
                lines.append(f"{k} = {val(v)}")
        lines.append("")
        for k, v in subs:
            table(f"{name}.{k}", v)

    ser = settings.to_serde()
    env = dict(ser["environment"])
    env.pop("version")  # skip_deserializing: not a key of a parameter file
    table("environment", env)
    table("simulation", ser["simulation"])
    table("parameters", ser["parameters"])
    with open(path, "w") as f:
        f.write("\n".join(lines))


def init_simulation(settings, params, init_type="distribution", particles=None):
    """init_simulation (init.rs:51-115): ``init_type`` is 'distribution' (the
    settings' initial distribution), 'file' (environment.init_file),
    'resume' (a snapshot in environment.init_file) or 'given' (``particles``)."""
    sim = Simulation(params)
    s = settings.simulation
    box = (s.box_size.x, s.box_size.y, s.box_size.z)
    if init_type == "given":
        sim.init(particles)
    elif init_type == "distribution":
        if s.init_distribution == "Isotropic":
            sim.init(create_isotropic(s.number_of_particles, box, s.seed, params.device))
        else:
            q = settings.parameters
            sim.init(create_homogeneous(s.number_of_particles, q.magnetic_reorientation / q.diffusion.rotational,
                                        box, s.seed, params.device))
    elif init_type in ("file", "resume"):
        fname = settings.environment.init_file
        if fname is None:
            sim.close()
            raise ValueError("No input file provided in the parameterfile.")
        if init_type == "file":
            sim.init(out.read_particles(fname))
        else:
            snap = out.read_snapshot(fname)
            rng = np.array(snap["rng_state"], dtype=np.uint64).reshape(-1, 2)
            sim.resume(Snapshot(snap["particles"], snap["compensation"], snap["timestep"], rng))
    else:
        sim.close()
        raise ValueError(f"unknown init_type {init_type!r}")
    return sim


def _particles(sim, o):
    return sim.get_particles_head(o.particles_head) if o.particles_head is not None else sim.get_particles()


def run_simulation(settings, sim, writer):
    """run_simulation (main.rs:139-289): the step loop with its outputs.
    Steps between outputs run as one ``ss_step`` call (replayed graphs)."""
    s = settings.simulation
    o = s.output_at_timestep
    box = (s.box_size.x, s.box_size.y, s.box_size.z)
    grid = (s.grid_size.x, s.grid_size.y, s.grid_size.z, s.grid_size.phi, s.grid_size.theta)
    if o.initial_condition:
        # the count grid only while a distribution is written (tiled path)
        sim.keep_distribution(True)
        writer.append(out.entry_value(sim.get_timestep(),
                                      distribution=out.distribution_value(sim.get_distribution(), box, grid),
                                      particles=_particles(sim, o) if o.particles is not None else None))
        sim.keep_distribution(False)
    start = sim.get_timestep() + 1
    last = s.number_of_timesteps + start
    every = [x for x in (o.distribution, o.flowfield, o.magneticfield, o.particles, o.snapshot) if x]
    t = start
    while t <= last:
        # next timestep with any output
        nxt = last
        for x in every:
            k = -(-t // x) * x
            if k <= last:
                nxt = min(nxt, k)
        if nxt > t:
            sim.step(nxt - t)
            t = nxt
        ts = t

        def due(x):
            return x is not None and ts % x == 0

        # the distribution of an output step is the one it samples: keep the
        # count grid from just before that step (it bins the particles then)
        if due(o.distribution):
            sim.keep_distribution(True)
        sim.step(1)
        dist = None
        if due(o.distribution):
            dist = out.distribution_value(sim.get_distribution(), box, grid)
            sim.keep_distribution(False)
        ff = sim.get_flow_field() if due(o.flowfield) else None
        mf = sim.get_magnetic_field() if due(o.magneticfield) else None
        parts = _particles(sim, o) if due(o.particles) else None
        if dist is not None or ff is not None or parts is not None:
            writer.append(out.entry_value(ts, distribution=dist, flowfield=ff, magneticfield=mf, particles=parts))
        if due(o.snapshot):
            snap = sim.get_snapshot()
            writer.write_snapshot(out.snapshot_value(snap.particles, snap.timestep, snap.compensation,
                                                     snap.rng_state))
        t += 1
    if o.final_snapshot:
        snap = sim.get_snapshot()
        writer.write_snapshot(out.snapshot_value(snap.particles, snap.timestep, snap.compensation, snap.rng_state))


def run(settings, out_dir, precision=FP32, init_type="distribution", noise="philox", noise_streams=None,
        **overrides):
    """A whole run like the reference binary's ``run`` (main.rs:66-132):
    output directory, the settings saved as TOML, the header, the loop.
    Returns the output file path."""
    flags = overrides.pop("flags", 0)
    if noise == "reference":
        if noise_streams is None:
            env = os.environ.get("RAYON_NUM_THREADS")
            if not env or not env.isdigit():
                raise ValueError("No environment variable 'RAYON_NUM_THREADS' set.")
            noise_streams = int(env)
        overrides.update(noise_mode=NOISE_REFERENCE, noise_streams=noise_streams)
    elif noise != "philox":
        raise ValueError(f"unknown noise {noise!r}")
    params = settings.to_params(precision=precision, flags=flags, **overrides)
    base = output_base(out_dir, settings.environment.prefix)
    save_toml(settings, base + ".toml")
    settings.set_version(VERSION)
    sim = init_simulation(settings, params, init_type)
    try:
        with out.OutputWriter(base, settings.environment.output_format) as w:
            w.write_metadata(settings)
            run_simulation(settings, sim, w)
            return w.path
    finally:
        sim.close()

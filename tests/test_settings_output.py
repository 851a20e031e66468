"""Parameter files (settings/mod.rs, si.rs), output files (worker.rs,
output.rs, tools/DataStreamer.py) and the run driver (main.rs:139-289).

The settings cases are the reference's own (settings/mod.rs:226-384) on its
fixture files (tests/golden/settings/, copied from the reference's test/
directory).  The byte formats of the third-party serialisers are checked
against hand-assembled encodings of small values."""
import lzma
import os
import struct

import msgpack
import numpy as np
import pytest

import stochasticsampling_amd as ss
from stochasticsampling_amd import driver, output as O, settings as S

SET = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "settings")


def test_read_settings_reference_case():
    """settings/mod.rs:233-361 (read_settings)."""
    st = S.read_parameter_file(os.path.join(SET, "parameter.toml"))
    st.set_version("version")
    d = S.read_parameter_file(os.path.join(SET, "parameter_no_defaults.toml"))
    assert d.environment.init_file is None
    assert st.environment.init_file == "foo/bar.cbor"
    assert d.environment.io_queue_size == S.DEFAULT_IO_QUEUE_SIZE == 1
    assert st.environment.io_queue_size == 50
    assert d.environment.output_format == S.DEFAULT_OUTPUT_FORMAT == "MsgPack"
    assert st.environment.output_format == "Bincode"
    assert st.environment.prefix == "foo" and st.environment.version == "version"
    q = st.parameters
    assert (q.diffusion.rotational, q.diffusion.translational) == (0.5, 1.0)
    assert (q.stress.active, q.stress.magnetic) == (1.0, 1.0)
    assert q.magnetic_dipole_dipole == 5.0 and d.parameters.magnetic_dipole_dipole == 0.0
    assert q.magnetic_reorientation == 1.0
    assert q.magnetic_drag == 123.4 and d.parameters.magnetic_drag == 0.0
    assert q.shape == 44.3
    assert d.parameters.hydro_screening == 0.0 and q.hydro_screening == 1.3
    assert d.parameters.volume_exclusion == 0.0 and q.volume_exclusion == 265.6
    s = st.simulation
    assert (s.box_size.x, s.box_size.y, s.box_size.z) == (1.0, 2.0, 3.0)
    assert (s.grid_size.x, s.grid_size.y, s.grid_size.z, s.grid_size.phi, s.grid_size.theta) == (11, 12, 13, 6, 7)
    assert s.init_distribution == "Homogeneous" and d.simulation.init_distribution == "Isotropic"
    assert (s.number_of_particles, s.number_of_timesteps, s.timestep, s.seed) == (100, 500, 0.1, 1)
    o, od = s.output_at_timestep, d.simulation.output_at_timestep
    assert o.distribution == 12 and od.distribution is None
    assert od.final_snapshot is True and o.final_snapshot is False
    assert o.flowfield == 42 and od.flowfield is None
    assert o.magneticfield == 41 and od.magneticfield is None
    assert o.particles == 100 and od.particles is None
    assert o.particles_head == 10 and od.particles_head is None
    assert od.initial_condition is True and o.initial_condition is False
    assert o.snapshot == 666 and od.snapshot is None


def test_settings_unused_keys_are_rejected():
    """settings/mod.rs:363-367 (the fixture itself does not even parse as
    TOML 1.0: `x = 1.`); an unknown key at any level is an error."""
    with pytest.raises(S.SettingsError):
        S.read_parameter_file(os.path.join(SET, "parameter_unused.toml"))
    text = open(os.path.join(SET, "parameter_no_defaults.toml")).read()
    for section, extra in [("[simulation]", "non_existing_key = 10"), ("[parameters.diffusion]", "foo = 1.0"),
                           ("[environment]", 'version = "x"')]:
        with pytest.raises(S.SettingsError, match="unknown field"):
            S.read_parameter_file(text=text.replace(section, section + "\n    " + extra, 1))
    with pytest.raises(S.SettingsError, match="missing field"):
        S.read_parameter_file(text=text.replace("    seed = 1\n", ""))
    with pytest.raises(S.SettingsError, match="unknown variant"):
        S.read_parameter_file(text=text.replace("[environment]", '[environment]\n    output_format = "Json"'))


def test_check_settings():
    text = open(os.path.join(SET, "parameter_no_defaults.toml")).read()
    with pytest.raises(S.SettingsError, match="Box size"):
        S.read_parameter_file(text=text.replace("x = 1.0", "x = 0.0", 1))
    with pytest.raises(S.SettingsError, match="particles_head"):
        S.read_parameter_file(text=text.replace("[simulation.output_at_timestep]",
                                                "[simulation.output_at_timestep]\n        particles_head = 101"))


def test_si_settings():
    # the fixture lacks the required hydro_screening and volume_exclusion
    with pytest.raises(S.SettingsError, match="hydro_screening: missing field"):
        S.read_parameter_file_si(os.path.join(SET, "parameter_si.toml"))
    # number density 1 and speed 1 make the length and time scales 1; at
    # temperature 0 only the active rotational diffusion 1/(2 tau) remains
    text = open(os.path.join(SET, "parameter_si.toml")).read()
    r = (3.0 * 0.5 / (4.0 * np.pi)) ** (1.0 / 3.0)  # volume fraction 0.5 -> n = 1
    text = (text.replace("viscocity = 1e-3", "viscocity = 2.0\n    hydro_screening = 0.25\n    volume_exclusion = 0.5")
            .replace("temperature = 300", "temperature = 0.0").replace("volume_fraction = 1e-2", "volume_fraction = 0.5")
            .replace("radius = 1e-6", f"radius = {r!r}").replace("self_propulsion_speed = 1e-4", "self_propulsion_speed = 1.0"))
    s = S.read_parameter_file_si(text=text)
    assert abs(s.simulation.box_size.x - 1e-4) < 1e-16 and abs(s.simulation.timestep - 0.1) < 1e-15
    q = s.parameters
    assert q.diffusion.translational == 0.0
    assert abs(q.diffusion.rotational - 1.0 / 1.5) < 1e-15  # 1 / (2 * 0.75)
    assert abs(q.stress.active - 1.88496e-18 / 2.0) < 1e-30
    assert abs(q.stress.magnetic - 1e-15 * 1e-3 / 2.0) < 1e-30
    fr = 8.0 * np.pi * 2.0 * r ** 3
    assert abs(q.magnetic_reorientation - 1e-15 * 1e-3 / fr) < 1e-28  # alignment * D_r, D_r = rotdiff
    assert (q.hydro_screening, q.volume_exclusion) == (0.25, 0.5)


def test_to_params_and_toml_round_trip(tmp_path):
    st = S.read_parameter_file(os.path.join(SET, "parameter.toml"))
    p = st.to_params(precision=ss.FP64)
    assert (p.box_size.x, p.grid_size.theta, p.number_of_particles, p.seed) == (1.0, 7, 100, 1)
    assert (p.magnetic_drag, p.shape, p.hydro_screening, p.volume_exclusion) == (123.4, 44.3, 1.3, 265.6)
    assert (p.stress_active, p.stress_magnetic, p.magnetic_dipole_dipole) == (1.0, 1.0, 5.0)
    assert (p.diffusion_translational, p.diffusion_rotational, p.timestep) == (1.0, 0.5, 0.1)
    path = str(tmp_path / "s.toml")
    driver.save_toml(st, path)
    assert S.read_parameter_file(path) == st


def test_bincode_msgpack_cbor_byte_layout():
    e = O.entry_value(7)
    # bincode: four None tags, then the u64 timestep
    assert O.encode("Bincode", e) == b"\x00\x00\x00\x00" + struct.pack("<Q", 7)
    # msgpack: a map of 5 named fields, nil values, fixint 7
    assert O.encode("MsgPack", e) == (b"\x85\xacdistribution\xc0\xa9flowfield\xc0\xadmagneticfield\xc0"
                                      b"\xa9particles\xc0\xa8timestep\x07")
    assert O.encode("CBOR", e) == (b"\xa5\x6cdistribution\xf6\x69flowfield\xf6\x6dmagneticfield\xf6"
                                   b"\x69particles\xf6\x68timestep\x07")
    # an ndarray: {v: 1, dim: (2,), data: [0.5, -1.0]}
    a = O.Arr(np.array([0.5, -1.0]))
    assert O.encode("Bincode", {"a": a}) == (b"\x01" + struct.pack("<Q", 2) + struct.pack("<Q", 2)
                                             + struct.pack("<2d", 0.5, -1.0))
    assert O.encode("MsgPack", {"a": a}) == (b"\x81\xa1a\x83\xa1v\x01\xa3dim\x91\x02\xa4data\x92"
                                             b"\xcb" + struct.pack(">d", 0.5) + b"\xcb" + struct.pack(">d", -1.0))
    # Option<String>, unit enum, bool
    v = {"s": O.Opt("ab"), "e": O.Enum("Bincode", S.OUTPUT_FORMATS), "b": True}
    assert O.encode("Bincode", v) == b"\x01" + struct.pack("<Q", 2) + b"ab" + struct.pack("<I", 1) + b"\x01"
    assert O.encode("MsgPack", v) == b"\x83\xa1s\xa2ab\xa1e\x92\x01\x90\xa1b\xc3"
    assert O.encode("CBOR", v) == b"\xa3\x61s\x62ab\x61e\x67Bincode\x61b\xf5"


@pytest.mark.parametrize("fmt", ["MsgPack", "CBOR", "Bincode"])
def test_output_files_round_trip(tmp_path, fmt):
    st = S.read_parameter_file(os.path.join(SET, "parameter.toml"))
    st.set_version("0.1.0")
    rng = np.random.default_rng(1)
    base = str(tmp_path / "run")
    dist = rng.random((3, 2, 2, 2, 2))
    ff = rng.random((3, 3, 2, 2))
    parts = rng.random((10, 5))
    with O.OutputWriter(base, fmt) as w:
        w.write_metadata(st)
        w.append(O.entry_value(0, distribution=O.distribution_value(dist, (1.0, 2.0, 3.0), (3, 2, 2, 2, 2))))
        w.append(O.entry_value(5, flowfield=ff, particles=parts))
        snap = w.write_snapshot(O.snapshot_value(parts, 5, compensation=rng.random((10, 3)) * 1e-8))
    r = O.OutputReader(w.path)
    assert len(r) == 2
    meta = r.metadata()
    assert meta["simulation"]["grid_size"]["x"] == 11 and meta["parameters"]["shape"] == 44.3
    assert meta["environment"]["version"] == "0.1.0"
    e0, e1 = r[0], r[1]
    assert e0["timestep"] == 0 and e1["timestep"] == 5
    np.testing.assert_array_equal(e0["distribution"]["dist"], dist)
    assert e0["distribution"]["grid_width"]["theta"] == np.pi / 2
    assert e0["flowfield"] is None and e0["particles"] is None
    np.testing.assert_array_equal(e1["flowfield"], ff)
    np.testing.assert_array_equal(e1["particles"], parts)
    s = O.read_snapshot(snap)
    assert s["timestep"] == 5 and s["rng_state"] == []
    np.testing.assert_array_equal(s["particles"], parts)
    # index: u64 offsets, the first right after the uncompressed header
    idx = np.fromfile(base + ".index", dtype="<u8")
    raw = open(w.path, "rb").read()
    assert raw[idx[1]:idx[1] + 6] == b"\xfd7zXZ\x00"  # an xz stream per blob
    lzma.decompress(raw[idx[0]:idx[1]])


def test_msgpack_output_reads_like_datastreamer(tmp_path):
    """The access pattern of tools/DataStreamer.py (get_metadata, __getitem__,
    data_to_dist, data_to_flowfield) on a MsgPack file."""
    st = S.read_parameter_file(os.path.join(SET, "parameter_no_defaults.toml"))
    base = str(tmp_path / "run")
    dist = np.arange(11 * 12 * 13 * 6 * 7, dtype=float).reshape(11, 12, 13, 6, 7)
    with O.OutputWriter(base, "MsgPack") as w:
        w.write_metadata(st)
        w.append(O.entry_value(3, distribution=O.distribution_value(dist, (1.0, 2.0, 3.0), (11, 12, 13, 6, 7)),
                               flowfield=np.ones((3, 11, 12, 13))))
    index = np.fromfile(base + ".index", dtype=np.uint64)
    with open(w.path, "rb") as f:
        meta = msgpack.unpackb(f.read(int(index[0])), raw=False)
        f.seek(int(index[0]))
        data = msgpack.unpackb(lzma.decompress(f.read()), raw=False)
    assert meta["simulation"]["box_size"] == {"x": 1.0, "y": 2.0, "z": 3.0}
    assert meta["parameters"]["diffusion"]["rotational"] == 0.5
    d = np.array(data["distribution"]["dist"]["data"]).reshape(data["distribution"]["dist"]["dim"])
    np.testing.assert_array_equal(d, dist)
    ff = np.array(data["flowfield"]["data"]).reshape(data["flowfield"]["dim"])
    assert ff.shape == (3, 11, 12, 13)


def test_particle_files_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    for fmt, ext in O.EXTENSIONS.items():
        path = str(tmp_path / f"init.{ext}")
        p = rng.random((50, 5))
        O.write_particles(path, p)
        np.testing.assert_array_equal(O.read_particles(path), p)


@pytest.mark.parametrize("fmt", ["MsgPack", "CBOR", "Bincode"])
def test_snapshot_rng_state_round_trip(tmp_path, fmt):
    """rng_state rows serialise like rand_pcg's Lcg64Xsh32 {state, increment}
    (full u64 range) and read back as (state, increment) pairs."""
    rng = np.random.default_rng(2)
    parts = rng.random((4, 5))
    st = [(2**64 - 1, 2**63 + 1), (12345, 6789)]
    with O.OutputWriter(str(tmp_path / "r"), fmt) as w:
        snap = w.write_snapshot(O.snapshot_value(parts, 9, rng_state=st))
    s = O.read_snapshot(snap)
    assert s["rng_state"] == st and s["timestep"] == 9
    if fmt == "Bincode":
        raw = lzma.decompress(open(snap, "rb").read())
        # Vec<Particle>: u64 length + 4 x 5 f64, then Vec<Pcg32>: u64 length + (state, increment) pairs
        off = 8 + 4 * 5 * 8
        assert struct.unpack_from("<5Q", raw, off) == (2, 2**64 - 1, 2**63 + 1, 12345, 6789)
    elif fmt == "MsgPack":
        assert O.msgpack_decode(lzma.decompress(open(snap, "rb").read()))["rng_state"][0] == {
            "state": 2**64 - 1, "increment": 2**63 + 1}


@pytest.mark.gpu
def test_driver_initial_conditions_are_the_reference_streams(oracle):
    p = driver.create_isotropic(200000, (1.0, 2.0, 3.0), 5)
    np.testing.assert_array_equal(p[:, :4], oracle.create_isotropic(200000, (1.0, 2.0, 3.0), 5)[:, :4])
    assert abs(np.cos(p[:, 4]).mean()) < 0.01 and p[:, 1].max() < 2.0
    h = driver.create_homogeneous(200000, 2.0, (1.0, 1.0, 1.0), 6)
    ny = np.sin(h[:, 4]) * np.sin(h[:, 3])
    # <n . y> for p(c) ~ exp(kappa c): coth(kappa) - 1/kappa
    assert abs(ny.mean() - (1.0 / np.tanh(2.0) - 0.5)) < 0.01
    assert abs((np.sin(h[:, 4]) * np.cos(h[:, 3])).mean()) < 0.01 and abs(np.cos(h[:, 4]).mean()) < 0.01
    with pytest.raises(ValueError, match="Alignment of zero"):
        driver.create_homogeneous(10, 0.0, (1.0, 1.0, 1.0), 6)


@pytest.mark.gpu
def test_driver_run_and_resume(tmp_path):
    """A small run through the driver: entries on schedule (one step more
    than number_of_timesteps, main.rs:176-182), fields and particles equal
    to the simulation's getters, and resuming the final snapshot continues
    exactly like an uninterrupted run."""
    text = open(os.path.join(SET, "parameter_no_defaults.toml")).read()
    text = (text.replace("number_of_particles = 100", "number_of_particles = 3000")
            .replace("number_of_timesteps = 500", "number_of_timesteps = 6")
            .replace("timestep = 0.1", "timestep = 0.001")
            .replace("x = 11", "x = 8").replace("y = 12", "y = 8").replace("z = 13", "z = 8")
            .replace("phi = 6", "phi = 4").replace("theta = 7", "theta = 4")
            .replace("x = 1.0", "x = 4.0").replace("y = 2.0", "y = 4.0").replace("z = 3.0", "z = 4.0")
            .replace("[simulation.output_at_timestep]",
                     "[simulation.output_at_timestep]\n        flowfield = 2\n        particles = 3\n"
                     "        distribution = 7\n        snapshot = 4"))
    st = S.read_parameter_file(text=text)
    path = driver.run(st, str(tmp_path), precision=ss.FP32)
    r = O.OutputReader(path)
    ts = [r[i]["timestep"] for i in range(len(r))]
    assert ts == [0, 2, 3, 4, 6, 7]
    assert r[0]["distribution"] is not None and r[-1]["distribution"] is not None
    base = path[: -len(".msgpack-lzma")]
    snaps = sorted(f for f in os.listdir(os.path.dirname(path)) if ".snapshot." in f)
    assert len(snaps) == 2  # step 4 and the final one
    final = O.read_snapshot(os.path.join(os.path.dirname(path), f"{os.path.basename(base)}.snapshot.2.msgpack-lzma"))
    assert final["timestep"] == 7
    # uninterrupted reference run with the same parameters and initial condition
    p = st.to_params(precision=ss.FP32, flags=ss.FLAG_KEEP_DISTRIBUTION)
    sim = ss.Simulation(p)
    box = (4.0, 4.0, 4.0)
    sim.init(driver.create_isotropic(3000, box, 1))
    # the driver keeps the count grid only around its distribution outputs
    np.testing.assert_array_equal(sim.get_distribution(), r[0]["distribution"]["dist"])
    sim.step(6)
    assert r[4]["timestep"] == 6
    np.testing.assert_array_equal(sim.get_particles(), r[4]["particles"])
    np.testing.assert_array_equal(sim.get_flow_field(), r[4]["flowfield"])
    sim.step(1)
    np.testing.assert_array_equal(sim.get_distribution(), r[5]["distribution"]["dist"])
    ref = sim.get_snapshot()  # FP32 positions and their low-order parts
    np.testing.assert_array_equal(ref.particles, final["particles"])
    np.testing.assert_array_equal(ref.compensation, final["compensation"])
    sim.step(3)
    # resume the step-4 snapshot and run 6 more steps
    mid = O.read_snapshot(os.path.join(os.path.dirname(path), f"{os.path.basename(base)}.snapshot.1.msgpack-lzma"))
    assert mid["timestep"] == 4
    res = ss.Simulation(p)
    res.resume(ss.Snapshot(mid["particles"], mid["compensation"], mid["timestep"]))
    res.step(6)
    np.testing.assert_array_equal(res.get_particles(), sim.get_particles())


@pytest.mark.gpu
def test_driver_reference_noise_snapshot_resume(tmp_path):
    """Reference noise through the driver: the snapshot files carry the
    generators, and a run resumed from the step-4 snapshot file continues the
    uninterrupted run exactly (init_simulation 'resume', init.rs:51-115)."""
    text = open(os.path.join(SET, "parameter_no_defaults.toml")).read()
    text = (text.replace("number_of_particles = 100", "number_of_particles = 2000")
            .replace("number_of_timesteps = 500", "number_of_timesteps = 6")
            .replace("timestep = 0.1", "timestep = 0.001")
            .replace("x = 11", "x = 8").replace("y = 12", "y = 8").replace("z = 13", "z = 8")
            .replace("phi = 6", "phi = 4").replace("theta = 7", "theta = 4")
            .replace("x = 1.0", "x = 4.0").replace("y = 2.0", "y = 4.0").replace("z = 3.0", "z = 4.0")
            .replace("[simulation.output_at_timestep]",
                     "[simulation.output_at_timestep]\n        particles = 7\n        snapshot = 4"))
    st = S.read_parameter_file(text=text)
    path = driver.run(st, str(tmp_path), precision=ss.FP64, noise="reference", noise_streams=3)
    r = O.OutputReader(path)
    last = [r[i] for i in range(len(r)) if r[i]["timestep"] == 7][0]["particles"]
    d = os.path.dirname(path)
    base = os.path.basename(path)[: -len(".msgpack-lzma")]
    mid = os.path.join(d, f"{base}.snapshot.1.msgpack-lzma")
    snap = O.read_snapshot(mid)
    assert snap["timestep"] == 4 and len(snap["rng_state"]) == 3
    st.environment.init_file = mid
    p = st.to_params(precision=ss.FP64, noise_mode=ss.NOISE_REFERENCE, noise_streams=3)
    sim = driver.init_simulation(st, p, "resume")
    sim.step(3)
    np.testing.assert_array_equal(sim.get_particles(), last)
    with pytest.raises(ValueError, match="RAYON_NUM_THREADS"):
        os.environ.pop("RAYON_NUM_THREADS", None)
        driver.run(st, str(tmp_path), precision=ss.FP64, noise="reference")

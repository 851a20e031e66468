"""Output files of the reference: a header with the settings, then one
LZMA-compressed blob per output entry, an ``.index`` file of the blobs' byte
offsets, and compressed snapshot files (src/bin/output/worker.rs:101-248,
README.md:38-58; read by tools/DataStreamer.py).

Serialisation follows serde's data model as the reference's crates encode it
(versions from Cargo.lock; the crates are not vendored, so their published
formats are restated here):

* MsgPack, rmp-serde 0.13.7 ``write_named``: structs as maps keyed by field
  name, sequences and tuples as arrays, ``Option`` as nil or the value,
  integers in their shortest form, f64 as float 64, unit enum variants as
  ``[variant index, []]``.
* CBOR, serde_cbor 0.10.1: structs as maps keyed by field name, unit enum
  variants as their name; floats are written as 64-bit floats (serde_cbor
  may pick a shorter float when the value allows; readers see the same
  number).
* Bincode 1.1.4 (default options): little endian, fixed-width integers,
  ``usize`` as u64, sequence and string lengths as u64, ``Option`` as a u8
  tag, enum variants as u32 indices, structs and tuples as their fields in
  order.
* ndarray 0.12.1 arrays: the struct ``{v: 1u8, dim: (usize, ...), data:
  [elements in logical order]}``.
* ``Particle`` as the tuple ``[x, y, z, phi, theta]`` (particle.rs:363-395).
* LZMA: one xz stream per blob (lzma crate, preset 1, CRC64 check).

Values are built as trees: dicts are structs (field order kept), lists are
sequences, ``Tup`` tuples, ``Opt`` options, ``Enum`` unit variants, ``U8`` a
u8, Python ints are usize, floats f64, ``Arr`` an ndarray.
"""
import io
import lzma
import os
import struct

import numpy as np

try:
    import msgpack
except ImportError:  # pragma: no cover
    msgpack = None

EXTENSIONS = {"MsgPack": "msgpack-lzma", "CBOR": "cbor-lzma", "Bincode": "bincode-lzma"}
FORMAT_OF_EXTENSION = {v: k for k, v in EXTENSIONS.items()}


class Tup(tuple):
    """A serde tuple (fixed length, no length prefix in bincode)."""


class Opt:
    def __init__(self, value):
        self.value = value


class Enum:
    def __init__(self, name, variants):
        self.name = name
        self.index = list(variants).index(name)


class U8(int):
    pass


class Arr:
    """An ndarray 0.12 array (any float dtype is written as f64)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a, dtype=np.float64)


def _lift(v):
    """Settings.to_serde() markers -> value tree."""
    if isinstance(v, dict):
        return {k: _lift(x) for k, x in v.items()}
    if isinstance(v, tuple) and v and v[0] == "option":
        return Opt(None if v[1] is None else _lift(v[1]))
    if isinstance(v, tuple) and v and v[0] == "enum":
        return Enum(v[1], v[2])
    return v


# --------------------------------------------------------------------------
# MsgPack
# --------------------------------------------------------------------------
def _mp_plain(v):
    if isinstance(v, dict):
        return {k: _mp_plain(x) for k, x in v.items()}
    if isinstance(v, Opt):
        return None if v.value is None else _mp_plain(v.value)
    if isinstance(v, Enum):
        return [v.index, []]
    if isinstance(v, Arr):
        return {"v": 1, "dim": list(v.a.shape), "data": v.a.ravel().tolist()}
    if isinstance(v, (list, tuple)):
        return [_mp_plain(x) for x in v]
    if isinstance(v, np.ndarray):
        return v.tolist()
    return v


def msgpack_encode(v):
    if msgpack is None:  # pragma: no cover
        raise RuntimeError("msgpack is not importable")
    return msgpack.packb(_mp_plain(v), use_bin_type=True, use_single_float=False)


def msgpack_decode(b):
    return msgpack.unpackb(b, raw=False, strict_map_key=False)


# --------------------------------------------------------------------------
# CBOR (RFC 7049 major types 0-7, definite lengths)
# --------------------------------------------------------------------------
def _cb_head(out, major, n):
    if n < 24:
        out.append(bytes([(major << 5) | n]))
    elif n < 1 << 8:
        out.append(bytes([(major << 5) | 24, n]))
    elif n < 1 << 16:
        out.append(bytes([(major << 5) | 25]) + struct.pack(">H", n))
    elif n < 1 << 32:
        out.append(bytes([(major << 5) | 26]) + struct.pack(">I", n))
    else:
        out.append(bytes([(major << 5) | 27]) + struct.pack(">Q", n))


def _cb(out, v):
    if isinstance(v, dict):
        _cb_head(out, 5, len(v))
        for k, x in v.items():
            _cb(out, k)
            _cb(out, x)
    elif isinstance(v, Opt):
        _cb(out, v.value)
    elif isinstance(v, Enum):
        _cb(out, v.name)
    elif isinstance(v, Arr):
        _cb(out, {"v": U8(1), "dim": Tup(v.a.shape), "data": v.a.ravel()})
    elif v is None:
        out.append(b"\xf6")
    elif isinstance(v, bool):
        out.append(b"\xf5" if v else b"\xf4")
    elif isinstance(v, (int, np.integer)):
        v = int(v)
        if v >= 0:
            _cb_head(out, 0, v)
        else:
            _cb_head(out, 1, -1 - v)
    elif isinstance(v, (float, np.floating)):
        out.append(b"\xfb" + struct.pack(">d", float(v)))
    elif isinstance(v, str):
        b = v.encode()
        _cb_head(out, 3, len(b))
        out.append(b)
    elif isinstance(v, np.ndarray):
        if v.ndim == 2:  # a sequence of fixed-size tuples (particles, compensation)
            _cb_head(out, 4, v.shape[0])
            row = bytes([0x80 | v.shape[1]]) if v.shape[1] < 24 else None
            for r in v:
                if row is None:
                    _cb_head(out, 4, v.shape[1])
                else:
                    out.append(row)
                out.append(b"".join(b"\xfb" + struct.pack(">d", float(x)) for x in r))
        else:
            _cb_head(out, 4, v.size)
            flat = v.ravel()
            if flat.dtype.kind == "f":
                out.append(b"".join(b"\xfb" + struct.pack(">d", float(x)) for x in flat))
            else:
                for x in flat:
                    _cb(out, x)
    elif isinstance(v, (list, tuple)):
        _cb_head(out, 4, len(v))
        for x in v:
            _cb(out, x)
    else:
        raise TypeError(f"cannot encode {type(v)}")


def cbor_encode(v):
    out = []
    _cb(out, v)
    return b"".join(out)


def cbor_decode(b):
    v, pos = _cbd(memoryview(b), 0)
    return v


def _cbd_len(b, pos, info):
    if info < 24:
        return info, pos
    if info == 24:
        return b[pos], pos + 1
    if info == 25:
        return struct.unpack_from(">H", b, pos)[0], pos + 2
    if info == 26:
        return struct.unpack_from(">I", b, pos)[0], pos + 4
    if info == 27:
        return struct.unpack_from(">Q", b, pos)[0], pos + 8
    raise ValueError("indefinite lengths are not used by serde_cbor")


def _half(h):
    s, e, m = h >> 15, (h >> 10) & 0x1F, h & 0x3FF
    if e == 0:
        v = m * 2.0 ** -24
    elif e == 31:
        v = float("inf") if m == 0 else float("nan")
    else:
        v = (1 + m / 1024.0) * 2.0 ** (e - 15)
    return -v if s else v


def _cbd(b, pos):
    ib = b[pos]
    pos += 1
    major, info = ib >> 5, ib & 0x1F
    if major == 7:
        if info == 20:
            return False, pos
        if info == 21:
            return True, pos
        if info == 22:
            return None, pos
        if info == 25:
            return _half(struct.unpack_from(">H", b, pos)[0]), pos + 2
        if info == 26:
            return struct.unpack_from(">f", b, pos)[0], pos + 4
        if info == 27:
            return struct.unpack_from(">d", b, pos)[0], pos + 8
        raise ValueError(f"unsupported CBOR simple value {info}")
    n, pos = _cbd_len(b, pos, info)
    if major == 0:
        return n, pos
    if major == 1:
        return -1 - n, pos
    if major == 2:
        return bytes(b[pos : pos + n]), pos + n
    if major == 3:
        return bytes(b[pos : pos + n]).decode(), pos + n
    if major == 4:
        out = []
        for _ in range(n):
            x, pos = _cbd(b, pos)
            out.append(x)
        return out, pos
    if major == 5:
        out = {}
        for _ in range(n):
            k, pos = _cbd(b, pos)
            x, pos = _cbd(b, pos)
            out[k] = x
        return out, pos
    if major == 6:  # tag: skip it
        return _cbd(b, pos)
    raise ValueError(f"bad CBOR major type {major}")


# --------------------------------------------------------------------------
# Bincode
# --------------------------------------------------------------------------
def _bc(out, v):
    if isinstance(v, dict):
        for x in v.values():
            _bc(out, x)
    elif isinstance(v, Opt):
        if v.value is None:
            out.append(b"\x00")
        else:
            out.append(b"\x01")
            _bc(out, v.value)
    elif isinstance(v, Enum):
        out.append(struct.pack("<I", v.index))
    elif isinstance(v, Arr):
        out.append(b"\x01")
        out.append(struct.pack(f"<{v.a.ndim}Q", *v.a.shape))
        out.append(struct.pack("<Q", v.a.size))
        out.append(v.a.astype("<f8").tobytes())
    elif isinstance(v, U8):
        out.append(struct.pack("<B", int(v)))
    elif isinstance(v, bool):
        out.append(b"\x01" if v else b"\x00")
    elif isinstance(v, (int, np.integer)):
        out.append(struct.pack("<Q", int(v)))
    elif isinstance(v, (float, np.floating)):
        out.append(struct.pack("<d", float(v)))
    elif isinstance(v, str):
        b = v.encode()
        out.append(struct.pack("<Q", len(b)))
        out.append(b)
    elif isinstance(v, Tup):
        for x in v:
            _bc(out, x)
    elif isinstance(v, np.ndarray):  # Vec<[f64; k]> or Vec<f64>
        out.append(struct.pack("<Q", v.shape[0]))
        out.append(np.ascontiguousarray(v, dtype="<f8").tobytes())
    elif isinstance(v, (list, tuple)):
        out.append(struct.pack("<Q", len(v)))
        for x in v:
            _bc(out, x)
    else:
        raise TypeError(f"cannot encode {type(v)}")


def bincode_encode(v):
    out = []
    _bc(out, v)
    return b"".join(out)


class _BR:
    def __init__(self, b):
        self.b = memoryview(b)
        self.pos = 0

    def take(self, fmt):
        r = struct.unpack_from(fmt, self.b, self.pos)
        self.pos += struct.calcsize(fmt)
        return r

    def u64(self):
        return self.take("<Q")[0]

    def f64(self):
        return self.take("<d")[0]

    def f64s(self, n):
        a = np.frombuffer(self.b, dtype="<f8", count=n, offset=self.pos).astype(np.float64)
        self.pos += 8 * n
        return a

    def opt(self, f):
        tag = self.take("<B")[0]
        return f() if tag == 1 else None

    def string(self):
        n = self.u64()
        s = bytes(self.b[self.pos : self.pos + n]).decode()
        self.pos += n
        return s

    def array(self, ndim):
        assert self.take("<B")[0] == 1
        dim = self.take(f"<{ndim}Q")
        n = self.u64()
        return self.f64s(n).reshape(dim)

    def particles(self):
        n = self.u64()
        return self.f64s(5 * n).reshape(n, 5)


def _bd_distribution(r):
    dist = r.array(5)
    gw = dict(zip(("x", "y", "z", "phi", "theta"), (r.f64() for _ in range(5))))
    bs = dict(zip(("x", "y", "z"), (r.f64() for _ in range(3))))
    gs = dict(zip(("x", "y", "z", "phi", "theta"), (r.u64() for _ in range(5))))
    return {"dist": dist, "grid_width": gw, "box_size": bs, "grid_size": gs}


def bincode_decode_entry(b):
    """OutputEntry (src/output.rs:11-17) from bincode."""
    r = _BR(b)
    return {
        "distribution": r.opt(lambda: _bd_distribution(r)),
        "flowfield": r.opt(lambda: r.array(4)),
        "magneticfield": r.opt(lambda: r.array(4)),
        "particles": r.opt(r.particles),
        "timestep": r.u64(),
    }


def bincode_decode_snapshot(b):
    """Snapshot (src/bin/simulation/mod.rs:80-86) from bincode; the trailing
    compensation field is this library's extension."""
    r = _BR(b)
    parts = r.particles()
    nrng = r.u64()
    rng = [tuple(r.take("<QQ")) for _ in range(nrng)]
    ts = r.u64()
    comp = None
    if r.pos < len(r.b):
        n = r.u64()
        comp = r.f64s(3 * n).reshape(n, 3)
    return {"particles": parts, "rng_state": rng, "timestep": ts, "compensation": comp}


def bincode_decode_settings(b):
    """Settings (settings/mod.rs:24-30) from bincode."""
    from .settings import INIT_DISTRIBUTIONS, OUTPUT_FORMATS

    r = _BR(b)
    sim = {
        "number_of_particles": r.u64(),
        "number_of_timesteps": r.u64(),
        "timestep": r.f64(),
        "init_distribution": INIT_DISTRIBUTIONS[r.take("<I")[0]],
        "seed": r.u64(),
    }
    o = {"initial_condition": bool(r.take("<B")[0]), "final_snapshot": bool(r.take("<B")[0])}
    for k in ("distribution", "flowfield", "magneticfield", "particles_head", "particles", "snapshot"):
        o[k] = r.opt(r.u64)
    sim["output_at_timestep"] = o
    sim["box_size"] = dict(zip("xyz", (r.f64() for _ in range(3))))
    sim["grid_size"] = dict(zip(("x", "y", "z", "phi", "theta"), (r.u64() for _ in range(5))))
    par = {k: r.f64() for k in ("magnetic_drag", "shape", "hydro_screening", "volume_exclusion",
                                 "magnetic_reorientation")}
    par["diffusion"] = {"translational": r.f64(), "rotational": r.f64()}
    par["stress"] = {"active": r.f64(), "magnetic": r.f64()}
    par["magnetic_dipole"] = {"magnetic_dipole_dipole": r.f64()}
    env = {"prefix": r.string(), "version": r.string(), "init_file": r.opt(r.string), "io_queue_size": r.u64(),
           "output_format": OUTPUT_FORMATS[r.take("<I")[0]]}
    return {"simulation": sim, "parameters": par, "environment": env}, r.pos


# --------------------------------------------------------------------------
# Values of the reference's output types
# --------------------------------------------------------------------------
def distribution_value(dist, box, grid):
    """Distribution (distribution/mod.rs:27-36) of a normalised 5-D array."""
    gw = {"x": box[0] / grid[0], "y": box[1] / grid[1], "z": box[2] / grid[2], "phi": 2 * np.pi / grid[3],
          "theta": np.pi / grid[4]}
    return {
        "dist": Arr(dist),
        "grid_width": gw,
        "box_size": {"x": float(box[0]), "y": float(box[1]), "z": float(box[2])},
        "grid_size": dict(zip(("x", "y", "z", "phi", "theta"), (int(g) for g in grid))),
    }


def entry_value(timestep, distribution=None, flowfield=None, magneticfield=None, particles=None):
    """OutputEntry (src/output.rs:11-17)."""
    return {
        "distribution": Opt(distribution),
        "flowfield": Opt(None if flowfield is None else Arr(flowfield)),
        "magneticfield": Opt(None if magneticfield is None else Arr(magneticfield)),
        "particles": Opt(None if particles is None else np.ascontiguousarray(particles, dtype=np.float64)),
        "timestep": int(timestep),
    }


def snapshot_value(particles, timestep, compensation=None, rng_state=None):
    """Snapshot (src/bin/simulation/mod.rs:80-86).  ``rng_state``: the
    reference-noise generators as (state, increment) rows, each serialised
    like rand_pcg's Lcg64Xsh32 {state, increment}; empty for the
    counter-based Philox noise, which has no generator state.  The FP32
    path's low-order position parts follow as an extra ``compensation``
    field."""
    rng = [] if rng_state is None else [{"state": int(s), "increment": int(i)} for s, i in rng_state]
    v = {"particles": np.ascontiguousarray(particles, dtype=np.float64), "rng_state": rng, "timestep": int(timestep)}
    if compensation is not None:
        v["compensation"] = np.ascontiguousarray(compensation, dtype=np.float64)
    return v


def _plain_particles(v):
    """msgpack/CBOR trees -> numpy particles where the value is a list of
    5-element rows."""
    return np.asarray(v, dtype=np.float64).reshape(-1, 5) if v is not None else None


def encode(fmt, v):
    if fmt == "MsgPack":
        return msgpack_encode(_np_to_lists(v))
    if fmt == "CBOR":
        return cbor_encode(v)
    if fmt == "Bincode":
        return bincode_encode(v)
    raise ValueError(f"unknown output format {fmt}")


def _np_to_lists(v):
    if isinstance(v, dict):
        return {k: _np_to_lists(x) for k, x in v.items()}
    if isinstance(v, Opt):
        return Opt(_np_to_lists(v.value))
    if isinstance(v, np.ndarray):
        return v.tolist()
    return v


def compress(b):
    return lzma.compress(b, format=lzma.FORMAT_XZ, check=lzma.CHECK_CRC64, preset=1)


# --------------------------------------------------------------------------
# Files
# --------------------------------------------------------------------------
class OutputWriter:
    """The IO worker's files for one run (worker.rs:101-248): ``base``
    + ``.<format>-lzma`` (header, then blobs), ``.index`` (u64 offsets of the
    blobs) and ``.snapshot.<k>.<format>-lzma``.  Synchronous: the caller's
    thread writes (the GPU step runs asynchronously anyway)."""

    def __init__(self, base, fmt="MsgPack"):
        if fmt not in EXTENSIONS:
            raise ValueError(f"unknown output format {fmt}")
        self.base = base
        self.fmt = fmt
        self.path = f"{base}.{EXTENSIONS[fmt]}"
        self.index_path = f"{base}.index"
        self.snapshots = 0
        self._f = open(self.path, "wb")
        self._idx = open(self.index_path, "wb")

    def write_metadata(self, settings):
        """The settings as the file's first, uncompressed object."""
        self._f.write(encode(self.fmt, _lift(settings.to_serde())))

    def append(self, entry):
        self._idx.write(struct.pack("<Q", self._f.tell()))
        self._f.write(compress(encode(self.fmt, entry)))

    def write_snapshot(self, snapshot):
        self.snapshots += 1
        path = f"{self.base}.snapshot.{self.snapshots}.{EXTENSIONS[self.fmt]}"
        with open(path, "wb") as f:
            f.write(compress(encode(self.fmt, snapshot)))
        return path

    def close(self):
        if self._f:
            self._f.close()
            self._idx.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class OutputReader:
    """Blobs of an output file by index, like tools/DataStreamer.py; entries
    come back as dicts with numpy arrays."""

    def __init__(self, path, index_path=None):
        base, ext = os.path.splitext(path)
        self.fmt = FORMAT_OF_EXTENSION[ext[1:]]
        self.path = path
        self.index = np.fromfile(index_path or base + ".index", dtype="<u8")
        size = os.path.getsize(path)
        self.sizes = np.diff(np.append(self.index, size)).astype(np.int64)

    def __len__(self):
        return len(self.index)

    def metadata(self):
        with open(self.path, "rb") as f:
            b = f.read(int(self.index[0]) if len(self.index) else os.path.getsize(self.path))
        if self.fmt == "MsgPack":
            return msgpack_decode(b)
        if self.fmt == "CBOR":
            return cbor_decode(b)
        return bincode_decode_settings(b)[0]

    def __getitem__(self, i):
        with open(self.path, "rb") as f:
            f.seek(int(self.index[i]))
            b = lzma.decompress(f.read(int(self.sizes[i])))
        return decode_entry(self.fmt, b)


def _tree_array(a):
    return None if a is None else np.asarray(a["data"], dtype=np.float64).reshape(a["dim"])


def decode_entry(fmt, b):
    if fmt == "Bincode":
        return bincode_decode_entry(b)
    t = msgpack_decode(b) if fmt == "MsgPack" else cbor_decode(b)
    d = t["distribution"]
    if d is not None:
        d = dict(d)
        d["dist"] = _tree_array(d["dist"])
    return {"distribution": d, "flowfield": _tree_array(t["flowfield"]),
            "magneticfield": _tree_array(t["magneticfield"]), "particles": _plain_particles(t["particles"]),
            "timestep": t["timestep"]}


def read_snapshot(path):
    """A snapshot file (any of the three formats, by extension)."""
    fmt = _format_of(path)
    b = lzma.decompress(open(path, "rb").read())
    if fmt == "Bincode":
        return bincode_decode_snapshot(b)
    t = msgpack_decode(b) if fmt == "MsgPack" else cbor_decode(b)
    comp = t.get("compensation")
    rng = [(g["state"], g["increment"]) if isinstance(g, dict) else tuple(g) for g in t["rng_state"]]
    return {"particles": _plain_particles(t["particles"]), "rng_state": rng, "timestep": t["timestep"],
            "compensation": None if comp is None else np.asarray(comp, dtype=np.float64).reshape(-1, 3)}


def read_particles(path):
    """An initial-condition file, ``Vec<Particle>`` (init.rs:23-49)."""
    fmt = _format_of(path)
    b = lzma.decompress(open(path, "rb").read())
    if fmt == "Bincode":
        return _BR(b).particles()
    return _plain_particles(msgpack_decode(b) if fmt == "MsgPack" else cbor_decode(b))


def write_particles(path, particles):
    fmt = _format_of(path)
    with open(path, "wb") as f:
        f.write(compress(encode(fmt, np.ascontiguousarray(particles, dtype=np.float64))))


def _format_of(path):
    for ext, fmt in FORMAT_OF_EXTENSION.items():
        if path.endswith("." + ext):
            return fmt
    raise ValueError(f"Do not recognise file extension of '{path}'")

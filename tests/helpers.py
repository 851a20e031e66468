"""Shared helpers of the test-suite: fixture readers, small configurations and
comparison utilities."""
import os
import struct

import numpy as np

from stochasticsampling_amd.params import make_params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_bincode_ndarray(path):
    """bincode-serialised ndarray 0.12 Array<f64, D>: u8 version, D x u64
    dims, u64 length, f64 little endian (as data only)."""
    b = open(path, "rb").read()
    assert b[0] == 1
    # the element count sits after the dims; find D from the file length
    for ndim in range(1, 8):
        off = 1 + 8 * ndim
        (n,) = struct.unpack("<Q", b[off : off + 8])
        dims = struct.unpack("<" + "Q" * ndim, b[1:off])
        if int(np.prod(dims)) == n and len(b) == off + 8 + 8 * n:
            return np.frombuffer(b[off + 8 :], "<f8").reshape(dims).copy()
    raise ValueError(f"cannot parse {path}")


def small_params(**kw):
    """A small configuration in the spirit of the CPU-runnable case."""
    base = dict(box=(10.0, 10.0, 10.0), grid=(8, 8, 8, 4, 4), n=2048, seed=7)
    base.update(kw)
    return make_params(**base)


def isotropic_particles(n, box, seed):
    """Uniform positions and isotropic orientations from numpy (test inputs)."""
    rng = np.random.default_rng(seed)
    p = np.empty((n, 5))
    p[:, 0] = rng.random(n) * box[0]
    p[:, 1] = rng.random(n) * box[1]
    p[:, 2] = rng.random(n) * box[2]
    p[:, 3] = rng.random(n) * 2 * np.pi
    p[:, 4] = np.arccos(1 - 2 * rng.random(n))
    return p


def rel_err(a, b):
    """max |a-b| / max |b| (with a floor so an all-zero field compares absolutely)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def periodic_diff(a, b, period):
    d = np.abs(a - b)
    return np.minimum(d, period - d)

"""CPU-side checks of libss: it loads, exports every entry point that
include/ss_api.h declares, rejects bad input on the host, and its host-only
initial-condition generator has the stated distribution."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

import stochasticsampling_amd as ss
from stochasticsampling_amd import lib as sslib
from stochasticsampling_amd.params import Params

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ss_api.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ss_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = sslib.load()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(sslib.EXPORTS)
    assert L.ss_version() == 4


def test_params_struct_layout_matches_header(tmp_path):
    """The ctypes mirror has the C compiler's layout of ss_params."""
    src = tmp_path / "layout.c"
    fields = [f[0] for f in Params._fields_]
    body = "\n".join(f'printf("%s %zu\\n", "{f}", offsetof(ss_params, {f}));' for f in fields)
    src.write_text(
        f'#include <stdio.h>\n#include <stddef.h>\n#include "{HEADER}"\n'
        f'int main(void){{ {body} printf("size %zu\\n", sizeof(ss_params)); return 0; }}\n'
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for f in fields:
        assert int(got[f]) == getattr(Params, f).offset, f
    assert int(got["size"]) == C.sizeof(Params)


def test_params_default_matches_python():
    L = sslib.load()
    p = Params()
    L.ss_params_default(C.byref(p))
    assert p.precision == ss.FP64 and p.noise_mode == ss.NOISE_PHILOX
    assert p.grid_size.x == 1 and p.box_size.x == 1.0 and p.number_of_particles == 1


def test_create_without_gpu_reports_device_error():
    if ss.device_count() > 0:
        pytest.skip("a GPU is visible")
    p = ss.make_params()
    with pytest.raises(ss.SSError) as e:
        ss.Simulation(p)
    assert e.value.code == -2 and "GPU" in str(e.value)


@pytest.mark.parametrize(
    "kw, code",
    [
        (dict(grid=(0, 4, 4, 4, 4)), -1),
        (dict(box=(0.0, 1.0, 1.0)), -1),
        (dict(n=0), -1),
        (dict(precision=7), -1),
        (dict(flags=ss.FLAG_QUASI2D), -1),
        (dict(grid=(1024, 1024, 1024, 64, 64)), -4),
    ],
)
def test_create_rejects_bad_parameters(kw, code):
    p = ss.make_params(**kw)
    with pytest.raises(ss.SSError) as e:
        ss.Simulation(p)
    assert e.value.code == code


def test_initial_condition_statistics():
    p = ss.make_params(box=(50.0, 40.0, 30.0), n=200000)
    parts = ss.initial_condition(p, amplitude=0.05, seed=2)
    assert parts.shape == (200000, 5)
    assert np.all((parts[:, 0] >= 0) & (parts[:, 0] < 50))
    assert np.all((parts[:, 1] >= 0) & (parts[:, 1] < 40))
    assert np.all((parts[:, 2] >= 0) & (parts[:, 2] < 30))
    assert np.all((parts[:, 3] >= 0) & (parts[:, 3] < 2 * math.pi))
    c = np.cos(parts[:, 4])
    s = np.sin(2 * math.pi * parts[:, 0] / 50)
    # E[cos theta | x] = amplitude * sin(2 pi x / L) / 3
    slope = (c * s).mean() / (s * s).mean()
    assert abs(slope - 0.05 / 3) < 0.006
    # marginally uniform in cos theta
    h, _ = np.histogram(c, bins=10, range=(-1, 1))
    assert np.all(np.abs(h / len(c) - 0.1) < 0.005)
    # deterministic
    again = ss.initial_condition(p, amplitude=0.05, seed=2)
    np.testing.assert_array_equal(parts, again)


def test_initial_condition_zero_amplitude_is_isotropic():
    p = ss.make_params(n=100000)
    parts = ss.initial_condition(p, amplitude=0.0, seed=9)
    c = np.cos(parts[:, 4])
    assert abs(c.mean()) < 0.01 and abs((c * c).mean() - 1 / 3) < 0.01

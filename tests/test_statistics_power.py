"""Power of the magnetic-suspension statistics comparison, on the CPU oracle
alone: the comparison of tests/test_gpu_statistics.py accepts a second set of
oracle seeds and rejects a copy of the oracle whose field torque
(external_field_alignment, modifiers.rs:103-118) is planted 10% too strong.
Fewer seeds and particles than the GPU test (4 seeds of 2^14 particles, 600
steps), so the GPU test's standard errors are smaller and its power larger."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import magnetic_stats as ms
from oracle_lib import ORACLE_DIR, Oracle

SRC = os.path.join(ORACLE_DIR, "ss_oracle.c")
PATTERN = "for (int c = 0; c < 3; ++c) b[c] = d[c] * p->magnetic_reorientation;"
PLANTED = "for (int c = 0; c < 3; ++c) b[c] = d[c] * (1.1 * p->magnetic_reorientation);"


@pytest.fixture(scope="module")
def strong_torque_oracle():
    text = open(SRC).read()
    assert text.count(PATTERN) == 1
    with tempfile.TemporaryDirectory(prefix="oracle_torque_") as tmp:
        src = os.path.join(tmp, "torque.c")
        with open(src, "w") as f:
            f.write(text.replace(PATTERN, PLANTED))
        so = os.path.join(tmp, "torque.so")
        subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=c11", "-shared", f"-I{ORACLE_DIR}", "-o", so, src,
                        "-lm"], check=True, capture_output=True)
        yield Oracle(so_path=so)


def _sample(o, seeds, n=1 << 14, steps=600, every=100):
    runs = [ms.oracle_run(o, n, s, steps, every) for s in seeds]
    return np.array([r[0] for r in runs]), np.array([r[1] for r in runs])


def test_comparison_accepts_oracle_and_rejects_planted_torque_error(oracle, strong_torque_oracle):
    ref_series, ref_hist = _sample(oracle, [200, 201, 202, 203])
    other_series, other_hist = _sample(oracle, [300, 301, 302, 303])
    assert ms.gap(other_series, ref_series) <= ms.SERIES_SE
    assert ms.gap(other_hist, ref_hist) <= ms.MARGINAL_SE
    bad_series, _ = _sample(strong_torque_oracle, [300, 301, 302, 303])
    assert ms.gap(bad_series, ref_series) > 2 * ms.SERIES_SE

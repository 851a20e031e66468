"""ss_fft_execute_host_noalloc (no device allocation per call) against
ss_fft_execute_host, the allocating entry point with the same kernels.

Two calls in a row on one GPU share the library's staging buffer; the
second, smaller batch must not see anything of the first.  The results are
bit-identical to the allocating entry point and meet the FFT bar against
numpy (1e-13 FP64, 2e-6 FP32 relative).
"""
import numpy as np
import pytest

import stochasticsampling_amd as ss
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _case(shape, batch, precision, direction, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((batch,) + shape) + 1j * rng.standard_normal((batch,) + shape)
    dt = np.complex128 if precision == ss.FP64 else np.complex64
    x = np.ascontiguousarray(x.astype(dt))
    plan = ss.FFTPlan(shape, direction, batch=batch, precision=precision)
    a = x.copy()
    plan.reexecute3d(a)
    b = x.copy()
    plan.reexecute3d_noalloc(b)
    np.testing.assert_array_equal(a, b)
    axes = (1, 2, 3)
    xr = x.astype(np.complex128)
    ref = np.fft.fftn(xr, axes=axes) if direction == -1 else np.fft.ifftn(xr, axes=axes) * np.prod(shape)
    assert rel_err(b, ref) < (1e-13 if precision == ss.FP64 else 2e-6)
    plan.close()


def test_noalloc_matches_allocating_entry_point_twice():
    # first call: eight FP64 arrays of 32^3 (4 MiB)
    _case((32, 32, 32), 8, ss.FP64, -1, 1)
    # second call on the same buffer: a smaller FP32 batch, backward
    _case((16, 12, 10), 2, ss.FP32, +1, 2)
    # and a small FP64 one after it
    _case((6, 5, 4), 3, ss.FP64, -1, 3)


def test_noalloc_rejects_batches_beyond_the_buffer():
    # 64^3 FP64 complex is 4 MiB per array: 17 of them exceed the 64 MiB buffer
    plan = ss.FFTPlan((64, 64, 64), -1, batch=17, precision=ss.FP64)
    a = np.zeros((17, 64, 64, 64), dtype=np.complex128)
    with pytest.raises(ss.SSError, match="exceeds the staging buffer"):
        plan.reexecute3d_noalloc(a)
    plan.close()
    # exactly at the limit is accepted
    plan = ss.FFTPlan((64, 64, 64), -1, batch=16, precision=ss.FP64)
    a = np.zeros((16, 64, 64, 64), dtype=np.complex128)
    a[:, 0, 0, 0] = 1.0
    plan.reexecute3d_noalloc(a)
    assert np.all(a == 1.0)
    plan.close()

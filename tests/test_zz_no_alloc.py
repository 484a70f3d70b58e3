"""harm_dump_columns_noalloc (no device allocation, staging in a fixed device
array) against harm_dump_columns (staging allocated per call), through the C ABI.

Two calls in a row on the same GPU, the second with a smaller zone range at
another offset: the shared array keeps the first call's rows, and none of them
may show through.  Sizes beyond the array are refused with error 3."""
import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params

pytestmark = pytest.mark.gpu


def _cols(g, fn, z0, n):
    out = np.empty((n, hc.DUMP_NCOL))
    rc = getattr(g.L, fn)(g.h, z0, n, out.ctypes.data)
    return rc, out


@pytest.mark.parametrize("shape", [(32, 32, 64), (64, 64, 1)])
def test_dump_columns_noalloc_matches_allocating_entry(shape):
    g = hc.Harm(torus_params(*shape))
    g.init()
    g.step_ch(2)
    na = shape[0] * shape[1] * shape[2]
    calls = [(37, hc.DUMP_NOALLOC_MAX_ZONES if na >= hc.DUMP_NOALLOC_MAX_ZONES + 37 else na - 37), (na - 1500, 1200)]
    for z0, n in calls:
        rc_a, want = _cols(g, "harm_dump_columns", z0, n)
        rc_b, got = _cols(g, "harm_dump_columns_noalloc", z0, n)
        assert rc_a == 0 and rc_b == 0, g.L.harm_last_error()
        assert np.array_equal(got, want), (z0, n)
    # a range beyond the fixed array: refused, nothing written
    big = min(na, hc.DUMP_NOALLOC_MAX_ZONES + 1)
    if big > hc.DUMP_NOALLOC_MAX_ZONES:
        out = np.full((big, hc.DUMP_NCOL), 7.0)
        rc = g.L.harm_dump_columns_noalloc(g.h, 0, big, out.ctypes.data)
        assert rc == 3
        assert b"HARM_DUMP_NOALLOC_MAX_ZONES" in g.L.harm_last_error()
        assert np.all(out == 7.0)
    # an empty range is a no-op
    assert g.L.harm_dump_columns_noalloc(g.h, 0, 0, None) == 0
    g.close()

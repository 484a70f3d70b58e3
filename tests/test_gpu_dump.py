"""ASCII dumps written on the GPU side (harm_dump_write) against the oracle's
restatement of dump() (dump.c:1710-1990): byte-identical files after the same
steps (3D wave-uniform, 3D, 2D, the tabulated EOS, and a failed run's short
lines), and harm_dump_columns equal to the file's values."""
import ctypes
import os

import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, OrcRestartHdr, lib as olib, orc_eos
from test_dump_oracle import read_dump

pytestmark = pytest.mark.gpu

DRIVER = dict(nstep=7, DTd=5.0, DTl=0.5, DTi=1.0, DTr=100, dump_cnt=3, image_cnt=4, rdump_cnt=1)


def _files_equal(a, b):
    x, y = open(a, "rb").read(), open(b, "rb").read()
    if x != y:
        k = next((i for i in range(min(len(x), len(y))) if x[i] != y[i]), min(len(x), len(y)))
        pytest.fail(f"files differ at byte {k} of {len(x)}/{len(y)}: {x[k - 60:k + 40]!r} vs {y[k - 60:k + 40]!r}")


@pytest.mark.parametrize("shape,tab,failed", [((32, 32, 64), False, 0), ((24, 16, 8), False, 0),
                                              ((48, 40, 1), False, 0), ((16, 16, 64), True, 0),
                                              ((16, 16, 64), False, 1)])
def test_dump_matches_oracle(tmp_path, shape, tab, failed):
    P = torus_params(*shape)
    g = hc.Harm(P)
    olib().orc_set_threads(16)
    o = Oracle(P)
    keep = None
    if tab:
        host = hc.eos_table_build(hc.eos_spec(rho_unit=1e8, nr=120, nt=100))
        dtab = hc.DeviceEos(host, P.gam)
        g.set_eos(dtab)
        keep = orc_eos(host, P.gam)
        olib().orc_set_eos(o.h, ctypes.byref(keep), 0)
    g.init()
    o.init_torus()
    for _ in range(3):
        g.step_ch(1)
        assert o.step() == 0
    if not np.array_equal(g.prims(), o.prims()):
        o.set_prims(g.prims())  # a Font-fix pow ulp: compare the writers on one state
    a, b = str(tmp_path / "gpu"), str(tmp_path / "orc")
    g.dump_write(a, failed=failed, **DRIVER)
    h = OrcRestartHdr(failed=failed, **DRIVER)
    assert olib().orc_dump_write(o.h, os.fsencode(b), ctypes.byref(h)) == 0
    _files_equal(a, b)
    if not failed:
        _, cols = read_dump(a)
        assert np.array_equal(g.dump_columns(), cols)
        assert np.array_equal(g.dump_columns(100, 50), cols[100:150])
    g.close()

"""The C-ABI library loads and exports every function include/harm_cool.h declares;
without a GPU the product refuses to run (no CPU fallback)."""
import ctypes
import os
import re

import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import HarmParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "harm_cool.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n not in ("sizeof",)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("harm_create", "harm_step", "harm_init_torus", "harm_utoprim2d_batch", "Utoprim_2d"):
        assert n in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_symbol(name):
    L = ctypes.CDLL(hc.LIB_PATH)
    assert hasattr(L, name), name


def test_fast_build_exports_the_same_abi():
    """libharmcool_fast.so (FMA-contracted device code) is a drop-in for the
    parity build: every declared entry point is there too."""
    F = ctypes.CDLL(hc.LIB_FAST_PATH)
    missing = [n for n in declared_functions() if not hasattr(F, n)]
    assert not missing, missing


def test_params_struct_layout():
    # 6 int32 + 18 doubles + 4 uint32, no padding holes beyond pad0
    assert ctypes.sizeof(HarmParams) == 6 * 4 + 18 * 8 + 4 * 4
    assert HarmParams.a.offset == 24


def test_product_fails_loudly_without_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(hc.HarmError):
        hc.Harm(hc.config1())
    with pytest.raises(hc.HarmError):
        hc.Harm(hc.config1(), fast=True)

"""The C ABI library: it loads without a GPU and exports every entry point
that include/fimhip.h declares; the host-side glue behaves on bad input."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from fimocn_subx_realtime_hera_apr2020_amd import fimhip, glvl, state
from helpers import golden_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "fimhip.h")).read()
    return sorted(set(re.findall(r"\b(fimhip_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_python_agree():
    assert declared_symbols() == sorted(fimhip.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = fimhip.load_library()
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_host_side_pow_is_the_device_routine():
    lib = fimhip.load_library()
    assert lib.fimhip_pow(2.0, 10.0) == 1024.0
    assert b"gfx950" in lib.fimhip_version()


def test_create_rejects_bad_sizes_without_gpu_calls():
    lib = fimhip.load_library()
    gd = fimhip.GridDesc()
    gd.nip, gd.nvl, gd.ntra, gd.dt = 10, 64, 4, 240.0
    h = ct.c_void_p()
    assert lib.fimhip_create(ct.byref(gd), 0, ct.byref(h)) == -1
    gd.nip, gd.nvl = 642, 65
    assert lib.fimhip_create(ct.byref(gd), 0, ct.byref(h)) == -1
    assert lib.fimhip_create(None, 0, ct.byref(h)) == -1


def test_every_entry_rejects_a_null_handle_without_gpu_calls():
    """Every entry point that takes an instance handle returns
    FIMHIP_EINVAL (-1) for a null one, with null arrays and zero sizes, and
    touches no device (this runs without a GPU)."""
    lib = fimhip.load_library()
    no_handle = {"fimhip_create", "fimhip_ens_create", "fimhip_icos_grid", "fimhip_icos_grid_info",
                 "fimhip_pow", "fimhip_ens_kernel_count", "fimhip_ens_kernel_name", "fimhip_synth_state",
                 "fimhip_version"}
    checked = 0
    for name in fimhip.EXPORTS:
        if name in no_handle:
            continue
        f = getattr(lib, name)
        assert f.argtypes, name
        args = [0.0 if t in (ct.c_double, ct.c_float) else 0 if t in (ct.c_int, ct.c_int64, ct.c_size_t) else None
                for t in f.argtypes]
        assert f(*args) == -1, name
        checked += 1
    assert checked >= 50


def test_glvl_roundtrip(tmp_path):
    g = golden_grid(2)
    p = tmp_path / "glvl.dat"
    glvl.write_glvl(g, str(p))
    assert p.read_bytes() == open(os.path.join(ROOT, "tests", "golden", "glvl_G2.dat"), "rb").read()


def test_state_generator_is_batch_invariant():
    g = golden_grid(3)
    a = state.make_state(g.lat, g.lon, 64, 4, wind_noise=True, perturb=True, member=3)
    b = state.make_state(g.lat[:100], g.lon[:100], 64, 4, wind_noise=True, perturb=True, member=3)
    for k in ("us", "dp"):
        assert np.array_equal(a[k][:100], b[k])
    c = state.make_state(g.lat, g.lon, 64, 4, wind_noise=True, perturb=True, member=4)
    assert not np.array_equal(a["us"], c["us"])

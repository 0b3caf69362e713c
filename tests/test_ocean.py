"""The ocean instance of the uneven batch on the CPU: the seeded ocean state
and land mask, and the checker's land/sea mask (no counterpart in the
reference's atmosphere; its rules follow the flux gating of
hycom_cnuity.F90:100-125).  The mask is pinned by what it must leave
unchanged: an all-wet mask is the reference path byte for byte, land columns
keep their thickness, winds and tracers, and global mass is conserved."""
import numpy as np
import pytest

from fimocn_subx_realtime_hera_apr2020_amd import constants as C
from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
from helpers import golden_grid


def _run(O, g, st, nsteps, wet, dt=720.0):
    th, ds = ocean.vertical_tables()
    o = O.OracleRun(O.grid_arrays(g, ocean.NVL, dt, th, ds), ocean.NVL, ocean.NTRA, dt, pow_kind=1)
    o.load(st)
    o.set_wet(wet)
    o.init(True)
    init = {f: o.a[f].copy() for f in ("dp", "us", "vs", "tr")}
    m0, t0 = o.globsum()
    for its in range(1, nsteps + 1):
        o.step(its, 0)
    return o, init, m0, t0


def test_ocean_state_and_mask():
    g = fimhip.icos_grid(5)
    st = ocean.make_ocean_state(g.lat, g.lon)
    wet = st["wet"]
    assert 0.29 < 1.0 - wet.mean() < 0.31
    # the land field is smooth: most cells agree with all their neighbours
    prox = np.asarray(g.prox) - 1
    same = np.array([np.all(wet[prox[c, :g.nprox[c]]] == wet[c]) for c in range(g.nip)])
    assert same.mean() > 0.9
    t, s = st["tr"]
    assert t.min() >= 270.0 and t.max() <= 303.0 and np.all(np.diff(t, axis=1) >= 0)
    assert s.min() >= 34.0 and s.max() <= 36.0
    dpmin = np.array(C.OCN_DPMIN32[::-1]) * ocean.PA_PER_M
    assert np.all(st["dp"] >= dpmin[None, :] * (1 - 1e-12))
    assert np.allclose(st["pr"][:, -1], ocean.P_TOP) and np.allclose(st["pr"][:, 0], ocean.P_TOP + st["dp"].sum(1))
    # seeded
    st2 = ocean.make_ocean_state(g.lat, g.lon)
    assert all(np.array_equal(st[k], st2[k]) for k in st)


def test_all_wet_mask_is_the_unmasked_path(oracle_lib):
    g = golden_grid(2)
    st = ocean.make_ocean_state(g.lat, g.lon)
    a, _, _, _ = _run(oracle_lib, g, st, 6, None)
    b, _, _, _ = _run(oracle_lib, g, st, 6, np.ones(g.nip, np.int32))
    for f in ("us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "sdot", "relvor"):
        assert np.array_equal(a.a[f], b.a[f]), f


def test_mask_keeps_land_and_conserves_mass(oracle_lib):
    g = golden_grid(3)
    st = ocean.make_ocean_state(g.lat, g.lon)
    land = st["wet"] == 0
    assert land.any() and not land.all()
    o, init, m0, t0 = _run(oracle_lib, g, st, 24, st["wet"])
    for f in ("dp", "us", "vs"):
        assert np.array_equal(o.a[f][land], init[f][land]), f
        assert not np.array_equal(o.a[f][~land], init[f][~land]), f
    assert np.array_equal(o.a["tr"][:, land], init["tr"][:, land])
    m1, t1 = o.globsum()
    assert abs(m1 - m0) / m0 < 1e-12
    # temperature and salinity totals move only by the FCT's bound clipping
    # and hybgen's theta redistribution, as in the unmasked reference path
    assert np.all(np.abs(t1 - t0) / t0 < 1e-3)
    for f in ("dp", "us", "vs", "tr", "pr", "ex"):
        assert np.isfinite(o.a[f]).all(), f
    # and the mask is not a no-op
    u, _, _, _ = _run(oracle_lib, g, st, 24, None)
    assert not np.array_equal(u.a["dp"], o.a["dp"])

"""The ocean instance of the uneven batch in the batched FP32 form on the
GPU: the land/sea mask (an all-wet mask is the unmasked path bit for bit,
land columns never change, global mass is conserved), agreement with the
checker under the same mask, and the atmosphere and ocean instances stepped
by one graph replay giving the bytes of each stepped alone."""
import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu

# relative-L2 bounds of the masked FP32 form against the masked FP64 checker
# after 24 full steps at G4 x 32 layers (dt = 720 s)
BOUNDS_OCEAN = {"us": 1e-3, "vs": 1e-3, "dp": 1e-3, "pr": 1e-4, "ex": 3e-5, "t": 1.5e-5, "s": 1.5e-5}


def _ocean(level, members=1, masked=True, fused=True, dt=None):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    g = fimhip.icos_grid(level)
    th, ds = ocean.vertical_tables()
    dt = dt or 1.2 * 4800.0 / float(2 ** level // 2)
    st = ocean.make_ocean_state(g.lat, g.lon)
    e = fimhip.Ensemble(g, members, ocean.NVL, ocean.NTRA, dt, th, ds)
    e.set_fused(fused)
    for m in range(members):
        e.load_member(m, st)
        if masked is True:
            e.set_wet(m, st["wet"])
        elif masked == "all_wet":
            e.set_wet(m, np.ones(g.nip, np.int32))
    e.init_state(True)
    return g, e, st


def _fields(e, m):
    return {n: e.download(m, n) for n in ("us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "trdp", "sdot")}


@pytest.mark.parametrize("fused", [True, False])
def test_all_wet_mask_is_the_unmasked_path(fused):
    out = []
    for masked in (False, "all_wet"):
        print("case: G3 x 32 ocean, 2 members, 6 steps, mask %s, fused %s" % (masked, fused), flush=True)
        g, e, _ = _ocean(3, members=2, masked=masked, fused=fused)
        try:
            e.step(6, 0)
            e.sync()
            out.append([_fields(e, m) for m in range(2)])
        finally:
            e.close()
    for m in range(2):
        for n in out[0][m]:
            assert np.array_equal(out[0][m][n], out[1][m][n]), (m, n)


def test_land_never_changes_and_mass_is_conserved():
    print("case: G4 x 32 ocean, masked, 24 steps", flush=True)
    g, e, st = _ocean(4)
    try:
        land = st["wet"] == 0
        init = _fields(e, 0)
        m0, t0, _, _ = e.globsum()
        e.step(24, 0)
        e.sync()
        f = _fields(e, 0)
        m1, t1, _, _ = e.globsum()
    finally:
        e.close()
    for n in ("dp", "us", "vs"):
        assert np.array_equal(f[n][land], init[n][land]), n
        assert not np.array_equal(f[n][~land], init[n][~land]), n
    assert np.array_equal(f["tr"][:, land], init["tr"][:, land])
    assert abs(m1[0] - m0[0]) / m0[0] < 1e-12, (m1 - m0) / m0
    print("T/S total drift:", (t1 - t0) / t0, flush=True)
    assert np.all(np.abs(t1 - t0) / t0 < 1e-3)
    for n, a in f.items():
        assert np.all(np.isfinite(a)), n


def test_masked_fp32_form_against_masked_checker():
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    from oracle import oracle as O
    level, nsteps = 4, 24
    print("case: G4 x 32 ocean, masked FP32 form vs masked FP64 checker, %d steps" % nsteps, flush=True)
    g, e, st = _ocean(level)
    dt = e.dt
    try:
        e.step(nsteps, 0)
        e.sync()
        f = _fields(e, 0)
    finally:
        e.close()
    th, ds = ocean.vertical_tables()
    o = O.OracleRun(O.grid_arrays(g, ocean.NVL, dt, th, ds), ocean.NVL, ocean.NTRA, dt, pow_kind=1)
    o.load(st)
    o.set_wet(st["wet"])
    o.init(True)
    for its in range(1, nsteps + 1):
        o.step(its, 0)
    errs = {n: rel_l2(f[n], o.a[n]) for n in ("us", "vs", "dp", "pr", "ex")}
    errs["t"] = rel_l2(f["tr"][0], o.a["tr"][0])
    errs["s"] = rel_l2(f["tr"][1], o.a["tr"][1])
    print("masked ocean rel L2 after %d steps:" % nsteps, {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    for n, b in BOUNDS_OCEAN.items():
        assert errs[n] < b, (n, errs[n], b)


@pytest.mark.skip(reason="the entry that captured several instances into one graph "
                         "(fimhip_ens_step_together) ended its first GPU run with a host segmentation fault "
                         "whose cause was not found; the entry is taken out (see the design notes)")
def test_atmosphere_and_ocean_in_one_graph_equal_each_alone():
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, glvl, state
    nsteps = 5
    res = []
    for together in (False, True):
        print("case: G2+G3 atmosphere x 64 and G3 ocean x 32 masked, %d steps, together %s" % (nsteps, together),
              flush=True)
        ga = glvl.concat([fimhip.icos_grid(2), fimhip.icos_grid(3)])
        th, ds = state.vertical_tables(64)
        atm = fimhip.Ensemble(ga, 1, 64, 4, 240.0, th, ds)
        try:
            atm.synth(0, seed=20240601, synth_member=0, flags=1)
            atm.init_state(True)
            _, oce, _ = _ocean(3)
            try:
                if together:
                    fimhip.step_together([atm, oce], nsteps, 0)
                else:
                    atm.step(nsteps, 0)
                    oce.step(nsteps, 0)
                    atm.sync()
                    oce.sync()
                res.append((_fields(atm, 0), _fields(oce, 0)))
            finally:
                oce.close()
        finally:
            atm.close()
    for i in range(2):
        for n in res[0][i]:
            assert np.array_equal(res[0][i][n], res[1][i][n]), (i, n)


def test_atmosphere_and_ocean_stepped_side_by_side_equal_each_alone():
    """The uneven batch stepped side by side (every instance's captured step
    enqueued on its own stream before any wait) gives the bytes of each
    instance stepped alone."""
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, glvl, state
    nsteps = 5
    res = []
    for together in (False, True):
        print("case: G2+G3 atmosphere x 64 and G3 ocean x 32 masked, %d steps, side by side %s" % (nsteps, together),
              flush=True)
        ga = glvl.concat([fimhip.icos_grid(2), fimhip.icos_grid(3)])
        th, ds = state.vertical_tables(64)
        atm = fimhip.Ensemble(ga, 1, 64, 4, 240.0, th, ds)
        try:
            atm.synth(0, seed=20240601, synth_member=0, flags=1)
            atm.init_state(True)
            _, oce, _ = _ocean(3)
            try:
                if together:
                    fimhip.step_concurrently([atm, oce], nsteps, 0)
                else:
                    atm.step(nsteps, 0)
                    atm.sync()
                    oce.step(nsteps, 0)
                    oce.sync()
                res.append((_fields(atm, 0), _fields(oce, 0)))
            finally:
                oce.close()
        finally:
            atm.close()
    for i in range(2):
        for n in res[0][i]:
            assert np.array_equal(res[0][i][n], res[1][i][n]), (i, n)

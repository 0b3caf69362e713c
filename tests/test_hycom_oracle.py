"""The checker's restatement of the HYCOM ocean continuity equation
(oracle/hycom_oracle.c, hycom_cnuity.F90 cnuity) against the reference built
from its own sources: byte for byte against the committed digests of the
reference's output (tests/golden/hycom_cnuity_ref.json) at G2 and G3 x 32
layers for every interface-smoothing branch and across the 10-day mass
correction, against the live reference at G4 where oracle/_ref is built, and
against the properties the equation fixes (mass conserved to rounding without
the drift correction, land columns untouched, duplicate edge fluxes equal and
opposite)."""
import hashlib
import json
import os

import numpy as np
import pytest

from fimocn_subx_realtime_hera_apr2020_amd import glvl, ocean
from helpers import GOLDEN, HYCOM_CASES

RECORDS = ("dp", "pres", "massflx", "mssfx_avg")


def _sha(v):
    return hashlib.sha256(np.ascontiguousarray(v, "<f8").tobytes()).hexdigest()




def _run(O, g, case, thkdff, biharm, nstep0, mass, nsteps=3):
    h = O.HycomCnuity(g, case, thkdff=thkdff, biharm=biharm, mass=mass)
    out = {n: [] for n in RECORDS + ("mass",)}
    for st in range(nstep0, nstep0 + nsteps):
        leapm = 1 + st % 2
        assert h.step(st, leapm, 3 - leapm) == 0, "duplicate edge fluxes differ"
        for n in RECORDS:
            out[n].append(h.a[n].ravel().copy())
        out["mass"].append(np.array(h.mass))
    return h, out


@pytest.mark.parametrize("level", [2, 3])
def test_hycom_cnuity_checker_equals_reference_digests(oracle_lib, level):
    with open(os.path.join(GOLDEN, "hycom_cnuity_ref.json")) as f:
        ref = json.load(f)
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G%d.dat" % level))
    case = ocean.hycom_cnuity_case(g, kdm=ref["kdm"])
    for name, thkdff, biharm, nstep0, mass in HYCOM_CASES:
        _, out = _run(oracle_lib, g, case, thkdff, biharm, nstep0, mass, ref["nsteps"])
        want = ref["runs"]["G%d_%s" % (level, name)]
        for n, vals in out.items():
            assert [_sha(v) for v in vals] == want[n], (name, n)


def test_hycom_cnuity_checker_equals_live_reference_g4(oracle_lib):
    from oracle import refrun
    if not os.path.exists(os.path.join(refrun.BIN, "hycomref")):
        pytest.skip("reference build (oracle/_ref) not present")
    path = refrun.ref_glvl(4)
    g = glvl.read_glvl(path)
    case = ocean.hycom_cnuity_case(g, kdm=32, seed=7)
    r = refrun.run_hycomref(path, case, nsteps=4)
    _, out = _run(oracle_lib, g, case, 0.05, 0.75, 1, (0.0, 0.0, 0.0), nsteps=4)
    for n in RECORDS + ("mass",):
        for a, b in zip(out[n], r[n]):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), n


def test_hycom_cnuity_conserves_mass_and_keeps_land(oracle_lib):
    """Without the drift correction the fluxes only move mass: over one step
    the ocean's total (sum of bottom pressure times area over the wet cells)
    changes by rounding only (later steps also mix in the other leapfrog
    level through the time smoothing, whose seeded total differs); land
    columns keep their thickness and pressure; every duplicate edge flux is
    the negative of its twin."""
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G3.dat"))
    case = ocean.hycom_cnuity_case(g, kdm=32)
    wet = case["wet"] > 0
    area = np.asarray(g.area, np.float64)
    m0 = (case["pres"][:, -1] * area)[wet].sum()
    for thkdff in (0.0, 0.05):
        h, out = _run(oracle_lib, g, case, thkdff, 0.75, 1, (0.0, 0.0, 0.0), nsteps=1)
        pres = h.a["pres"]
        m1 = (pres[:, -1] * area)[wet].sum()
        assert abs(m1 - m0) <= 1e-12 * m0, (thkdff, (m1 - m0) / m0)
        assert np.array_equal(h.a["dp"][:, ~wet], case["dp"][:, ~wet])
        assert np.array_equal(pres[~wet], case["pres"][~wet])
        prox = g.dyn_prox() - 1
        proxs = g.proxs - 1
        fx = h.a["massflx"]
        for i in np.flatnonzero(wet)[::7]:
            for e in range(g.nprox[i]):
                ix, ex = prox[i, e], proxs[i, e]
                if wet[ix]:
                    assert np.array_equal(fx[i, e], -fx[ix, ex])


EDG_RECORDS = ("pres", "um_edg", "vm_edg", "un_edg", "vn_edg", "dp_edg", "pr_edg", "geop_edg", "mont_edg",
               "uvsq_edg", "spcv_edg")


def _edgvar_chain(O, g, case, janjic):
    """bclin_edgvar then cnuity at step 1 (leapm 2), as hycom_run calls them;
    the records by the reference driver's names."""
    h = O.HycomCnuity(g, case)
    h.edgvar(g, 2, 1, janjic)
    out = {"pres": h.a["pres"].copy(), "dp_edg": h.a["dpm_edg"].copy()}
    for n in ("um_edg", "vm_edg", "un_edg", "vn_edg"):
        out[n] = h.a[n].copy()
    for n in h.EDG_OUT:
        out[n] = h.e[n].copy()
    assert h.step(1, 2, 1) == 0
    out.update(dp=h.a["dp"], pres2=h.a["pres"], massflx=h.a["massflx"], mssfx_avg=h.a["mssfx_avg"],
               mass=np.array(h.mass))
    return h, out


@pytest.mark.parametrize("level", [2, 3])
@pytest.mark.parametrize("janjic", [False, True])
def test_hycom_edgvar_checker_equals_reference_digests(oracle_lib, level, janjic):
    with open(os.path.join(GOLDEN, "hycom_edgvar_ref.json")) as f:
        ref = json.load(f)
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G%d.dat" % level))
    case = ocean.hycom_edgvar_case(g, kdm=ref["kdm"])
    _, out = _edgvar_chain(oracle_lib, g, case, janjic)
    want = ref["runs"]["G%d_janjic%d" % (level, int(janjic))]
    assert set(out) == set(want)
    for n, v in out.items():
        assert _sha(v) == want[n], n


def test_hycom_edgvar_checker_equals_live_reference_g4(oracle_lib):
    from oracle import refrun
    if not os.path.exists(os.path.join(refrun.BIN, "hycedgref")):
        pytest.skip("reference build (oracle/_ref) not present")
    path = refrun.ref_glvl(4)
    g = glvl.read_glvl(path)
    case = ocean.hycom_edgvar_case(g, kdm=32, seed=9)
    for janjic in (False, True):
        r = refrun.run_hycedgref(path, case, nstep=1, leapm=2, janjic=janjic, chain=True)
        _, out = _edgvar_chain(oracle_lib, g, case, janjic)
        for n, v in out.items():
            assert np.array_equal(np.ravel(v).view(np.uint64), np.ravel(r[n]).view(np.uint64)), (janjic, n)


def test_hycom_edgvar_properties(oracle_lib):
    """What the interpolation fixes: a horizontally uniform Montgomery
    potential and specific volume stay uniform on the edges (the weights
    8+8+1+1 over 18 and the sidewall blend sum to one), edge thicknesses and
    kinetic energies are non-negative, land cells and a pentagon's sixth
    slot get zero edge values, and without janjic_ocn geop_edg is not
    touched."""
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G3.dat"))
    case = ocean.hycom_edgvar_case(g, kdm=32)
    case["montg"] = np.tile(np.linspace(-3.0, 5.0, 32), (g.nip, 1))
    case["spcifv"] = np.tile(np.linspace(9.7e-4, 9.8e-4, 32), (g.nip, 1))
    h = oracle_lib.HycomCnuity(g, case)
    h.edgvar(g, 2, 1, False)
    wet = case["wet"] > 0
    m, sv = h.e["mont_edg"], h.e["spcv_edg"]
    for i in np.flatnonzero(wet):
        np_ = g.nprox[i]
        assert np.allclose(m[i, :np_], case["montg"][i], rtol=0, atol=4e-15), i
        assert np.allclose(sv[i, :np_], case["spcifv"][i], rtol=3e-16, atol=0), i
    assert (h.a["dpm_edg"] >= 0).all() and (h.e["uvsq_edg"] >= 0).all()
    for n in ("mont_edg", "uvsq_edg", "spcv_edg"):
        assert (h.e[n][~wet] == 0).all(), n
    assert (h.a["dpm_edg"][~wet] == 0).all() and (h.e["pr_edg"][~wet] == 0).all()
    pent = np.flatnonzero((g.nprox == 5) & wet)
    assert len(pent) > 0
    assert (h.a["um_edg"][pent, 5] == 0).all() and (h.a["dpm_edg"][pent, 5] == 0).all()
    assert (h.e["geop_edg"] == 7.0).all()


MOM_RECORDS = ("uclin", "vclin", "pres", "ustarb")


def _momtum(O, g, janjic, seed=None):
    import sys
    sys.path.insert(0, GOLDEN)
    import make_golden
    case = ocean.hycom_momtum_case(g, kdm=32) if seed is None else ocean.hycom_momtum_case(g, kdm=32, seed=seed)
    h, edges = make_golden.momtum_inputs(O, g, case)
    h.momtum(g, 2, 1, janjic)
    return case, edges, h, {"uclin": h.e["uclin"], "vclin": h.e["vclin"], "pres": h.a["pres"],
                            "ustarb": h.e["ustarb"]}


@pytest.mark.parametrize("level", [2, 3])
@pytest.mark.parametrize("janjic", [False, True])
def test_hycom_momtum_checker_equals_reference_digests(oracle_lib, level, janjic):
    with open(os.path.join(GOLDEN, "hycom_momtum_ref.json")) as f:
        ref = json.load(f)
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G%d.dat" % level))
    _, _, _, out = _momtum(oracle_lib, g, janjic)
    want = ref["runs"]["G%d_janjic%d" % (level, int(janjic))]
    assert set(out) == set(want)
    for n, v in out.items():
        assert _sha(v) == want[n], n


def test_hycom_momtum_checker_equals_live_reference_g4(oracle_lib):
    from oracle import refrun
    if not os.path.exists(os.path.join(refrun.BIN, "hycmomref")):
        pytest.skip("reference build (oracle/_ref) not present")
    path = refrun.ref_glvl(4)
    g = glvl.read_glvl(path)
    for janjic in (False, True):
        case, edges, _, out = _momtum(oracle_lib, g, janjic, seed=13)
        r = refrun.run_hycmomref(path, case, edges, nstep=1, leapm=2, janjic=janjic)
        for n, v in out.items():
            assert np.array_equal(np.ravel(v).view(np.uint64), np.ravel(r[n]).view(np.uint64)), (janjic, n)


def test_hycom_momtum_properties(oracle_lib):
    """Land columns keep their velocities, pressure and friction velocity; the
    friction velocity of the ocean is non-negative; the bottom layers that
    the downward mixing reaches (thinner than their mixing slab) exist in the
    input, so that the recurrence is exercised."""
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G3.dat"))
    case, _, h, out = _momtum(oracle_lib, g, False)
    wet = case["wet"] > 0
    for n in ("uclin", "vclin"):
        assert np.array_equal(out[n][:, ~wet], case[n][:, ~wet]), n
        assert not np.array_equal(out[n][:, wet], case[n][:, wet]), n
    assert np.array_equal(out["pres"][~wet], case["pres"][~wet])
    assert np.array_equal(out["ustarb"][~wet], case["ustarb"][~wet]) and (out["ustarb"][wet] >= 0).all()
    pres, dpn, kdm = out["pres"], case["dp"][0], case["kdm"]
    mixed = 0
    for k in range(1, kdm):
        slab = np.maximum(0.1 * ocean.ONEM, pres[:, k + 1] - pres[:, kdm] + 10 * ocean.ONEM)
        mixed += int(np.sum((slab > np.minimum(dpn[:, k], slab)) & wet))
    assert mixed > 100


TR_RECORDS = ("cumuflx0", "dpinit0", "cumuflx1", "temp", "saln", "passv")


def _transport(O, g, seed=None):
    case = ocean.hycom_transport_case(g, kdm=32) if seed is None else ocean.hycom_transport_case(g, kdm=32, seed=seed)
    h0 = O.HycomCnuity(g, case)
    h0.transport(leapn=1, transp0=True)
    h = O.HycomCnuity(g, case)
    h.transport()
    return case, h, {"cumuflx0": h0.t["cumuflx"], "dpinit0": h0.t["dpinit"], "cumuflx1": h.t["cumuflx"],
                     "temp": h.t["temp"], "saln": h.t["saln"], "passv": h.t["passv"]}


@pytest.mark.parametrize("level", [2, 3])
def test_hycom_transport_checker_equals_reference_digests(oracle_lib, level):
    with open(os.path.join(GOLDEN, "hycom_transport_ref.json")) as f:
        ref = json.load(f)
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G%d.dat" % level))
    _, _, out = _transport(oracle_lib, g)
    want = ref["runs"]["G%d" % level]
    assert set(out) == set(want)
    for n, v in out.items():
        assert _sha(v) == want[n], n


def test_hycom_transport_checker_equals_live_reference_g4(oracle_lib):
    from oracle import refrun
    if not os.path.exists(os.path.join(refrun.BIN, "hyctrref")):
        pytest.skip("reference build (oracle/_ref) not present")
    path = refrun.ref_glvl(4)
    g = glvl.read_glvl(path)
    case, _, out = _transport(oracle_lib, g, seed=17)
    r = refrun.run_hyctrref(path, case, leapn=1)
    for n, v in out.items():
        assert np.array_equal(np.ravel(v).view(np.uint64), np.ravel(r[n]).view(np.uint64)), n


def test_hycom_transport_properties(oracle_lib):
    """What the transport fixes: each tracer's mass-weighted mean over the
    ocean is restored to rounding by the conservation step (the diffusion of
    temperature and salinity that follows conserves it too, since each edge's
    two fluxes cancel), a tracer that is constant stays constant, land
    columns keep their values, and the vertical displacements reach past a
    whole layer in both directions (the multi-layer walks are exercised)."""
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G3.dat"))
    case, h, out = _transport(oracle_lib, g)
    wet = case["wet"] > 0
    area = np.asarray(g.area, np.float64)
    for n in ("temp", "saln"):
        m0 = (case[n] * case["dpinit"] * area[:, None])[wet].sum() / (case["dpinit"] * area[:, None])[wet].sum()
        m1 = (out[n] * case["dpfinl"] * area[:, None])[wet].sum() / (case["dpfinl"] * area[:, None])[wet].sum()
        assert abs(m1 - m0) <= 1e-12 * abs(m0), (n, m1 - m0)
        assert np.array_equal(out[n][~wet], case[n][~wet]), n
    c = dict(case)
    c["passv"] = np.full_like(case["passv"], 0.25)
    h2 = oracle_lib.HycomCnuity(g, c)
    h2.transport()
    assert np.allclose(h2.t["passv"][0][wet], 0.25, rtol=0, atol=1e-13)


HYB_RUNS = ((1, False), (2, True))


@pytest.mark.parametrize("level", [2, 3])
def test_hycom_hybgen_checker_equals_reference_digests(oracle_lib, level):
    with open(os.path.join(GOLDEN, "hycom_hybgen_ref.json")) as f:
        ref = json.load(f)
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G%d.dat" % level))
    case = ocean.hycom_hybgen_case(g, kdm=32)
    for leap, wt in HYB_RUNS:
        out = oracle_lib.hycom_hybgn1(case, leap, with_tracers=wt)
        want = ref["runs"]["G%d_leap%d%s" % (level, leap, "_tr" if wt else "")]
        assert set(out) == set(want)
        for n, v in out.items():
            assert _sha(v) == want[n], (leap, wt, n)


def test_hycom_hybgen_checker_equals_live_reference_g4(oracle_lib):
    from oracle import refrun
    if not os.path.exists(os.path.join(refrun.BIN, "hychybref")):
        pytest.skip("reference build (oracle/_ref) not present")
    g = glvl.read_glvl(refrun.ref_glvl(4))
    case = ocean.hycom_hybgen_case(g, kdm=32, seed=23, numtr=2)
    for leap, wt in HYB_RUNS:
        out = oracle_lib.hycom_hybgn1(case, leap, with_tracers=wt)
        r = refrun.run_hychybref(case, leap=leap, with_tracers=wt)
        for n, v in out.items():
            assert np.array_equal(np.ravel(v).view(np.uint64), np.ravel(r[n]).view(np.uint64)), (leap, n)


def test_hycom_hybgen_properties(oracle_lib):
    """What the grid generator fixes: each column's heat and salt content
    (the thickness-weighted sums) and its bottom pressure are kept to
    rounding, interfaces stay ordered, land and the other leapfrog level are
    untouched; and the case moves interfaces in most ocean columns (the
    checker's branches are exercised: a coverage run of it on this case
    leaves only pcwise's decreasing-coordinate branch and one ratio branch of
    equilb's scenario 4 unexecuted)."""
    g = glvl.read_glvl(os.path.join(GOLDEN, "glvl_G3.dat"))
    case = ocean.hycom_hybgen_case(g, kdm=32)
    wet = case["wet"] > 0
    out = oracle_lib.hycom_hybgn1(case, 2, with_tracers=True)
    for n in ("temp", "saln"):
        c0 = (case[n] * case["dp"][1]).sum(axis=1)
        c1 = (out[n] * out["dp"][1]).sum(axis=1)
        assert np.max(np.abs(c1 - c0)[wet] / np.abs(c0[wet])) < 1e-13, n
    tr0 = (case["passv"][0] * case["dp"][1]).sum(axis=1)
    tr1 = (out["passv"][0] * out["dp"][1]).sum(axis=1)
    assert np.max(np.abs(tr1 - tr0)[wet] / np.abs(tr0[wet])) < 1e-12
    bottom = case["pres"][:, 0].copy()          # the column's bottom from level 2's thicknesses
    for k in range(case["kdm"]):
        bottom = bottom + case["dp"][1][:, k]
    assert np.array_equal(out["pres"][wet, -1], bottom[wet])
    assert np.array_equal(out["pres"][~wet], case["pres"][~wet])
    assert (np.diff(out["pres"][wet], axis=1) >= 0).all()
    assert (out["dp"][1][wet] >= 0).all()
    for n in ("dens", "temp", "saln"):
        assert np.array_equal(out[n][~wet], case[n][~wet]), n
    for n in ("dp", "uvel", "vvel"):
        assert np.array_equal(out[n][0], case[n][0]), n
    moved = (out["dp"][1] != case["dp"][1]).any(axis=1)
    assert moved[wet].mean() > 0.5

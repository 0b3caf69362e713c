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

"""The HYCOM ocean's continuity equation on the GPU (fimhip_ocn_*, the FP64
kernels of hycom_cnuity.hip) against the checker's restatement
(oracle/hycom_oracle.c, pinned to the reference's hycom_cnuity.F90 bit for
bit): every output array and the mass state must hold the checker's bytes,
through the host-array drop-in and the device-resident path, for every
interface-smoothing branch (biharmonic, Laplacian, none), across the 10-day
mass correction, with 32 layers (two columns per wavefront) and with 48 (one
column per wavefront), at G3 to G7 (163842 cells)."""
import numpy as np
import pytest

from helpers import HYCOM_CASES

pytestmark = pytest.mark.gpu

ARRAYS = ("dp", "pres", "massflx", "mssfx_avg")
CASES = {c[0]: c[1:] for c in HYCOM_CASES}




def _grid(level):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip
    return fimhip.icos_grid(level)


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _check(O, g, case, thkdff, biharm, nstep0, mass, nsteps, resident):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip
    chk = O.HycomCnuity(g, case, thkdff=thkdff, biharm=biharm, mass=mass)
    ocn = fimhip.OceanCnuity(g, case["kdm"], case["wet"], case["dtleap"], thkdff=thkdff, biharm=biharm)
    try:
        ocn.mass = mass
        a = {n: np.ascontiguousarray(case[n], np.float64).copy() for n in fimhip.OCN_FIELDS}
        if resident:
            for n in fimhip.OCN_FIELDS:
                ocn.upload(n, a[n])
        for st in range(nstep0, nstep0 + nsteps):
            leapm = 1 + st % 2
            assert chk.step(st, leapm, 3 - leapm) == 0
            if resident:
                ocn.step(st, leapm, 3 - leapm)
                got = {n: ocn.download(n) for n in ARRAYS}
            else:
                ocn.cnuity(st, leapm, 3 - leapm, *(a[n] for n in fimhip.OCN_FIELDS[2:7]),
                           *(a[n] for n in ARRAYS))
                got = a
            for n in ARRAYS:
                assert _same(got[n], chk.a[n].reshape(got[n].shape)), (st, n)
            assert ocn.mass == chk.mass, (st, ocn.mass, chk.mass)
    finally:
        ocn.close()


@pytest.mark.parametrize("name", ["biharm", "laplace", "nosmooth", "masscor"])
def test_dropin_matches_checker_bitwise_g3(oracle_lib, name):
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    thkdff, biharm, nstep0, mass = CASES[name]
    print("case: G3 x 32, host-array drop-in, %s" % name, flush=True)
    g = _grid(3)
    _check(oracle_lib, g, ocean.hycom_cnuity_case(g, kdm=32), thkdff, biharm, nstep0, mass, 3, resident=False)


@pytest.mark.parametrize("level,kdm,name", [(4, 48, "biharm"), (4, 48, "laplace"), (5, 32, "biharm"),
                                            (5, 32, "masscor")])
def test_resident_matches_checker_bitwise(oracle_lib, level, kdm, name):
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    thkdff, biharm, nstep0, mass = CASES[name]
    print("case: G%d x %d, resident, %s" % (level, kdm, name), flush=True)
    g = _grid(level)
    _check(oracle_lib, g, ocean.hycom_cnuity_case(g, kdm=kdm, seed=11 + level), thkdff, biharm, nstep0, mass, 4,
           resident=True)


def test_leapfrog_start_with_changing_dtleap(oracle_lib):
    """hycom_run sets dtleap = batrop at the first step and 2 * batrop after
    (hycom_run.F90:136): the instance takes the new value between calls."""
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    print("case: G4 x 32, dtleap 900 s then 1800 s, drop-in", flush=True)
    g = _grid(4)
    case = ocean.hycom_cnuity_case(g, kdm=32, dtleap=900.0)
    chk = oracle_lib.HycomCnuity(g, case)
    ocn = fimhip.OceanCnuity(g, 32, case["wet"], 900.0)
    try:
        a = {n: np.ascontiguousarray(case[n], np.float64).copy() for n in fimhip.OCN_FIELDS}
        for st in (1, 2, 3):
            dtleap = 900.0 * min(st, 2)
            chk.g.dtleap = dtleap
            ocn.set_dtleap(dtleap)
            leapm = 1 + st % 2
            assert chk.step(st, leapm, 3 - leapm) == 0
            ocn.cnuity(st, leapm, 3 - leapm, *(a[n] for n in fimhip.OCN_FIELDS[2:7]), *(a[n] for n in ARRAYS))
            for n in ARRAYS:
                assert _same(a[n], chk.a[n].reshape(a[n].shape)), (st, n)
            assert ocn.mass == chk.mass
        with pytest.raises(fimhip.FimhipError):
            ocn.set_dtleap(0.0)
    finally:
        ocn.close()


def test_g7_full_size_matches_checker_bitwise(oracle_lib):
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    print("case: G7 x 32, resident, biharm, 3 steps", flush=True)
    g = _grid(7)
    _check(oracle_lib, g, ocean.hycom_cnuity_case(g, kdm=32), 0.05, 0.75, 1, (0.0, 0.0, 0.0), 3, resident=True)


def test_bad_arguments_are_rejected():
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    g = _grid(2)
    case = ocean.hycom_cnuity_case(g, kdm=16)
    with pytest.raises(fimhip.FimhipError):
        fimhip.OceanCnuity(g, 65, case["wet"], 1800.0)
    ocn = fimhip.OceanCnuity(g, 16, case["wet"], 1800.0)
    try:
        for n in fimhip.OCN_FIELDS:
            ocn.upload(n, case[n])
        for bad in ((1, 1, 1), (1, 0, 2), (0, 1, 2), (1, 2, 3)):
            with pytest.raises(fimhip.FimhipError):
                ocn.step(*bad)
        with pytest.raises(ValueError):
            ocn.upload("dp", case["dp"][0])
    finally:
        ocn.close()

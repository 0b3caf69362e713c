"""The HYCOM ocean's dynamics on the GPU (fimhip_ocn_*, the FP64 kernels of
hycom_edgvar.hip, hycom_cnuity.hip, hycom_momtum.hip and hycom_transp.hip)
against the checker's restatement (oracle/hycom_oracle.c, pinned to the
reference's hycom_edgvar.F90, hycom_cnuity.F90, hycom_momtum.F90 with
hycom_dissip.F90, and hycom_transp3d.F90 with hycom_fct3d.F90 and
hycom_dffusn.F90, bit for bit): every output array and the mass state must
hold the checker's bytes, through the host-array drop-ins and the
device-resident path, for every interface-smoothing branch (biharmonic,
Laplacian, none), across the 10-day mass correction, with and without
janjic_ocn, with 0 to 3 passive tracers and with and without the tracers'
diffusion, with 32 layers (two columns per wavefront) and with 48 (one
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
        a = {n: np.ascontiguousarray(case[n], np.float64).copy() for n in fimhip.OCN_CNUITY_FIELDS}
        if resident:
            for n in fimhip.OCN_CNUITY_FIELDS:
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
        a = {n: np.ascontiguousarray(case[n], np.float64).copy() for n in fimhip.OCN_CNUITY_FIELDS}
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
        for n in fimhip.OCN_CNUITY_FIELDS:
            ocn.upload(n, case[n])
        for bad in ((1, 1, 1), (1, 0, 2), (0, 1, 2), (1, 2, 3)):
            with pytest.raises(fimhip.FimhipError):
                ocn.step(*bad)
        with pytest.raises(ValueError):
            ocn.upload("dp", case["dp"][0])
    finally:
        ocn.close()


EDG_CHECK = ("um_edg", "vm_edg", "un_edg", "vn_edg", "dpm_edg", "pres")


def _edg_same(h, got, tag):
    for n in EDG_CHECK:
        assert _same(got[n], h.a[n].reshape(got[n].shape)), (tag, n)
    for n in h.EDG_OUT:
        assert _same(got[n], h.e[n].reshape(got[n].shape)), (tag, n)


@pytest.mark.parametrize("janjic", [False, True])
def test_edgvar_dropin_matches_checker_bitwise_g3(oracle_lib, janjic):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    print("case: G3 x 32, bclin_edgvar drop-in, janjic %s" % janjic, flush=True)
    g = _grid(3)
    case = ocean.hycom_edgvar_case(g, kdm=32)
    h = oracle_lib.HycomCnuity(g, case)
    o = fimhip.Ocean(g, 32, case["wet"], case["dtleap"])
    try:
        o.set_janjic(janjic)
        a = {n: np.ascontiguousarray(case[n], np.float64).copy() for n in fimhip.OCN_EDGVAR_FIELDS}
        for st in (1, 2):
            leapm = 1 + st % 2
            h.edgvar(g, leapm, 3 - leapm, janjic)
            o.edgvar(st, leapm, 3 - leapm, a["uclin"], a["vclin"], a["utrop"], a["vtrop"], a["dp"], a["pres"],
                     a["geop"], a["montg"], a["spcifv"], a["um_edg"], a["vm_edg"], a["un_edg"], a["vn_edg"],
                     a["dpm_edg"], a["pr_edg"], a["geop_edg"], a["mont_edg"], a["uvsq_edg"], a["spcv_edg"])
            _edg_same(h, a, st)
    finally:
        o.close()


@pytest.mark.parametrize("level,kdm,janjic", [(4, 48, True), (5, 32, False), (5, 32, True)])
def test_edgvar_then_cnuity_resident_matches_checker(oracle_lib, level, kdm, janjic):
    """bclin_edgvar then cnuity each step, as hycom_run calls them, on the
    device-resident arrays: 3 leapfrog steps from step 1 (the mass correction
    included)."""
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    print("case: G%d x %d, edgvar + cnuity resident, janjic %s" % (level, kdm, janjic), flush=True)
    g = _grid(level)
    case = ocean.hycom_edgvar_case(g, kdm=kdm, seed=31 + level)
    h = oracle_lib.HycomCnuity(g, case)
    o = fimhip.Ocean(g, kdm, case["wet"], case["dtleap"])
    try:
        o.set_janjic(janjic)
        for n in fimhip.OCN_EDGVAR_FIELDS:
            o.upload(n, case[n])
        for st in (1, 2, 3):
            leapm = 1 + st % 2
            h.edgvar(g, leapm, 3 - leapm, janjic)
            assert h.step(st, leapm, 3 - leapm) == 0
            o.step_edgvar(leapm, 3 - leapm)
            o.step(st, leapm, 3 - leapm)
            got = {n: o.download(n) for n in EDG_CHECK[:5] + tuple(h.EDG_OUT)}
            got["pres"] = h.a["pres"]     # pres is checked after cnuity below
            _edg_same(h, got, st)
            for n in ARRAYS:
                assert _same(o.download(n), h.a[n]), (st, n)
            assert o.mass == h.mass, st
    finally:
        o.close()


def test_edgvar_then_cnuity_g7_full_size(oracle_lib):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    print("case: G7 x 32, edgvar + cnuity resident, 2 steps", flush=True)
    g = _grid(7)
    case = ocean.hycom_edgvar_case(g, kdm=32)
    h = oracle_lib.HycomCnuity(g, case)
    o = fimhip.Ocean(g, 32, case["wet"], case["dtleap"])
    try:
        for n in fimhip.OCN_EDGVAR_FIELDS:
            o.upload(n, case[n])
        for st in (1, 2):
            leapm = 1 + st % 2
            h.edgvar(g, leapm, 3 - leapm, False)
            assert h.step(st, leapm, 3 - leapm) == 0
            o.step_edgvar(leapm, 3 - leapm)
            o.step(st, leapm, 3 - leapm)
        for n in ARRAYS:
            assert _same(o.download(n), h.a[n]), n
        for n in ("um_edg", "dpm_edg"):
            assert _same(o.download(n), h.a[n]), n
        for n in h.EDG_OUT:
            assert _same(o.download(n), h.e[n]), n
    finally:
        o.close()


MOM_OUT = ("uclin", "vclin", "pres", "ustarb")


def _mom_setup(O, g, case, janjic_edges=True):
    """The checker instance with the edge values momtum reads formed by its
    bclin_edgvar (janjic_ocn on, so geop_edg is set) and the case's pres."""
    h = O.HycomCnuity(g, case)
    h.edgvar(g, 2, 1, janjic_edges)
    h.a["pres"][:] = case["pres"]
    return h


@pytest.mark.parametrize("janjic", [False, True])
def test_momtum_dropin_matches_checker_bitwise_g3(oracle_lib, janjic):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    print("case: G3 x 32, momtum drop-in, janjic %s" % janjic, flush=True)
    g = _grid(3)
    case = ocean.hycom_momtum_case(g, kdm=32)
    h = _mom_setup(oracle_lib, g, case)
    a = {n: h.a[n].copy() for n in ("um_edg", "vm_edg", "dp", "pres")}
    a.update({n: h.e[n].copy() for n in ("utrop", "vtrop", "uclin", "vclin", "geop_edg", "mont_edg", "uvsq_edg",
                                          "spcv_edg")})
    a.update({n: np.ascontiguousarray(case[n], np.float64).copy() for n in ("taux", "tauy", "ustarb")})
    ub, vb = np.full(g.nip, 5.0), np.full(g.nip, 5.0)
    o = fimhip.Ocean(g, 32, case["wet"], case["dtleap"])
    try:
        o.set_momtum(case["batrop"], case["veldff"], case["smagcf"])
        o.set_janjic(janjic)
        h.momtum(g, 2, 1, janjic)
        o.momtum(1, 2, 1, a["utrop"], a["vtrop"], a["uclin"], a["vclin"], a["um_edg"], a["vm_edg"], a["geop_edg"],
                 a["mont_edg"], a["uvsq_edg"], a["spcv_edg"], a["taux"], a["tauy"], ub, vb, a["dp"], a["pres"],
                 a["ustarb"])
        for n in MOM_OUT:
            want = h.a[n] if n == "pres" else h.e[n]
            assert _same(a[n], want), n
        assert (ub == 5.0).all() and (vb == 5.0).all()
    finally:
        o.close()


@pytest.mark.parametrize("level,kdm,janjic", [(4, 48, True), (5, 32, False)])
def test_ocean_step_resident_matches_checker(oracle_lib, level, kdm, janjic):
    """bclin_edgvar, cnuity and momtum each step, in hycom_run's order, on the
    device-resident arrays: 3 leapfrog steps from step 1, the velocities that
    momtum writes feeding the next step's edge values."""
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    print("case: G%d x %d, edgvar + cnuity + momtum resident, janjic %s" % (level, kdm, janjic), flush=True)
    g = _grid(level)
    case = ocean.hycom_momtum_case(g, kdm=kdm, seed=41 + level)
    h = oracle_lib.HycomCnuity(g, case)
    o = fimhip.Ocean(g, kdm, case["wet"], case["dtleap"])
    try:
        o.set_janjic(janjic)
        o.set_momtum(case["batrop"], case["veldff"], case["smagcf"])
        for n in fimhip.OCN_MOMTUM_FIELDS:
            o.upload(n, case[n])
        for st in (1, 2, 3):
            leapm = 1 + st % 2
            h.edgvar(g, leapm, 3 - leapm, janjic)
            assert h.step(st, leapm, 3 - leapm) == 0
            h.momtum(g, leapm, 3 - leapm, janjic)
            o.step_edgvar(leapm, 3 - leapm)
            o.step(st, leapm, 3 - leapm)
            o.step_momtum(leapm, 3 - leapm)
            for n in ARRAYS:
                assert _same(o.download(n), h.a[n]), (st, n)
            for n in ("uclin", "vclin", "ustarb"):
                assert _same(o.download(n), h.e[n]), (st, n)
            assert o.mass == h.mass, st
    finally:
        o.close()


def test_ocean_step_g7_full_size(oracle_lib):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    print("case: G7 x 32, edgvar + cnuity + momtum resident, 2 steps", flush=True)
    g = _grid(7)
    case = ocean.hycom_momtum_case(g, kdm=32)
    h = oracle_lib.HycomCnuity(g, case)
    o = fimhip.Ocean(g, 32, case["wet"], case["dtleap"])
    try:
        o.set_momtum(case["batrop"], case["veldff"], case["smagcf"])
        for n in fimhip.OCN_MOMTUM_FIELDS:
            o.upload(n, case[n])
        for st in (1, 2):
            leapm = 1 + st % 2
            h.edgvar(g, leapm, 3 - leapm, False)
            assert h.step(st, leapm, 3 - leapm) == 0
            h.momtum(g, leapm, 3 - leapm, False)
            o.step_edgvar(leapm, 3 - leapm)
            o.step(st, leapm, 3 - leapm)
            o.step_momtum(leapm, 3 - leapm)
        for n in ARRAYS:
            assert _same(o.download(n), h.a[n]), n
        for n in ("uclin", "vclin", "ustarb"):
            assert _same(o.download(n), h.e[n]), n
    finally:
        o.close()


TR_OUT = ("temp", "saln", "passv")


def _tr_case(g, kdm, numtr, seed, temdff=None):
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    case = ocean.hycom_transport_case(g, kdm=kdm, numtr=numtr, seed=seed)
    if temdff is not None:
        case["temdff"] = temdff
    return case


def _tr_ocean(g, case):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip
    o = fimhip.Ocean(g, case["kdm"], case["wet"], case["dtleap"])
    o.set_momtum(case["batrop"])
    o.set_transport(case["numtr"], case["temdff"])
    return o


@pytest.mark.parametrize("numtr,temdff", [(1, None), (0, 0.0), (3, None)])
def test_transport_dropin_matches_checker_bitwise_g3(oracle_lib, numtr, temdff):
    """transp0, transp1 and transp2 through the host-array drop-ins."""
    print("case: G3 x 32, transport drop-in, numtr %d, temdff %s" % (numtr, temdff), flush=True)
    g = _grid(3)
    case = _tr_case(g, 32, numtr, 29 + numtr, temdff)
    o = _tr_ocean(g, case)
    try:
        h0 = oracle_lib.HycomCnuity(g, case)
        h0.transport(leapn=2, transp0=True)
        cu = np.ascontiguousarray(case["cumuflx"], np.float64).copy()
        di = np.ascontiguousarray(case["dpinit"], np.float64).copy()
        o.transp0(1, 2, cu, np.ascontiguousarray(case["dp"], np.float64), di)
        assert _same(cu, h0.t["cumuflx"]) and _same(di, h0.t["dpinit"])
        h = oracle_lib.HycomCnuity(g, case)
        h.transport()
        a = {n: np.ascontiguousarray(case[n], np.float64).copy()
             for n in ("cumuflx", "massflx_tr", "dpinit", "dpfinl", "temp", "saln", "passv")}
        o.transp1(2, a["cumuflx"], a["massflx_tr"])
        assert _same(a["cumuflx"], h.t["cumuflx"])
        o.transp2(2, a["temp"], a["saln"], a["passv"] if numtr else None, a["cumuflx"], a["dpinit"], a["dpfinl"])
        for n in TR_OUT:
            assert _same(a[n], h.t[n]), n
    finally:
        o.close()


@pytest.mark.parametrize("level,kdm,numtr", [(4, 48, 2), (5, 32, 1)])
def test_transport_resident_matches_checker(oracle_lib, level, kdm, numtr):
    """The device-resident entries as hycom_run calls them: transp0 from dp's
    new level, transp1 on the step's mass flux, transp2 with dpfinl = dp's
    level leapn; two transport intervals, the second starting from the
    first's tracers."""
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip
    print("case: G%d x %d, transport resident, numtr %d" % (level, kdm, numtr), flush=True)
    g = _grid(level)
    case = _tr_case(g, kdm, numtr, 53 + level, None)
    o = _tr_ocean(g, case)
    try:
        dp = np.stack([case["dpinit"], case["dpfinl"]])
        c = dict(case, dp=dp)
        for n in ("temp", "saln", "passv", "cumuflx", "dpinit"):
            o.upload(n, case[n])
        o.upload("dp", dp)
        o.upload("massflx", case["massflx_tr"])
        h = oracle_lib.HycomCnuity(g, c)
        h.transport()
        o.step_transp1()
        o.step_transp2(2)
        for n in TR_OUT + ("cumuflx",):
            assert _same(o.download(n), h.t[n]), n
        # the next interval: dpinit from dp's level 2, a fresh flux integral
        h.transport(leapn=2, transp0=True)
        o.step_transp0(2)
        for n in ("cumuflx", "dpinit"):
            assert _same(o.download(n), h.t[n]), n
        h.t["cumuflx"][...] = case["cumuflx"]
        o.upload("cumuflx", case["cumuflx"])
        h.t["dpfinl"][...] = case["dpinit"]
        o.upload("dp", np.stack([case["dpfinl"], case["dpinit"]]))
        h.transport()
        o.step_transp1()
        o.step_transp2(2)
        for n in TR_OUT + ("cumuflx",):
            assert _same(o.download(n), h.t[n]), ("second", n)
    finally:
        o.close()


def test_transport_g7_full_size(oracle_lib):
    print("case: G7 x 32, transport resident, numtr 1", flush=True)
    g = _grid(7)
    case = _tr_case(g, 32, 1, 71, None)
    o = _tr_ocean(g, case)
    try:
        for n in ("temp", "saln", "passv", "cumuflx", "dpinit"):
            o.upload(n, case[n])
        o.upload("dp", np.stack([case["dpinit"], case["dpfinl"]]))
        o.upload("massflx", case["massflx_tr"])
        h = oracle_lib.HycomCnuity(g, case)
        h.transport()
        o.step_transp1()
        o.step_transp2(2)
        for n in TR_OUT:
            assert _same(o.download(n), h.t[n]), n
    finally:
        o.close()


def test_transport_bad_arguments_are_rejected():
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip
    g = _grid(2)
    case = _tr_case(g, 32, 1, 5, None)
    o = fimhip.Ocean(g, 32, case["wet"], case["dtleap"])
    try:
        with pytest.raises(fimhip.FimhipError):
            o.step_transp1()                 # before set_transport
        o.set_transport(1, 0.02)
        o.step_transp0(1)                    # transp0 needs no batrop
        with pytest.raises(fimhip.FimhipError):
            o.step_transp1()                 # transp1 does
        o.set_momtum(case["batrop"])
        with pytest.raises(fimhip.FimhipError):
            o.set_transport(9, 0.02)
        with pytest.raises(fimhip.FimhipError):
            o.set_transport(2, 0.02)         # numtr is fixed once set
        o.set_transport(1, 0.0)
        with pytest.raises(fimhip.FimhipError):
            o.step_transp2(3)
        with pytest.raises(fimhip.FimhipError):
            o.step_transp0(0)
    finally:
        o.close()


def _hyb_ocean(g, case, numtr=0):
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip
    o = fimhip.Ocean(g, case["kdm"], case["wet"], 1800.0)
    o.set_momtum(case["batrop"])
    if numtr:
        o.set_transport(numtr, 0.02)
    o.set_hybgen(case["targt"], case["salmin"], case["dpmin"], case["latij"])
    return o


@pytest.mark.parametrize("leap,with_tracers", [(1, False), (2, True)])
def test_hybgen_dropin_matches_checker_bitwise_g3(oracle_lib, leap, with_tracers):
    """hybgn1 through the host-array drop-in, as hycom_run calls it (no
    tracers) and with two passive tracers."""
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    print("case: G3 x 32, hybgn1 drop-in, leap %d, tracers %s" % (leap, with_tracers), flush=True)
    g = _grid(3)
    case = ocean.hycom_hybgen_case(g, kdm=32, numtr=2)
    want = oracle_lib.hycom_hybgn1(case, leap, with_tracers=with_tracers)
    o = _hyb_ocean(g, case, numtr=2 if with_tracers else 0)
    try:
        a = {n: np.ascontiguousarray(case[n], np.float64).copy() for n in oracle_lib.HYB_FIELDS}
        o.hybgn1(1, leap, case["targt"], a["dens"], a["dp"], a["pres"], a["temp"], a["saln"], a["uvel"], a["vvel"],
                 case["latij"], a["passv"] if with_tracers else None)
        for n in oracle_lib.HYB_FIELDS:
            assert _same(a[n], want[n]), n
    finally:
        o.close()


@pytest.mark.parametrize("level,kdm,numtr", [(4, 48, 1), (5, 32, 0)])
def test_hybgen_resident_matches_checker(oracle_lib, level, kdm, numtr):
    """hybgn1 on the device-resident arrays at 32 layers and at 48 (the
    64-layer build), twice in a row on alternating levels."""
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    print("case: G%d x %d, hybgn1 resident, numtr %d" % (level, kdm, numtr), flush=True)
    g = _grid(level)
    case = ocean.hycom_hybgen_case(g, kdm=kdm, seed=61 + level, numtr=max(numtr, 1))
    o = _hyb_ocean(g, case, numtr=numtr)
    try:
        o.upload("dens", case["dens"])
        o.upload("dp", case["dp"])
        o.upload("pres", case["pres"])
        o.upload("temp", case["temp"])
        o.upload("saln", case["saln"])
        o.upload("uclin", case["uvel"])
        o.upload("vclin", case["vvel"])
        if numtr:
            o.upload("passv", case["passv"])
        c = dict(case)
        for leap in (2, 1):
            want = oracle_lib.hycom_hybgn1(c, leap, with_tracers=bool(numtr))
            o.step_hybgn1(leap, with_tracers=bool(numtr))
            got = {n: o.download(m) for n, m in (("dens", "dens"), ("dp", "dp"), ("pres", "pres"), ("temp", "temp"),
                                                  ("saln", "saln"), ("uvel", "uclin"), ("vvel", "vclin"))}
            if numtr:
                got["passv"] = o.download("passv")
            for n, v in got.items():
                assert _same(v, want[n]), (leap, n)
            c = dict(c, **want)
    finally:
        o.close()


def test_hybgen_g7_full_size(oracle_lib):
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    print("case: G7 x 32, hybgn1 resident", flush=True)
    g = _grid(7)
    case = ocean.hycom_hybgen_case(g, kdm=32)
    want = oracle_lib.hycom_hybgn1(case, 1)
    o = _hyb_ocean(g, case)
    try:
        for n, m in (("dens", "dens"), ("dp", "dp"), ("pres", "pres"), ("temp", "temp"), ("saln", "saln"),
                     ("uvel", "uclin"), ("vvel", "vclin")):
            o.upload(m, case[n])
        o.step_hybgn1(1)
        for n, m in (("dens", "dens"), ("dp", "dp"), ("pres", "pres"), ("temp", "temp"), ("saln", "saln"),
                     ("uvel", "uclin"), ("vvel", "vclin")):
            assert _same(o.download(m), want[n]), n
    finally:
        o.close()


def test_hybgen_bad_arguments_are_rejected():
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    g = _grid(2)
    case = ocean.hycom_hybgen_case(g, kdm=32)
    o = fimhip.Ocean(g, 32, case["wet"], 1800.0)
    try:
        with pytest.raises(fimhip.FimhipError):
            o.step_hybgn1(1)                  # before set_hybgen
        o.set_hybgen(case["targt"], case["salmin"], case["dpmin"], case["latij"])
        with pytest.raises(fimhip.FimhipError):
            o.step_hybgn1(1)                  # no batrop yet
        o.set_momtum(case["batrop"])
        with pytest.raises(fimhip.FimhipError):
            o.step_hybgn1(3)
        with pytest.raises(fimhip.FimhipError):
            o.step_hybgn1(1, with_tracers=True)   # no tracers set
        with pytest.raises(ValueError):
            o.set_hybgen(case["targt"][:5], case["salmin"], case["dpmin"], case["latij"])
    finally:
        o.close()


def _interval(oracle_lib, level, kdm, nsteps, numtr=1):
    """hycom_run's dynamics over nsteps steps (hycom_run.F90:364-556):
    bclin_edgvar, cnuity and momtum every step, transp1 every second step,
    transp2 (dpfinl = dp's new level), hybgn1 and transp0 every 30th; the
    checker on shared arrays beside the device-resident chain."""
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, ocean
    g = _grid(level)
    case = ocean.hycom_ocean_case(g, kdm=kdm, seed=73 + level, numtr=numtr)
    h = oracle_lib.HycomCnuity(g, case)
    o = fimhip.Ocean(g, kdm, case["wet"], case["dtleap"])
    o.set_momtum(case["batrop"], case["veldff"], case["smagcf"])
    o.set_transport(numtr, case["temdff"])
    o.set_hybgen(case["targt"], case["salmin"], case["dpmin"], case["latij"])
    for n in fimhip.OCN_MOMTUM_FIELDS + ["temp", "saln", "dens", "passv", "cumuflx", "dpinit"]:
        o.upload(n, case[n])
    oracle_lib.hycom_interval_op(h, "transp0", 1)
    o.step_transp0(1)
    for st in range(1, nsteps + 1):
        leapm = 1 + st % 2
        leapn = 3 - leapm
        h.edgvar(g, leapm, leapn, False)
        assert h.step(st, leapm, leapn) == 0
        h.momtum(g, leapm, leapn, False)
        o.step_edgvar(leapm, leapn)
        o.step(st, leapm, leapn)
        o.step_momtum(leapm, leapn)
        if st % 2 == 0:
            oracle_lib.hycom_interval_op(h, "transp1", leapn)
            o.step_transp1()
        if st % 30 == 0:
            for op in ("transp2", "hybgn1", "transp0"):
                oracle_lib.hycom_interval_op(h, op, leapn)
            o.step_transp2(leapn)
            o.step_hybgn1(leapn)
            o.step_transp0(leapn)
    return h, o


@pytest.mark.parametrize("level,kdm", [(4, 32), (3, 48), (3, 64), (3, 16)])
def test_ocean_interval_resident_matches_checker(oracle_lib, level, kdm):
    """The whole interval at 32 layers (two columns per wavefront), 48 and
    the maximum 64 (one column per wavefront; at 64 the bottom interface has
    no lane of its own), and 16 (half of each column's lanes idle)."""
    print("case: G%d x %d, 31 ocean steps (edgvar, cnuity, momtum, transport, hybgn1) resident" % (level, kdm),
          flush=True)
    h, o = _interval(oracle_lib, level, kdm, 31)
    try:
        for n in ARRAYS:
            assert _same(o.download(n), h.a[n]), n
        for n in ("uclin", "vclin", "ustarb"):
            assert _same(o.download(n), h.e[n]), n
        for n in ("temp", "saln", "dens", "passv", "cumuflx", "dpinit"):
            assert _same(o.download(n), h.o[n]), n
        assert o.mass == h.mass
    finally:
        o.close()

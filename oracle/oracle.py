"""ctypes driver for the C restatement in fim_oracle.c (the checker).

Test infrastructure only.  ``OracleRun`` owns a full FP64 model state in the
reference layout (level fastest) and advances it with the C restatement.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
CSRC = os.path.join(os.path.dirname(HERE), "fimocn_subx_realtime_hera_apr2020_amd", "csrc")


def build(force: bool = False) -> str:
    srcs = [os.path.join(HERE, "fim_oracle.c"), os.path.join(HERE, "hycom_oracle.c")]
    newest = max(os.path.getmtime(s) for s in srcs)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        cmd = ["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
               "-shared", "-I" + HERE, "-I" + CSRC, *srcs, "-o", LIB + ".tmp", "-lm"]
        subprocess.check_call(cmd)
        os.replace(LIB + ".tmp", LIB)
    return LIB


_PD = ct.POINTER(ct.c_double)
_PI = ct.POINTER(ct.c_int)


class _Grid(ct.Structure):
    _fields_ = [("nip", ct.c_int), ("nvl", ct.c_int), ("ntra", ct.c_int), ("pow_kind", ct.c_int),
                ("flux_scheme", ct.c_int), ("ntrb", ct.c_int), ("dtratio", ct.c_int), ("dt", ct.c_double),
                ("prox", _PI), ("proxs", _PI), ("nprox", _PI),
                ("cs", _PD), ("sn", _PD), ("sidevec_c", _PD), ("sidevec_e", _PD), ("sideln", _PD),
                ("hfcoef", _PD), ("area", _PD), ("rarea", _PD), ("corio", _PD), ("grndspd", _PD),
                ("dfflen", _PD), ("smoo_coeff", _PD), ("thetac", _PD), ("dpsig", _PD), ("wet", _PI)]


_STATE_ARRAYS = ["us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "trdp",
                 "u_tdcy", "v_tdcy", "dp_tdcy", "dpl_tdcy", "trc_tdcy", "trl_tdcy", "massfx",
                 "ws", "relvor", "sdot", "psrf", "ptdcy", "psurf",
                 "u_edg", "v_edg", "dp_edg", "lp_edg", "mont_edg", "uvsq_edg", "trc_edg",
                 "work_e1", "work_e2", "work_c1", "work_c2", "work_c3", "work_c4", "work_c5",
                 "work_l1", "work_l2"]


class _State(ct.Structure):
    _fields_ = [(n, _PD) for n in _STATE_ARRAYS] + [
        ("nf", ct.c_int), ("of", ct.c_int), ("vof", ct.c_int),
        ("ab1", ct.c_double), ("ab2", ct.c_double), ("ab3", ct.c_double),
        ("trb", _PD), ("cumufx", _PD), ("dpinit", _PD)]


def _lib():
    lib = ct.CDLL(build())
    for fn in ("fo_edgvar1", "fo_edgvar3", "fo_cnuity", "fo_mass_adv", "fo_trcadv", "fo_momtum"):
        getattr(lib, fn).argtypes = [ct.POINTER(_Grid), ct.POINTER(_State)]
    lib.fo_hybgen.argtypes = [ct.POINTER(_Grid), ct.POINTER(_State), ct.c_int]
    lib.fo_hystat.argtypes = [ct.POINTER(_Grid), ct.POINTER(_State), ct.c_int]
    lib.fo_init.argtypes = [ct.POINTER(_Grid), ct.POINTER(_State), ct.c_int]
    for fn in ("fo_transp0", "fo_transp1", "fo_transp2"):
        getattr(lib, fn).argtypes = [ct.POINTER(_Grid), ct.POINTER(_State)]
    lib.fo_step.argtypes = [ct.POINTER(_Grid), ct.POINTER(_State), ct.c_int, ct.c_int]
    lib.fo_abstart.argtypes = [ct.POINTER(_State), ct.c_int, ct.c_double]
    lib.fo_globsum.argtypes = [ct.POINTER(_Grid), ct.POINTER(_State), _PD, _PD]
    lib.fo_vertical_params.argtypes = [ct.c_int, ct.c_double, _PD, _PD]
    lib.fo_cell_params.argtypes = [ct.c_int, _PD, _PD, _PD, _PD, _PD]
    lib.fo_pow.argtypes = [ct.c_double, ct.c_double]
    lib.fo_pow.restype = ct.c_double
    lib.fo_pow_fast.argtypes = [ct.c_double, ct.c_double]
    lib.fo_pow_fast.restype = ct.c_double
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def _pd(a):
    return a.ctypes.data_as(_PD)


def _pi(a):
    return a.ctypes.data_as(_PI)


def grid_arrays(grid, nvl: int, dt: float, thetac, dpsig) -> dict:
    """FP64 / 0-based grid arrays in the oracle layout from a glvl.Grid."""
    nip = grid.nip
    prox = grid.dyn_prox().astype(np.int32) - 1
    prox[prox < 0] = 0
    proxs = grid.proxs.astype(np.int32) - 1
    proxs[proxs < 0] = 0
    d = dict(
        prox=np.ascontiguousarray(prox), proxs=np.ascontiguousarray(proxs),
        nprox=np.ascontiguousarray(grid.nprox.astype(np.int32)),
        cs=np.ascontiguousarray(grid.cs, np.float64), sn=np.ascontiguousarray(grid.sn, np.float64),
        sidevec_c=np.ascontiguousarray(grid.sidevec_c, np.float64),
        sidevec_e=np.ascontiguousarray(grid.sidevec_e, np.float64),
        sideln=np.ascontiguousarray(grid.sideln, np.float64),
        hfcoef=np.ascontiguousarray(grid.hfcoef, np.float64),
        area=np.ascontiguousarray(grid.area, np.float64),
        lat=np.ascontiguousarray(grid.lat, np.float64),
        thetac=np.ascontiguousarray(thetac, np.float64), dpsig=np.ascontiguousarray(dpsig, np.float64),
    )
    L = lib()
    d["rarea"] = np.empty(nip)
    d["corio"] = np.empty(nip)
    d["grndspd"] = np.empty(nip)
    L.fo_cell_params(nip, _pd(d["lat"]), _pd(d["area"]), _pd(d["corio"]), _pd(d["rarea"]), _pd(d["grndspd"]))
    d["dfflen"] = np.empty(nvl)
    d["smoo_coeff"] = np.empty(nvl)
    L.fo_vertical_params(nvl, dt, _pd(d["dfflen"]), _pd(d["smoo_coeff"]))
    return d


class OracleRun:
    """A full FP64 state advanced by the C restatement."""

    def __init__(self, garr: dict, nvl: int, ntra: int, dt: float, pow_kind: int = 1, flux_scheme: int = 1,
                 ntrb: int = 0, dtratio: int = 6):
        self.garr = garr
        nip = garr["nprox"].shape[0]
        self.nip, self.nvl, self.ntra, self.dt = nip, nvl, ntra, dt
        g = _Grid()
        g.nip, g.nvl, g.ntra, g.pow_kind, g.dt = nip, nvl, ntra, pow_kind, dt
        g.flux_scheme = flux_scheme
        g.ntrb, g.dtratio = ntrb, dtratio
        self.ntrb = ntrb
        for n in ("prox", "proxs", "nprox"):
            setattr(g, n, _pi(garr[n]))
        for n in ("cs", "sn", "sidevec_c", "sidevec_e", "sideln", "hfcoef", "area", "rarea", "corio",
                  "grndspd", "dfflen", "smoo_coeff", "thetac", "dpsig"):
            setattr(g, n, _pd(garr[n]))
        self.g = g
        c, l, e = (nip, nvl), (nip, nvl + 1), (nip, 6, nvl)
        shapes = dict(us=c, vs=c, dp=c, pr=l, ex=l, ph=l, mp=c, tr=(ntra,) + c, trdp=(ntra,) + c,
                      u_tdcy=(3,) + c, v_tdcy=(3,) + c, dp_tdcy=(3,) + c, dpl_tdcy=(3,) + c,
                      trc_tdcy=(ntra, 3) + c, trl_tdcy=(ntra, 3) + c, massfx=(3,) + e,
                      ws=c, relvor=c, sdot=l, psrf=(nip,), ptdcy=(2, nip), psurf=(nip,),
                      u_edg=e, v_edg=e, dp_edg=e, lp_edg=e, mont_edg=e, uvsq_edg=e, trc_edg=(ntra,) + e,
                      work_e1=e, work_e2=e, work_c1=c, work_c2=c, work_c3=c, work_c4=c, work_c5=c,
                      work_l1=l, work_l2=l)
        shapes.update(trb=(max(ntrb, 1),) + c, cumufx=e, dpinit=c)
        self.a = {n: np.zeros(shapes[n]) for n in _STATE_ARRAYS + ["trb", "cumufx", "dpinit"]}
        s = _State()
        for n in _STATE_ARRAYS + ["trb", "cumufx", "dpinit"]:
            setattr(s, n, _pd(self.a[n]))
        self.s = s

    def set_wet(self, wet):
        """Land/sea mask (nip values, 0 = land), or None for every cell wet
        (the reference's behaviour)."""
        if wet is None:
            self._wet = None
            self.g.wet = None
            return
        self._wet = np.ascontiguousarray(np.asarray(wet) != 0, np.int32)
        assert self._wet.shape == (self.nip,)
        self.g.wet = _pi(self._wet)

    def load(self, st: dict):
        for n in ("us", "vs", "dp", "pr", "ex", "ph", "tr"):
            self.a[n][...] = st[n]
        if self.ntrb and "trb" in st:
            self.a["trb"][...] = st["trb"]

    def init(self, init_hybgen: bool = True):
        lib().fo_init(ct.byref(self.g), ct.byref(self.s), 1 if init_hybgen else 0)

    def step(self, its: int, mode: int = 0):
        lib().fo_step(ct.byref(self.g), ct.byref(self.s), its, mode)

    def call(self, name: str, *args):
        getattr(lib(), "fo_" + name)(ct.byref(self.g), ct.byref(self.s), *args)

    def abstart(self, its: int):
        lib().fo_abstart(ct.byref(self.s), its, self.dt)

    def globsum(self):
        m = np.zeros(1)
        t = np.zeros(self.ntra + self.ntrb)
        lib().fo_globsum(ct.byref(self.g), ct.byref(self.s), _pd(m), _pd(t))
        return float(m[0]), t

    @property
    def slots(self):
        return self.s.nf, self.s.of, self.s.vof


# ---- HYCOM ocean continuity equation (hycom_oracle.c) -----------------------
class _HycGrid(ct.Structure):
    _fields_ = [("nip", ct.c_int), ("kdm", ct.c_int), ("prox", _PI), ("proxs", _PI), ("nprox", _PI), ("wet", _PI),
                ("area", _PD), ("rarea", _PD), ("sidevec_e", _PD), ("sideln", _PD), ("rsideln", _PD),
                ("rprox_ln", _PD), ("dtleap", ct.c_double), ("thkdff", ct.c_double), ("biharm", ct.c_double)]


class _HycMass(ct.Structure):
    _fields_ = [("totmass", ct.c_double), ("massglb0", ct.c_double), ("masscor", ct.c_double)]


class _HycMom(ct.Structure):
    _fields_ = [("cs", _PD), ("sn", _PD), ("sidevec_c", _PD), ("lat", _PD), ("batrop", ct.c_double),
                ("veldff", ct.c_double), ("smagcf", ct.c_double), ("janjic", ct.c_int)]


class HycomCnuity:
    """The checker's HYCOM cnuity on a glvl.Grid: arrays as
    ocean.hycom_cnuity_case makes them (C order of the Fortran layouts),
    updated in place by step(); mass holds totmass, massglb0, masscor."""

    def __init__(self, grid, case: dict, thkdff: float = 0.05, biharm: float = 0.75, mass=(0.0, 0.0, 0.0)):
        build()
        self._keep = {}
        nip = grid.nip
        prox = grid.dyn_prox().astype(np.int32) - 1
        prox[prox < 0] = 0
        proxs = grid.proxs.astype(np.int32) - 1
        proxs[proxs < 0] = 0
        area = np.asarray(grid.area, np.float64)
        sideln = np.ascontiguousarray(grid.sideln, np.float64)
        arr = dict(prox=np.ascontiguousarray(prox), proxs=np.ascontiguousarray(proxs),
                   nprox=np.ascontiguousarray(grid.nprox, np.int32), wet=np.ascontiguousarray(case["wet"], np.int32),
                   area=np.ascontiguousarray(area), rarea=np.ascontiguousarray(1.0 / area),
                   sidevec_e=np.ascontiguousarray(grid.sidevec_e, np.float64), sideln=sideln,
                   rsideln=np.ascontiguousarray(np.where(sideln != 0, 1.0 / np.where(sideln != 0, sideln, 1), 0.0)),
                   rprox_ln=np.ascontiguousarray(grid.rprox_ln, np.float64))
        self._keep.update(arr)
        g = _HycGrid()
        g.nip, g.kdm = nip, case["kdm"]
        for n in ("prox", "proxs", "nprox", "wet"):
            setattr(g, n, arr[n].ctypes.data_as(_PI))
        for n in ("area", "rarea", "sidevec_e", "sideln", "rsideln", "rprox_ln"):
            setattr(g, n, arr[n].ctypes.data_as(_PD))
        g.dtleap, g.thkdff, g.biharm = case["dtleap"], thkdff, biharm
        self.g = g
        self.m = _HycMass(*mass)
        self.case = case
        self.a = {n: np.ascontiguousarray(case[n], np.float64).copy()
                  for n in ("dp", "pres", "um_edg", "vm_edg", "un_edg", "vn_edg", "dpm_edg", "massflx", "mssfx_avg")}

    def step(self, nstep: int, leapm: int, leapn: int) -> int:
        a = self.a
        f = lib().fo_hycom_cnuity
        return f(ct.byref(self.g), ct.byref(self.m), nstep, leapm, leapn,
                 *[a[n].ctypes.data_as(_PD) for n in ("um_edg", "vm_edg", "un_edg", "vn_edg", "dpm_edg", "dp", "pres",
                                                     "massflx", "mssfx_avg")])

    @property
    def mass(self):
        return self.m.totmass, self.m.massglb0, self.m.masscor

    def momtum(self, grid, leapm: int, leapn: int, janjic: bool = False):
        """momtum (hycom_momtum.F90:16) on this instance's arrays (after
        edgvar, which leaves its edge values in self.a / self.e): uclin,
        vclin (self.e), pres (self.a) and ustarb (self.e) are updated; the
        case supplies taux, tauy, ustarb, batrop, veldff and smagcf."""
        a, e, case = self.a, self.e, self.case
        for n in ("taux", "tauy", "ustarb"):
            if n not in e:
                e[n] = np.ascontiguousarray(case[n], np.float64).copy()
        if "sidevec_c" not in self._keep:
            self._keep["sidevec_c"] = np.ascontiguousarray(grid.sidevec_c, np.float64)
            self._keep["lat"] = np.ascontiguousarray(grid.lat, np.float64)
        mm = _HycMom()
        for n in ("cs", "sn", "sidevec_c", "lat"):
            setattr(mm, n, self._keep[n].ctypes.data_as(_PD))
        mm.batrop, mm.veldff, mm.smagcf, mm.janjic = case["batrop"], case["veldff"], case["smagcf"], int(janjic)
        lib().fo_hycom_momtum(ct.byref(self.g), ct.byref(mm), leapm, leapn,
                              *[e[n].ctypes.data_as(_PD) for n in ("utrop", "vtrop", "uclin", "vclin")],
                              a["um_edg"].ctypes.data_as(_PD), a["vm_edg"].ctypes.data_as(_PD),
                              *[e[n].ctypes.data_as(_PD) for n in ("geop_edg", "mont_edg", "uvsq_edg", "spcv_edg",
                                                                   "taux", "tauy")],
                              a["dp"].ctypes.data_as(_PD), a["pres"].ctypes.data_as(_PD),
                              e["ustarb"].ctypes.data_as(_PD))

    def transport(self, leapn: int = None, transp0: bool = False, transp1_calls: int = 1):
        """HYCOM's tracer transport on copies of the case's arrays in self.t:
        transp0 (with leapn) when asked, otherwise transp1 with massflx_tr on
        cumuflx (transp1_calls times) and then transp2 on temp, saln and
        passv."""
        case = self.case
        if not hasattr(self, "t"):
            self.t = {n: np.ascontiguousarray(case[n], np.float64).copy()
                      for n in ("cumuflx", "massflx_tr", "dpinit", "dpfinl", "temp", "saln", "passv")}
        t = self.t
        P = lambda n: t[n].ctypes.data_as(_PD)
        if transp0:
            lib().fo_hycom_transp0(ct.byref(self.g), leapn, self.a["dp"].ctypes.data_as(_PD), P("dpinit"),
                                   P("cumuflx"))
            return
        for _ in range(transp1_calls):
            lib().fo_hycom_transp1(ct.byref(self.g), ct.c_double(case["batrop"]), P("massflx_tr"), P("cumuflx"))
        lib().fo_hycom_transp2(ct.byref(self.g), case["numtr"], ct.c_double(case["batrop"]),
                               ct.c_double(case["temdff"]), P("temp"), P("saln"), P("passv"), P("cumuflx"),
                               P("dpinit"), P("dpfinl"))

    EDG_IN = ("uclin", "vclin", "utrop", "vtrop", "geop", "montg", "spcifv")
    EDG_OUT = ("pr_edg", "geop_edg", "mont_edg", "uvsq_edg", "spcv_edg")

    def edgvar(self, grid, leapm: int, leapn: int, janjic: bool = False):
        """bclin_edgvar (hycom_edgvar.F90:128) on this instance's arrays: the
        case must hold ocean.hycom_edgvar_case's inputs; um/vm/un/vn_edg and
        dp_edg (as dpm_edg, which cnuity then reads) land in self.a, pr_edg,
        geop_edg, mont_edg, uvsq_edg and spcv_edg in self.e."""
        a = self.a
        if not hasattr(self, "e"):
            self.e = {n: np.ascontiguousarray(self.case[n], np.float64).copy() for n in self.EDG_IN + self.EDG_OUT}
            self._keep["cs"] = np.ascontiguousarray(grid.cs, np.float64)
            self._keep["sn"] = np.ascontiguousarray(grid.sn, np.float64)
        e = self.e
        lib().fo_hycom_edgvar(ct.byref(self.g), self._keep["cs"].ctypes.data_as(_PD),
                              self._keep["sn"].ctypes.data_as(_PD), int(janjic), leapm, leapn,
                              *[e[n].ctypes.data_as(_PD) for n in ("uclin", "vclin", "utrop", "vtrop")],
                              a["dp"].ctypes.data_as(_PD), a["pres"].ctypes.data_as(_PD),
                              *[e[n].ctypes.data_as(_PD) for n in ("geop", "montg", "spcifv")],
                              *[a[n].ctypes.data_as(_PD) for n in ("um_edg", "vm_edg", "un_edg", "vn_edg", "dpm_edg")],
                              *[e[n].ctypes.data_as(_PD) for n in self.EDG_OUT])


# ---- HYCOM hybrid vertical coordinate (hycom_oracle.c fo_hycom_hybgn1) ------
HYB_FIELDS = ("dens", "dp", "pres", "temp", "saln", "uvel", "vvel", "passv")


def hycom_hybgn1(case: dict, leap: int, with_tracers: bool = False) -> dict:
    """The checker's hybgn1 (hycom_hybgn1.F90:9) on copies of the case's
    arrays (ocean.hycom_hybgen_case): returns them updated.  with_tracers
    moves passv as well (hycom_run.F90:516 passes no tracers)."""
    build()
    a = {n: np.ascontiguousarray(case[n], np.float64).copy() for n in HYB_FIELDS}
    kdm, nip = case["kdm"], case["wet"].shape[0]
    wet = np.ascontiguousarray(case["wet"], np.int32)
    tab = {n: np.ascontiguousarray(case[n], np.float64) for n in ("targt", "salmin", "dpmin", "latij")}
    P = lambda x: x.ctypes.data_as(_PD)
    numtr = case["passv"].shape[0] if with_tracers else 0
    lib().fo_hycom_hybgn1(ct.c_int(nip), ct.c_int(kdm), wet.ctypes.data_as(_PI), ct.c_int(leap), P(tab["targt"]),
                          P(tab["salmin"]), P(tab["dpmin"]), ct.c_double(case["batrop"]), P(tab["latij"]),
                          P(a["dens"]), P(a["dp"]), P(a["pres"]), P(a["temp"]), P(a["saln"]), P(a["uvel"]),
                          P(a["vvel"]), ct.c_int(numtr), P(a["passv"]) if numtr else None)
    return a


def _hyc_ocean_arrays(h):
    if not hasattr(h, "o"):
        h.o = {n: np.ascontiguousarray(h.case[n], np.float64).copy()
               for n in ("temp", "saln", "dens", "passv", "cumuflx", "dpinit")}
    return h.o


def hycom_interval_op(h, op: str, leapn: int):
    """One of hycom_run's interval operations on a HycomCnuity instance whose
    case is ocean.hycom_ocean_case, sharing its arrays the way hycom_run does:
    'transp1' (the step's massflx into cumuflx), 'transp2' (dpfinl = dp's
    level leapn), 'hybgn1' (dens, dp, pres, temp, saln, uclin, vclin; no
    tracers, hycom_run.F90:516) or 'transp0'.  The tracers, dens, cumuflx and
    dpinit live in h.o."""
    o, a, case = _hyc_ocean_arrays(h), h.a, h.case
    P = lambda x: x.ctypes.data_as(_PD)
    g = ct.byref(h.g)
    if op == "transp1":
        lib().fo_hycom_transp1(g, ct.c_double(case["batrop"]), P(a["massflx"]), P(o["cumuflx"]))
    elif op == "transp2":
        cz = case["kdm"] * case["wet"].shape[0]
        dpfinl = np.ascontiguousarray(a["dp"].reshape(2, cz)[leapn - 1])
        lib().fo_hycom_transp2(g, case["numtr"], ct.c_double(case["batrop"]), ct.c_double(case["temdff"]),
                               P(o["temp"]), P(o["saln"]), P(o["passv"]), P(o["cumuflx"]), P(o["dpinit"]), P(dpfinl))
    elif op == "transp0":
        lib().fo_hycom_transp0(g, leapn, P(a["dp"]), P(o["dpinit"]), P(o["cumuflx"]))
    elif op == "hybgn1":
        tab = {n: np.ascontiguousarray(case[n], np.float64) for n in ("targt", "salmin", "dpmin", "latij")}
        wet = np.ascontiguousarray(case["wet"], np.int32)
        lib().fo_hycom_hybgn1(ct.c_int(case["wet"].shape[0]), ct.c_int(case["kdm"]), wet.ctypes.data_as(_PI),
                              ct.c_int(leapn), P(tab["targt"]), P(tab["salmin"]), P(tab["dpmin"]),
                              ct.c_double(case["batrop"]), P(tab["latij"]), P(o["dens"]), P(a["dp"]), P(a["pres"]),
                              P(o["temp"]), P(o["saln"]), P(h.e["uclin"]), P(h.e["vclin"]), ct.c_int(0), None)
    else:
        raise ValueError(op)

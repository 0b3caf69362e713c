"""Host-side mirror of the reference dynamics interface over libfimhip.so.

The functions below keep the names, argument order and argument meaning of
the reference subroutines (FIMsrc/fim/horizontal: edgvar1.F90:27,
edgvar3.F90:25, cnuity.F90:23, trcadv.F90:19, momtum.F90:16, hybgen.F90:20,
hystat.F90:23, abstart.F90:14) and forward to the C ABI declared in
include/fimhip.h.  Arrays are numpy float64 arrays in the reference layout
with the level index fastest (Fortran (nvl,nip) == C-order (nip,nvl)); inout
arrays are updated in place, as Fortran does.  A non-zero status from the
library raises FimhipError.

``Instance`` owns one device instance; ``Model`` adds the device-resident
fast path (upload, dyn_init start-up, asynchronous steps, download).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import glvl as _glvl

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libfimhip.so")

FIELDS = ["us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "trdp", "ws", "relvor", "sdot", "psrf", "ptdcy",
          "u_tdcy", "v_tdcy", "dp_tdcy", "dpl_tdcy", "trc_tdcy", "trl_tdcy", "massfx", "u_edg", "v_edg",
          "dp_edg", "lp_edg", "mont_edg", "uvsq_edg", "trc_edg", "trb", "cumufx", "dpinit"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

# symbols declared in include/fimhip.h
EXPORTS = ["fimhip_create", "fimhip_destroy", "fimhip_abstart", "fimhip_set_ab", "fimhip_get_ab",
           "fimhip_edgvar1_r8",
           "fimhip_edgvar3_r8", "fimhip_cnuity_r8", "fimhip_mass_adv_r8", "fimhip_set_flux_scheme",
           "fimhip_set_class_b", "fimhip_transp0_r8", "fimhip_transp1_r8", "fimhip_transp2_r8",
           "fimhip_trcadv_r8", "fimhip_momtum_r8",
           "fimhip_hybgen_r8", "fimhip_hystat_r8", "fimhip_field_size", "fimhip_upload", "fimhip_download",
           "fimhip_init_state", "fimhip_step", "fimhip_sync", "fimhip_globsum", "fimhip_icos_grid",
           "fimhip_icos_grid_info",
           "fimhip_synth_state", "fimhip_synth_device", "fimhip_profile", "fimhip_profile_report",
           "fimhip_pow", "fimhip_version", "fimhip_ens_create", "fimhip_ens_destroy", "fimhip_ens_synth",
           "fimhip_ens_load_member", "fimhip_ens_init_state", "fimhip_ens_step", "fimhip_ens_use_graph",
           "fimhip_ens_set_flux_scheme", "fimhip_ens_set_class_b", "fimhip_ens_load_class_b", "fimhip_ens_sync", "fimhip_ens_field_size", "fimhip_ens_download",
           "fimhip_ens_globsum",
           "fimhip_ens_kernel_count", "fimhip_ens_kernel_name", "fimhip_ens_set_tile", "fimhip_ens_get_tile",
           "fimhip_ens_autotune", "fimhip_ens_profile", "fimhip_ens_profile_report",
           "fimhip_globsum_noalloc", "fimhip_ens_set_fused", "fimhip_ens_get_fused", "fimhip_ens_patches",
           "fimhip_ens_set_wet", "fimhip_ls_pseudo_inverse",
           "fimhip_ocn_create", "fimhip_ocn_destroy", "fimhip_ocn_cnuity_r8", "fimhip_ocn_cnuity",
           "fimhip_ocn_field_size", "fimhip_ocn_upload", "fimhip_ocn_download", "fimhip_ocn_mass",
           "fimhip_ocn_set_mass", "fimhip_ocn_set_dtleap", "fimhip_ocn_profile", "fimhip_ocn_profile_report",
           "fimhip_ocn_edgvar_r8", "fimhip_ocn_edgvar", "fimhip_ocn_set_janjic", "fimhip_ocn_momtum_r8",
           "fimhip_ocn_momtum", "fimhip_ocn_set_momtum", "fimhip_ocn_set_transport", "fimhip_ocn_transp0_r8",
           "fimhip_ocn_transp1_r8", "fimhip_ocn_transp2_r8", "fimhip_ocn_transp0", "fimhip_ocn_transp1",
           "fimhip_ocn_transp2", "fimhip_ocn_set_hybgen", "fimhip_ocn_hybgn1_r8", "fimhip_ocn_hybgn1",
           "fimhip_ocn_sync"]

# real*4 drop-ins: name -> (number of float arrays, trailing int argument)
R4_ENTRIES = {"edgvar1": (16, False), "edgvar3": (16, False), "cnuity": (13, True), "mass_adv": (12, True),
              "trcadv": (7, True), "momtum": (13, False), "hybgen": (8, True), "hystat": (8, False),
              "transp0": (3, False), "transp2": (4, True)}
EXPORTS += ["fimhip_%s_r4" % n for n in R4_ENTRIES] + ["fimhip_abstart_r4", "fimhip_transp1_r4"]

SYNTH_WIND_NOISE, SYNTH_PERTURB, SYNTH_LOGNORMAL = 1, 2, 4
GLOBSUM_NOALLOC_MAX_Q = 64        # FIMHIP_GLOBSUM_NOALLOC_MAX_Q: mass + tracer totals per call


class FimhipError(RuntimeError):
    pass


_PD = ct.POINTER(ct.c_double)
_PI = ct.POINTER(ct.c_int)


class GridDesc(ct.Structure):
    _fields_ = [("nip", ct.c_int), ("nvl", ct.c_int), ("ntra", ct.c_int), ("dt", ct.c_double),
                ("nprox", _PI), ("prox", _PI), ("proxs", _PI), ("cs", _PD), ("sn", _PD),
                ("sidevec_c", _PD), ("sidevec_e", _PD), ("sideln", _PD), ("hfcoef", _PD),
                ("area", _PD), ("lat", _PD), ("lon", _PD), ("thetac", _PD), ("dpsig", _PD)]


_LIB = None


def load_library() -> ct.CDLL:
    """Load libfimhip.so; raises if it has not been built (no CPU fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise FimhipError("libfimhip.so not built: run __graft_entry__.build() (hipcc, gfx950)")
    lib = ct.CDLL(LIBPATH)
    vp = ct.c_void_p
    lib.fimhip_create.argtypes = [ct.POINTER(GridDesc), ct.c_int, ct.POINTER(vp)]
    lib.fimhip_destroy.argtypes = [vp]
    lib.fimhip_abstart.argtypes = [vp, ct.c_int, ct.c_int, _PI, _PI, _PI, _PD, _PD, _PD]
    lib.fimhip_set_ab.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double]
    lib.fimhip_get_ab.argtypes = [vp, _PI, _PI, _PI, _PD, _PD, _PD]
    lib.fimhip_edgvar1_r8.argtypes = [vp, ct.c_int] + [_PD] * 16
    lib.fimhip_edgvar3_r8.argtypes = [vp, ct.c_int] + [_PD] * 16
    lib.fimhip_cnuity_r8.argtypes = [vp, ct.c_int] + [_PD] * 13 + [ct.c_int]
    lib.fimhip_mass_adv_r8.argtypes = [vp, ct.c_int] + [_PD] * 12 + [ct.c_int]
    lib.fimhip_set_flux_scheme.argtypes = [vp, ct.c_int]
    lib.fimhip_set_class_b.argtypes = [vp, ct.c_int, ct.c_int]
    _PF4 = ct.POINTER(ct.c_float)
    for n, (na, tail) in R4_ENTRIES.items():
        getattr(lib, "fimhip_%s_r4" % n).argtypes = [vp, ct.c_int] + [_PF4] * na + ([ct.c_int] if tail else [])
    lib.fimhip_abstart_r4.argtypes = [vp, ct.c_int, ct.c_int, _PI, _PI, _PI, _PF4, _PF4, _PF4]
    lib.fimhip_transp1_r4.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float, ct.c_float,
                                      ct.c_float, _PF4, _PF4]
    lib.fimhip_transp0_r8.argtypes = [vp, ct.c_int, _PD, _PD, _PD]
    lib.fimhip_transp1_r8.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                      ct.c_double, _PD, _PD]
    lib.fimhip_transp2_r8.argtypes = [vp, ct.c_int, _PD, _PD, _PD, _PD, ct.c_int]
    lib.fimhip_trcadv_r8.argtypes = [vp, ct.c_int] + [_PD] * 7 + [ct.c_int]
    lib.fimhip_momtum_r8.argtypes = [vp, ct.c_int] + [_PD] * 13
    lib.fimhip_hybgen_r8.argtypes = [vp, ct.c_int] + [_PD] * 8 + [ct.c_int]
    lib.fimhip_hystat_r8.argtypes = [vp, ct.c_int] + [_PD] * 8
    lib.fimhip_field_size.argtypes = [vp, ct.c_int]
    lib.fimhip_field_size.restype = ct.c_longlong
    lib.fimhip_upload.argtypes = [vp, ct.c_int, _PD]
    lib.fimhip_download.argtypes = [vp, ct.c_int, _PD]
    lib.fimhip_init_state.argtypes = [vp, ct.c_int]
    lib.fimhip_step.argtypes = [vp, ct.c_int, ct.c_int]
    lib.fimhip_sync.argtypes = [vp]
    lib.fimhip_globsum.argtypes = [vp, _PD, _PD]
    lib.fimhip_globsum_noalloc.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, vp, vp, vp, vp, vp]
    _PF = ct.POINTER(ct.c_float)
    lib.fimhip_icos_grid.argtypes = [ct.c_int, _PF, _PF, _PI, _PI, _PI, _PF, _PF, _PF, _PF, _PF, _PF, _PF,
                                     _PF, _PF, _PI]
    lib.fimhip_icos_grid_info.argtypes = [ct.c_int, _PF, _PF, _PI, _PI, _PF, _PI]
    lib.fimhip_synth_state.argtypes = [ct.c_int, ct.c_int, ct.c_int, _PD, _PD, ct.c_ulonglong, ct.c_int,
                                       ct.c_int] + [_PD] * 7
    lib.fimhip_synth_device.argtypes = [vp, ct.c_ulonglong, ct.c_int, ct.c_int]
    lib.fimhip_profile.argtypes = [vp, ct.c_int]
    lib.fimhip_profile_report.argtypes = [vp, ct.c_char_p, ct.c_int]
    lib.fimhip_pow.argtypes = [ct.c_double, ct.c_double]
    lib.fimhip_pow.restype = ct.c_double
    lib.fimhip_version.restype = ct.c_char_p
    lib.fimhip_ens_create.argtypes = [ct.POINTER(GridDesc), ct.c_int, ct.c_int, ct.POINTER(vp)]
    lib.fimhip_ens_destroy.argtypes = [vp]
    lib.fimhip_ens_synth.argtypes = [vp, ct.c_int, ct.c_ulonglong, ct.c_int, ct.c_int]
    lib.fimhip_ens_load_member.argtypes = [vp, ct.c_int] + [_PD] * 7
    lib.fimhip_ens_init_state.argtypes = [vp, ct.c_int]
    lib.fimhip_ens_step.argtypes = [vp, ct.c_int, ct.c_int]
    lib.fimhip_ens_use_graph.argtypes = [vp, ct.c_int]
    lib.fimhip_ens_set_fused.argtypes = [vp, ct.c_int]
    lib.fimhip_ens_get_fused.argtypes = [vp]
    lib.fimhip_ens_set_wet.argtypes = [vp, ct.c_int, _PI]
    lib.fimhip_ls_pseudo_inverse.argtypes = [ct.POINTER(ct.c_float), ct.c_int, ct.POINTER(ct.c_float)]
    lib.fimhip_ens_patches.argtypes = [ct.POINTER(GridDesc), _PI, _PI, ct.c_longlong, _PI, ct.c_longlong]
    lib.fimhip_ens_set_flux_scheme.argtypes = [vp, ct.c_int]
    lib.fimhip_ens_set_class_b.argtypes = [vp, ct.c_int, ct.c_int]
    lib.fimhip_ens_load_class_b.argtypes = [vp, ct.c_int, _PD]
    lib.fimhip_ens_sync.argtypes = [vp]
    lib.fimhip_ens_field_size.argtypes = [vp, ct.c_int]
    lib.fimhip_ens_field_size.restype = ct.c_longlong
    lib.fimhip_ens_download.argtypes = [vp, ct.c_int, ct.c_int, _PD]
    lib.fimhip_ens_globsum.argtypes = [vp, _PD, _PD, _PD, _PD]
    lib.fimhip_ens_kernel_name.argtypes = [ct.c_int]
    lib.fimhip_ens_kernel_name.restype = ct.c_char_p
    lib.fimhip_ens_set_tile.argtypes = [vp, ct.c_int, ct.c_int]
    lib.fimhip_ens_get_tile.argtypes = [vp, ct.c_int]
    lib.fimhip_ens_autotune.argtypes = [vp, ct.c_int, ct.c_int, _PD]
    lib.fimhip_ens_profile.argtypes = [vp, ct.c_int]
    lib.fimhip_ens_profile_report.argtypes = [vp, ct.c_char_p, ct.c_int]
    lib.fimhip_ocn_create.argtypes = [ct.POINTER(GridDesc), _PD, ct.c_int, _PI, ct.c_double, ct.c_double,
                                      ct.c_double, ct.c_int, ct.POINTER(vp)]
    lib.fimhip_ocn_destroy.argtypes = [vp]
    lib.fimhip_ocn_cnuity_r8.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int] + [_PD] * 9
    lib.fimhip_ocn_cnuity.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int]
    lib.fimhip_ocn_field_size.argtypes = [vp, ct.c_int]
    lib.fimhip_ocn_field_size.restype = ct.c_longlong
    lib.fimhip_ocn_upload.argtypes = [vp, ct.c_int, _PD]
    lib.fimhip_ocn_download.argtypes = [vp, ct.c_int, _PD]
    lib.fimhip_ocn_mass.argtypes = [vp, _PD, _PD, _PD]
    lib.fimhip_ocn_set_mass.argtypes = [vp, ct.c_double, ct.c_double, ct.c_double]
    lib.fimhip_ocn_set_dtleap.argtypes = [vp, ct.c_double]
    lib.fimhip_ocn_edgvar_r8.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int] + [_PD] * 19
    lib.fimhip_ocn_edgvar.argtypes = [vp, ct.c_int, ct.c_int]
    lib.fimhip_ocn_set_janjic.argtypes = [vp, ct.c_int]
    lib.fimhip_ocn_momtum_r8.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int] + [_PD] * 17
    lib.fimhip_ocn_momtum.argtypes = [vp, ct.c_int, ct.c_int]
    lib.fimhip_ocn_set_momtum.argtypes = [vp, ct.c_double, ct.c_double, ct.c_double]
    lib.fimhip_ocn_set_transport.argtypes = [vp, ct.c_int, ct.c_double]
    lib.fimhip_ocn_transp0_r8.argtypes = [vp, ct.c_int, ct.c_int, _PD, _PD, _PD]
    lib.fimhip_ocn_transp1_r8.argtypes = [vp, ct.c_int, _PD, _PD]
    lib.fimhip_ocn_transp2_r8.argtypes = [vp, ct.c_int] + [_PD] * 6
    lib.fimhip_ocn_transp0.argtypes = [vp, ct.c_int]
    lib.fimhip_ocn_transp1.argtypes = [vp]
    lib.fimhip_ocn_transp2.argtypes = [vp, ct.c_int]
    lib.fimhip_ocn_set_hybgen.argtypes = [vp, _PD, _PD, _PD, _PD]
    lib.fimhip_ocn_hybgn1_r8.argtypes = [vp, ct.c_int, ct.c_int] + [_PD] * 10
    lib.fimhip_ocn_hybgn1.argtypes = [vp, ct.c_int, ct.c_int]
    lib.fimhip_ocn_sync.argtypes = [vp]
    lib.fimhip_ocn_profile.argtypes = [vp, ct.c_int]
    lib.fimhip_ocn_profile_report.argtypes = [vp, ct.c_char_p, ct.c_int]
    _LIB = lib
    return lib


def _chk(rc: int, what: str):
    if rc != 0:
        raise FimhipError("%s failed with status %d" % (what, rc))


def _p(a: np.ndarray):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_PD)


def icos_grid(level: int) -> _glvl.Grid:
    """Generate the glvl.dat tables of an all-bisection grid (fimhip_icos_grid)."""
    lib = load_library()
    nip = _glvl.nip_of(level)
    f = lambda *shape: np.zeros(shape, np.float32)
    i = lambda *shape: np.zeros(shape, np.int32)
    a = dict(lat=f(nip), lon=f(nip), nprox=i(nip), prox=i(nip, 6), proxs=i(nip, 6), area=f(nip),
             cs=f(nip, 6, 4), sn=f(nip, 6, 4), sidevec_c=f(nip, 6, 2), sidevec_e=f(nip, 6, 2),
             sideln=f(nip, 6), rprox_ln=f(nip, 6), corner_xy=f(nip, 2, 6), hfcoef=f(nip, 6, 6),
             inv_perm=i(nip))
    order = ["lat", "lon", "nprox", "prox", "proxs", "area", "cs", "sn", "sidevec_c", "sidevec_e", "sideln",
             "rprox_ln", "corner_xy", "hfcoef", "inv_perm"]
    ptrs = [a[n].ctypes.data_as(_PI if a[n].dtype == np.int32 else ct.POINTER(ct.c_float)) for n in order]
    _chk(lib.fimhip_icos_grid(level, *ptrs), "fimhip_icos_grid")
    return _glvl.Grid(level, 0, **a)


def ls_pseudo_inverse(xy) -> np.ndarray:
    """getlvl's least-squares pseudo-inverse (fimhip_ls_pseudo_inverse) of
    nobs = 5 or 6 points xy (nobs, 2) float32: the (5, nobs) array inv."""
    lib = load_library()
    p = np.ascontiguousarray(xy, np.float32)
    assert p.ndim == 2 and p.shape[1] == 2
    inv = np.zeros((p.shape[0], 5), np.float32)       # row per observation: inv(ibf, iobs), ibf fastest
    _chk(lib.fimhip_ls_pseudo_inverse(p.ctypes.data_as(ct.POINTER(ct.c_float)), p.shape[0],
                                      inv.ctypes.data_as(ct.POINTER(ct.c_float))), "ls_pseudo_inverse")
    return np.ascontiguousarray(inv.T)


def icos_grid_info(level: int) -> _glvl.GridInfo:
    """Generate the contents of the ginfo file icos_grid_info_level.dat
    (fimhip_icos_grid_info)."""
    lib = load_library()
    nip = _glvl.nip_of(level)
    lat, lon = np.zeros(nip, np.float32), np.zeros(nip, np.float32)
    prox, nprox = np.zeros((nip, 6), np.int32), np.zeros(nip, np.int32)
    edge, inv = np.zeros((nip, 2, 2, 6), np.float32), np.zeros(nip, np.int32)
    fp = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_float))
    _chk(lib.fimhip_icos_grid_info(level, fp(lat), fp(lon), _p32(prox), _p32(nprox), fp(edge), _p32(inv)),
         "fimhip_icos_grid_info")
    return _glvl.GridInfo(level, 0, lat, lon, prox, nprox, edge, inv)


def _p32(a):
    return a.ctypes.data_as(_PI)


def synth_state(lat, lon, nvl: int, ntra: int, seed: int = 20240601, member: int = 0, flags: int = 0) -> dict:
    """One member's seeded synthetic state on the host (fimhip_synth_state);
    the device generator fimhip_synth_device writes the same bytes."""
    lib = load_library()
    lat = np.ascontiguousarray(lat, np.float64)
    lon = np.ascontiguousarray(lon, np.float64)
    nip = lat.shape[0]
    c, l = (nip, nvl), (nip, nvl + 1)
    st = dict(us=np.empty(c), vs=np.empty(c), dp=np.empty(c), pr=np.empty(l), ex=np.empty(l), ph=np.empty(l),
              tr=np.empty((ntra,) + c))
    _chk(lib.fimhip_synth_state(nip, nvl, ntra, _p(lat), _p(lon), seed, member, flags,
                                *[_p(st[n]) for n in ("us", "vs", "dp", "pr", "ex", "ph", "tr")]),
         "fimhip_synth_state")
    return st


def _grid_desc(grid: _glvl.Grid, nvl: int, ntra: int, dt: float, thetac, dpsig):
    """The fimhip_grid_desc of a grid, with the arrays it points to."""
    keep = dict(
        nprox=np.ascontiguousarray(grid.nprox, np.int32),
        prox=np.ascontiguousarray(grid.prox, np.int32),
        proxs=np.ascontiguousarray(grid.proxs, np.int32),
        cs=np.ascontiguousarray(grid.cs, np.float64), sn=np.ascontiguousarray(grid.sn, np.float64),
        sidevec_c=np.ascontiguousarray(grid.sidevec_c, np.float64),
        sidevec_e=np.ascontiguousarray(grid.sidevec_e, np.float64),
        sideln=np.ascontiguousarray(grid.sideln, np.float64),
        hfcoef=np.ascontiguousarray(grid.hfcoef, np.float64),
        area=np.ascontiguousarray(grid.area, np.float64),
        lat=np.ascontiguousarray(grid.lat, np.float64),
        lon=np.ascontiguousarray(grid.lon, np.float64),
        thetac=np.ascontiguousarray(thetac, np.float64), dpsig=np.ascontiguousarray(dpsig, np.float64))
    gd = GridDesc()
    gd.nip, gd.nvl, gd.ntra, gd.dt = grid.nip, nvl, ntra, dt
    for n in ("nprox", "prox", "proxs"):
        setattr(gd, n, keep[n].ctypes.data_as(_PI))
    for n in ("cs", "sn", "sidevec_c", "sidevec_e", "sideln", "hfcoef", "area", "lat", "lon", "thetac", "dpsig"):
        setattr(gd, n, keep[n].ctypes.data_as(_PD))
    return gd, keep


def ens_patches(grid: _glvl.Grid, nvl: int = 64):
    """The patch partition of the batched form's fused tracer transport:
    (cnt, ent) with cnt (npatch, 4) the cumulative ends of each patch's core
    and rings 1-3 in its entries, ent (nentries, 4) each entry's 0-based cell
    and packed local neighbour indices."""
    import numpy as _np
    lib = load_library()
    gd, keep = _grid_desc(grid, nvl, 1, 60.0, _np.zeros(nvl), _np.zeros(nvl))
    sz = (ct.c_int * 2)()
    _chk(lib.fimhip_ens_patches(ct.byref(gd), sz, None, 0, None, 0), "ens_patches")
    cnt = _np.zeros((sz[0], 4), _np.int32)
    ent = _np.zeros((sz[1], 4), _np.int32)
    _chk(lib.fimhip_ens_patches(ct.byref(gd), sz, cnt.ctypes.data_as(_PI), cnt.size, ent.ctypes.data_as(_PI),
                                ent.size), "ens_patches")
    return cnt, ent


class Instance:
    """One device instance of the dynamics on one GPU (fimhip_create)."""

    def __init__(self, grid: _glvl.Grid, nvl: int, ntra: int, dt: float, thetac, dpsig, device: int = 0):
        self.lib = load_library()
        self.grid, self.nvl, self.ntra, self.dt = grid, nvl, ntra, dt
        self.ntrb = 0
        self.nip = grid.nip
        gd, self._keep = _grid_desc(grid, nvl, ntra, dt, thetac, dpsig)
        h = ct.c_void_p()
        _chk(self.lib.fimhip_create(ct.byref(gd), device, ct.byref(h)), "fimhip_create")
        self.h = h
        self.nf, self.of, self.vof = 0, 2, 1
        self.adbash = [0.0, 0.0, 0.0]

    def close(self):
        if getattr(self, "h", None):
            self.lib.fimhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference subroutines --------------------------------------------
    def abstart(self, its: int, itsDFI: int = 1):
        nf, of, vof = ct.c_int(self.nf), ct.c_int(self.of), ct.c_int(self.vof)
        a = [ct.c_double(x) for x in self.adbash]
        _chk(self.lib.fimhip_abstart(self.h, its, itsDFI, ct.byref(nf), ct.byref(of), ct.byref(vof),
                                     ct.byref(a[0]), ct.byref(a[1]), ct.byref(a[2])), "abstart")
        self.nf, self.of, self.vof = nf.value, of.value, vof.value
        self.adbash = [x.value for x in a]

    def set_ab(self, nf: int, of: int, vof: int, adbash1: float, adbash2: float, adbash3: float):
        """Set the instance's AB slots (1-based) and weights (a restart)."""
        _chk(self.lib.fimhip_set_ab(self.h, nf, of, vof, adbash1, adbash2, adbash3), "set_ab")
        self.nf, self.of, self.vof = nf, of, vof
        self.adbash = [adbash1, adbash2, adbash3]

    def get_ab(self):
        """(nf, of, vof, adbash1, adbash2, adbash3) the instance holds now."""
        i = [ct.c_int() for _ in range(3)]
        d = [ct.c_double() for _ in range(3)]
        _chk(self.lib.fimhip_get_ab(self.h, *[ct.byref(x) for x in i + d]), "get_ab")
        return tuple(x.value for x in i + d)

    def edgvar1(self, its, u_vel, v_vel, delp, pres, exner, geop, montg, tracr, u_edg, v_edg, dp_edg,
                lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg):
        _chk(self.lib.fimhip_edgvar1_r8(self.h, its, *map(_p, (u_vel, v_vel, delp, pres, exner, geop, montg,
             tracr, u_edg, v_edg, dp_edg, lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg))), "edgvar1")

    def edgvar3(self, its, u_vel, v_vel, delp, pres, exner, geop, montg, tracr, u_edg, v_edg, dp_edg,
                lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg):
        _chk(self.lib.fimhip_edgvar3_r8(self.h, its, *map(_p, (u_vel, v_vel, delp, pres, exner, geop, montg,
             tracr, u_edg, v_edg, dp_edg, lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg))), "edgvar3")

    def cnuity(self, its, u_velo, v_velo, u_edg, v_edg, dp_edg, lp_edg, delp, pres, exner, dp_tndcy,
               dplo_tndcy, massfx, omega, TimingBarriers=False):
        _chk(self.lib.fimhip_cnuity_r8(self.h, its, *map(_p, (u_velo, v_velo, u_edg, v_edg, dp_edg, lp_edg,
             delp, pres, exner, dp_tndcy, dplo_tndcy, massfx, omega)), int(TimingBarriers)), "cnuity")

    def mass_adv(self, its, u_velo, v_velo, u_edg, v_edg, dp_edg, lp_edg, delp, pres, exner, dp_tndcy,
                 massfx, omega, TimingBarriers=False):
        _chk(self.lib.fimhip_mass_adv_r8(self.h, its, *map(_p, (u_velo, v_velo, u_edg, v_edg, dp_edg, lp_edg,
             delp, pres, exner, dp_tndcy, massfx, omega)), int(TimingBarriers)), "mass_adv")

    def set_flux_scheme(self, flux_conserv_schm: int):
        """1 = cnuity (FCT), 2 = mass_adv, for the steps that follow (dyn_run.F90:377-386)."""
        _chk(self.lib.fimhip_set_flux_scheme(self.h, int(flux_conserv_schm)), "set_flux_scheme")

    def set_class_b(self, ntrb: int, dtratio: int = 6):
        """Class-B tracers (transp3d.F90 / fct3d.F90) moved every dtratio steps."""
        _chk(self.lib.fimhip_set_class_b(self.h, int(ntrb), int(dtratio)), "set_class_b")
        self.ntrb = int(ntrb)

    def transp0(self, its, cumufx, dp3d, dpinit):
        _chk(self.lib.fimhip_transp0_r8(self.h, its, _p(cumufx), _p(dp3d), _p(dpinit)), "transp0")

    def transp1(self, its, nf, of, vof, adbash1, adbash2, adbash3, cumufx, massfx):
        _chk(self.lib.fimhip_transp1_r8(self.h, its, nf, of, vof, adbash1, adbash2, adbash3, _p(cumufx),
                                        _p(massfx)), "transp1")

    def transp2(self, its, tracr, cumufx, dpinit, dpfinl, TimingBarriers=False):
        """tracr: (ntra + ntrb, nip, nvl); the class-B fields are transported in place."""
        _chk(self.lib.fimhip_transp2_r8(self.h, its, _p(tracr), _p(cumufx), _p(dpinit), _p(dpfinl),
                                        int(TimingBarriers)), "transp2")

    def call_r4(self, name: str, its: int, *arrays, tail=None):
        """The real*4 drop-in fimhip_<name>_r4 on float32 arrays (updated in
        place), with the trailing integer argument where the routine has one."""
        na, has_tail = R4_ENTRIES[name]
        if len(arrays) != na:
            raise ValueError("%s_r4 takes %d arrays" % (name, na))
        ptrs = []
        for a in arrays:
            if a.dtype != np.float32 or not a.flags.c_contiguous:
                raise ValueError("%s_r4: arrays must be contiguous float32" % name)
            ptrs.append(a.ctypes.data_as(ct.POINTER(ct.c_float)))
        args = [self.h, its] + ptrs + ([int(tail or 0)] if has_tail else [])
        _chk(getattr(self.lib, "fimhip_%s_r4" % name)(*args), name + "_r4")

    def trcadv(self, its, trcr_edg, tracr, trcdp, trc_tdcy, trclo_tdcy, massfx, delp, TimingBarriers=False):
        _chk(self.lib.fimhip_trcadv_r8(self.h, its, *map(_p, (trcr_edg, tracr, trcdp, trc_tdcy, trclo_tdcy,
             massfx, delp)), int(TimingBarriers)), "trcadv")

    def momtum(self, its, u_velo, v_velo, exner, relvor, u_edg, v_edg, trcr_edg, geop_edg, mont_edg,
               uvsq_edg, u_tndcy, v_tndcy, dp3d):
        _chk(self.lib.fimhip_momtum_r8(self.h, its, *map(_p, (u_velo, v_velo, exner, relvor, u_edg, v_edg,
             trcr_edg, geop_edg, mont_edg, uvsq_edg, u_tndcy, v_tndcy, dp3d))), "momtum")

    def hybgen(self, its, targt, us3d, vs3d, tr3d, sdot, ex3d, dp3d, pr3d, TimingBarriers=False):
        _chk(self.lib.fimhip_hybgen_r8(self.h, its, *map(_p, (np.ascontiguousarray(targt, np.float64), us3d,
             vs3d, tr3d, sdot, ex3d, dp3d, pr3d)), int(TimingBarriers)), "hybgen")

    def hystat(self, its, ph3d, ex3d, mp3d, dp3d, tr3d, trdp, psrf, ptdcy):
        _chk(self.lib.fimhip_hystat_r8(self.h, its, *map(_p, (ph3d, ex3d, mp3d, dp3d, tr3d, trdp, psrf,
             ptdcy))), "hystat")


class Model(Instance):
    """Device-resident state advanced by fimhip_step."""

    def shape(self, name: str):
        nip, nvl, ntra = self.nip, self.nvl, self.ntra
        c, l, e = (nip, nvl), (nip, nvl + 1), (nip, 6, nvl)
        return dict(us=c, vs=c, dp=c, pr=l, ex=l, ph=l, mp=c, tr=(ntra,) + c, trdp=(ntra,) + c, ws=c,
                    relvor=c, sdot=l, psrf=(nip,), ptdcy=(2, nip), u_tdcy=(3,) + c, v_tdcy=(3,) + c,
                    dp_tdcy=(3,) + c, dpl_tdcy=(3,) + c, trc_tdcy=(ntra, 3) + c, trl_tdcy=(ntra, 3) + c,
                    massfx=(3,) + e, u_edg=e, v_edg=e, dp_edg=e, lp_edg=e, mont_edg=e, uvsq_edg=e,
                    trc_edg=(ntra,) + e, trb=(self.ntrb,) + c, cumufx=e, dpinit=c)[name]

    def upload(self, name: str, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.shape != self.shape(name):
            raise ValueError("%s: shape %s, expected %s" % (name, a.shape, self.shape(name)))
        _chk(self.lib.fimhip_upload(self.h, FIELD_ID[name], _p(a)), "upload " + name)

    def download(self, name: str) -> np.ndarray:
        a = np.empty(self.shape(name))
        _chk(self.lib.fimhip_download(self.h, FIELD_ID[name], _p(a)), "download " + name)
        return a

    def load(self, st: dict):
        for n in ("us", "vs", "dp", "pr", "ex", "ph", "tr"):
            self.upload(n, st[n])
        if self.ntrb and "trb" in st:
            self.upload("trb", st["trb"])

    def synth(self, seed: int = 20240601, member: int = 0, flags: int = 0):
        """Fill the device state with a seeded synthetic member (fimhip_synth_device)."""
        _chk(self.lib.fimhip_synth_device(self.h, seed, member, flags), "synth_device")

    def init_state(self, init_hybgen: bool = True):
        _chk(self.lib.fimhip_init_state(self.h, 1 if init_hybgen else 0), "init_state")

    def step(self, its: int, mode: int = 0):
        _chk(self.lib.fimhip_step(self.h, its, mode), "step")

    def sync(self):
        _chk(self.lib.fimhip_sync(self.h), "sync")

    def profile(self, on: bool = True):
        _chk(self.lib.fimhip_profile(self.h, 1 if on else 0), "profile")

    def profile_report(self) -> dict:
        """{kernel: (launches, total_ms)} for the steps run while profiling."""
        buf = ct.create_string_buffer(1 << 16)
        _chk(self.lib.fimhip_profile_report(self.h, buf, len(buf)), "profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    def globsum(self):
        m = ct.c_double()
        t = np.zeros(self.ntra + self.ntrb)
        _chk(self.lib.fimhip_globsum(self.h, ct.byref(m), _p(t)), "globsum")
        return m.value, t


def step_concurrently(instances, nsteps: int = 1, mode: int = 0):
    """nsteps steps of several Ensemble instances on one GPU (an uneven
    batch: grids or layer counts that cannot share a launch): each step
    enqueues every instance's captured step on its own stream before any
    wait, so the instances run side by side; returns when all are done."""
    for _ in range(nsteps):
        for e in instances:
            e.step(1, mode)
    for e in instances:
        e.sync()


class Ensemble:
    """Members advanced together in the batched mixed-precision form
    (fimhip_ens_*): FP32 winds, tracers and edge values, FP64 thickness,
    pressure, Exner function, Montgomery potential and hybgen."""

    FIELDS_OUT = ("us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "trdp", "ws", "relvor", "sdot", "psrf")

    def __init__(self, grid: _glvl.Grid, members: int, nvl: int, ntra: int, dt: float, thetac, dpsig,
                 device: int = 0):
        self.lib = load_library()
        self.grid, self.members, self.nvl, self.ntra, self.dt = grid, members, nvl, ntra, dt
        self.nip = grid.nip
        self.ntrb = 0
        gd, self._keep = _grid_desc(grid, nvl, ntra, dt, thetac, dpsig)
        h = ct.c_void_p()
        _chk(self.lib.fimhip_ens_create(ct.byref(gd), members, device, ct.byref(h)), "fimhip_ens_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.fimhip_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synth(self, member: int, seed: int = 20240601, synth_member: int | None = None, flags: int = 0):
        sm = member if synth_member is None else synth_member
        _chk(self.lib.fimhip_ens_synth(self.h, member, seed, sm, flags), "ens_synth")

    def load_member(self, member: int, st: dict):
        a = [np.ascontiguousarray(st[n], np.float64) for n in ("us", "vs", "dp", "pr", "ex", "ph", "tr")]
        _chk(self.lib.fimhip_ens_load_member(self.h, member, *map(_p, a)), "ens_load_member")

    def init_state(self, init_hybgen: bool = True):
        _chk(self.lib.fimhip_ens_init_state(self.h, 1 if init_hybgen else 0), "ens_init_state")

    def step(self, nsteps: int = 1, mode: int = 0):
        _chk(self.lib.fimhip_ens_step(self.h, nsteps, mode), "ens_step")

    def set_wet(self, member: int, wet):
        """Land/sea mask of one member (nip values, 0 = land)."""
        w = np.ascontiguousarray(wet, np.int32)
        assert w.shape == (self.nip,)
        _chk(self.lib.fimhip_ens_set_wet(self.h, member, w.ctypes.data_as(_PI)), "ens_set_wet")

    def set_fused(self, on: bool):
        """The tracer transport as one kernel over patches (True) or as one
        kernel per stage (False, the default); both give the same bits."""
        _chk(self.lib.fimhip_ens_set_fused(self.h, int(on)), "ens_set_fused")

    def fused(self) -> bool:
        """Whether the tracer transport runs as the fused kernel (the
        autotuner keeps the faster form)."""
        r = self.lib.fimhip_ens_get_fused(self.h)
        _chk(min(r, 0), "ens_get_fused")
        return r == 1

    def use_graph(self, on: bool):
        _chk(self.lib.fimhip_ens_use_graph(self.h, 1 if on else 0), "ens_use_graph")

    def set_flux_scheme(self, flux_conserv_schm: int):
        """1 = cnuity (FCT), 2 = mass_adv, for the steps that follow (dyn_run.F90:377-386)."""
        _chk(self.lib.fimhip_ens_set_flux_scheme(self.h, int(flux_conserv_schm)), "ens_set_flux_scheme")

    def set_class_b(self, ntrb: int, dtratio: int = 6):
        """Class-B tracers (transp3d.F90 / fct3d.F90) moved every dtratio steps."""
        _chk(self.lib.fimhip_ens_set_class_b(self.h, int(ntrb), int(dtratio)), "ens_set_class_b")
        self.ntrb = int(ntrb)

    def load_class_b(self, member: int, trb):
        a = np.ascontiguousarray(trb, np.float64)
        if a.shape != (self.ntrb, self.nip, self.nvl):
            raise ValueError("trb: shape %s, expected %s" % (a.shape, (self.ntrb, self.nip, self.nvl)))
        _chk(self.lib.fimhip_ens_load_class_b(self.h, member, _p(a)), "ens_load_class_b")

    def sync(self):
        _chk(self.lib.fimhip_ens_sync(self.h), "ens_sync")

    def shape(self, name: str):
        nip, nvl, ntra = self.nip, self.nvl, self.ntra
        c, l = (nip, nvl), (nip, nvl + 1)
        return dict(us=c, vs=c, dp=c, mp=c, ws=c, relvor=c, pr=l, ex=l, ph=l, sdot=l, tr=(ntra,) + c,
                    trdp=(ntra,) + c, psrf=(nip,), trb=(self.ntrb,) + c)[name]

    def download(self, member: int, name: str) -> np.ndarray:
        a = np.empty(self.shape(name))
        _chk(self.lib.fimhip_ens_download(self.h, member, FIELD_ID[name], _p(a)), "ens_download " + name)
        return a

    def globsum(self):
        """(mass[m], trtot[m, t], mass32[m], trtot32[m, t]): FP64 sums and FP32
        Kahan-compensated sums of every member."""
        m, t = np.zeros(self.members), np.zeros((self.members, self.ntra))
        m32, t32 = np.zeros(self.members), np.zeros((self.members, self.ntra))
        _chk(self.lib.fimhip_ens_globsum(self.h, _p(m), _p(t), _p(m32), _p(t32)), "ens_globsum")
        return m, t, m32, t32

    def kernel_names(self):
        return [self.lib.fimhip_ens_kernel_name(k).decode() for k in range(self.lib.fimhip_ens_kernel_count())]

    def tiles(self) -> dict:
        return {n: self.lib.fimhip_ens_get_tile(self.h, k) for k, n in enumerate(self.kernel_names())}

    def set_tiles(self, tiles: dict):
        for k, n in enumerate(self.kernel_names()):
            if n in tiles:
                _chk(self.lib.fimhip_ens_set_tile(self.h, k, int(tiles[n])), "ens_set_tile")

    def autotune(self, mode: int = 0, reps: int = 2) -> np.ndarray:
        """Pick each kernel's fastest launch shape (advances the state);
        returns the kernel x tile table of mean milliseconds."""
        nk = self.lib.fimhip_ens_kernel_count()
        t = np.full((nk, 8), np.nan)
        _chk(self.lib.fimhip_ens_autotune(self.h, mode, reps, _p(t)), "ens_autotune")
        return t

    def profile(self, on: bool = True):
        _chk(self.lib.fimhip_ens_profile(self.h, 1 if on else 0), "ens_profile")

    def profile_report(self) -> dict:
        buf = ct.create_string_buffer(1 << 16)
        _chk(self.lib.fimhip_ens_profile_report(self.h, buf, len(buf)), "ens_profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out


OCN_FIELDS = ["dp", "pres", "um_edg", "vm_edg", "un_edg", "vn_edg", "dpm_edg", "massflx", "mssfx_avg",
              "uclin", "vclin", "utrop", "vtrop", "geop", "montg", "spcifv", "pr_edg", "geop_edg", "mont_edg",
              "uvsq_edg", "spcv_edg", "taux", "tauy", "ustarb", "cumuflx", "dpinit", "dpfinl", "temp", "saln",
              "passv", "dens"]
OCN_CNUITY_FIELDS = OCN_FIELDS[:9]       # the arrays cnuity reads or writes
OCN_EDGVAR_FIELDS = OCN_FIELDS[:21]      # with those of bclin_edgvar
OCN_MOMTUM_FIELDS = OCN_FIELDS[:24]      # with those of momtum
OCN_TRANSP_FIELDS = OCN_FIELDS[24:30]    # the tracer transport's (passv after set_transport)
# bclin_edgvar's inputs and its outputs besides um/vm/un/vn_edg and dp_edg (dpm_edg here)
OCN_EDG_IN = ["uclin", "vclin", "utrop", "vtrop", "geop", "montg", "spcifv"]
OCN_EDG_OUT = ["pr_edg", "geop_edg", "mont_edg", "uvsq_edg", "spcv_edg"]
FIMHIP_EDUP = -4


class Ocean:
    """The HYCOM ocean's dynamics on one GPU (fimhip_ocn_*): the edge
    interpolation, continuity equation, momentum equation and tracer
    transport.  ``cnuity`` keeps the argument list of hycom_cnuity.F90:21
    (nstep, leapm, leapn, um_edg, vm_edg, un_edg, vn_edg, dpm_edg, dp, pres,
    massflx, mssfx_avg), ``edgvar`` that of bclin_edgvar
    (hycom_edgvar.F90:128), ``momtum`` that of hycom_momtum.F90:16 and
    ``transp0/1/2`` those of hycom_transp3d.F90:18, 54, 88, with numpy float64
    arrays in the reference layout (C order: dp, uclin, vclin (2,nip,kdm);
    pres, geop (nip,kdm+1); utrop, vtrop (nip,3); montg, spcifv, dpinit,
    dpfinl, temp, saln (nip,kdm); passv (numtr,nip,kdm); edges (nip,6,kdm),
    pr_edg (nip,6,kdm+1)), the written arrays updated in place.  ``mass`` is
    hycom_variables' (totmass, massglb0, masscor).  As the reference stops on
    differing duplicate edge fluxes (chkdup), so this raises FimhipError."""

    def __init__(self, grid: _glvl.Grid, kdm: int, wet, dtleap: float, thkdff: float = 0.05,
                 biharm: float = 0.75, device: int = 0):
        self.lib = load_library()
        self.grid, self.kdm, self.nip = grid, kdm, grid.nip
        self.numtr = 0
        gd, self._keep = _grid_desc(grid, 1, 1, dtleap, np.zeros(1), np.zeros(1))
        self._keep["rprox_ln"] = np.ascontiguousarray(grid.rprox_ln, np.float64)
        self._keep["wet"] = np.ascontiguousarray(wet, np.int32)
        assert self._keep["wet"].shape == (grid.nip,)
        h = ct.c_void_p()
        _chk(self.lib.fimhip_ocn_create(ct.byref(gd), _p(self._keep["rprox_ln"]), kdm,
                                        self._keep["wet"].ctypes.data_as(_PI), dtleap, thkdff, biharm, device,
                                        ct.byref(h)), "fimhip_ocn_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.fimhip_ocn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, name: str):
        n, k = self.nip, self.kdm
        return {"dp": (2, n, k), "pres": (n, k + 1), "uclin": (2, n, k), "vclin": (2, n, k), "utrop": (n, 3),
                "vtrop": (n, 3), "geop": (n, k + 1), "montg": (n, k), "spcifv": (n, k),
                "pr_edg": (n, 6, k + 1), "taux": (n,), "tauy": (n,), "ustarb": (n,), "dpinit": (n, k),
                "dpfinl": (n, k), "temp": (n, k), "saln": (n, k), "dens": (n, k),
                "passv": (self.numtr, n, k)}.get(name, (n, 6, k))

    @staticmethod
    def _step_rc(rc: int, what: str):
        if rc == FIMHIP_EDUP:
            raise FimhipError("%s: duplicate edge fluxes differ (chkdup)" % what)
        _chk(rc, what)

    def cnuity(self, nstep, leapm, leapn, um_edg, vm_edg, un_edg, vn_edg, dpm_edg, dp, pres, massflx, mssfx_avg):
        arrs = (um_edg, vm_edg, un_edg, vn_edg, dpm_edg, dp, pres, massflx, mssfx_avg)
        for n, a in zip(OCN_FIELDS[2:7] + OCN_FIELDS[:2] + OCN_FIELDS[7:], arrs):
            if a.shape != self._shape(n):
                raise ValueError("%s: shape %s, expected %s" % (n, a.shape, self._shape(n)))
        self._step_rc(self.lib.fimhip_ocn_cnuity_r8(self.h, nstep, leapm, leapn, *map(_p, arrs)), "ocn_cnuity_r8")

    def edgvar(self, nstep, leapm, leapn, uclin, vclin, utrop, vtrop, dp, pres, geop, montg, spcifv, um_edg,
               vm_edg, un_edg, vn_edg, dp_edg, pr_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg):
        arrs = (uclin, vclin, utrop, vtrop, dp, pres, geop, montg, spcifv, um_edg, vm_edg, un_edg, vn_edg, dp_edg,
                pr_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg)
        names = OCN_EDG_IN[:4] + ["dp", "pres"] + OCN_EDG_IN[4:] + ["um_edg", "vm_edg", "un_edg", "vn_edg",
                                                                     "dpm_edg"] + OCN_EDG_OUT
        for n, a in zip(names, arrs):
            if a.shape != self._shape(n):
                raise ValueError("%s: shape %s, expected %s" % (n, a.shape, self._shape(n)))
        _chk(self.lib.fimhip_ocn_edgvar_r8(self.h, nstep, leapm, leapn, *map(_p, arrs)), "ocn_edgvar_r8")

    def step_edgvar(self, leapm: int, leapn: int):
        """bclin_edgvar on the device-resident arrays: its edge values are the
        ones ``step`` (cnuity) then reads."""
        _chk(self.lib.fimhip_ocn_edgvar(self.h, leapm, leapn), "ocn_edgvar")

    def momtum(self, nstep, leapm, leapn, utrop, vtrop, uclin, vclin, um_edg, vm_edg, geop_edg, mont_edg,
               uvsq_edg, spcv_edg, taux, tauy, ubforc, vbforc, dp, pres, ustarb):
        """momtum (hycom_momtum.F90:16); ubforc and vbforc are not written
        (modesplit is off in the reference)."""
        arrs = (utrop, vtrop, uclin, vclin, um_edg, vm_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg, taux, tauy,
                ubforc, vbforc, dp, pres, ustarb)
        names = ["utrop", "vtrop", "uclin", "vclin", "um_edg", "vm_edg", "geop_edg", "mont_edg", "uvsq_edg",
                 "spcv_edg", "taux", "tauy", "taux", "tauy", "dp", "pres", "ustarb"]
        for n, a in zip(names, arrs):
            if a.shape != self._shape(n):
                raise ValueError("%s: shape %s, expected %s" % (n, a.shape, self._shape(n)))
        _chk(self.lib.fimhip_ocn_momtum_r8(self.h, nstep, leapm, leapn, *map(_p, arrs)), "ocn_momtum_r8")

    def step_momtum(self, leapm: int, leapn: int):
        """momtum on the device-resident arrays."""
        _chk(self.lib.fimhip_ocn_momtum(self.h, leapm, leapn), "ocn_momtum")

    def set_momtum(self, batrop: float, veldff: float = 0.1, smagcf: float = 0.3):
        """batrop (hycom_init.F90:46) and the namelist's veldff, smagcf."""
        _chk(self.lib.fimhip_ocn_set_momtum(self.h, float(batrop), float(veldff), float(smagcf)), "ocn_set_momtum")

    def set_janjic(self, on: bool):
        """janjic_ocn (fimnamelist.F90:246, default off)."""
        _chk(self.lib.fimhip_ocn_set_janjic(self.h, int(on)), "ocn_set_janjic")

    def set_transport(self, numtr: int = 1, temdff: float = 0.02):
        """numtr passive tracers (hycom_control.F90:3) and temdff
        (fimnamelist.F90:255); batrop comes from ``set_momtum``."""
        _chk(self.lib.fimhip_ocn_set_transport(self.h, int(numtr), float(temdff)), "ocn_set_transport")
        self.numtr = int(numtr)

    def _check(self, names, arrs):
        for n, a in zip(names, arrs):
            if a.shape != self._shape(n) or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("%s: shape %s %s, expected %s float64" % (n, a.shape, a.dtype, self._shape(n)))

    def transp0(self, nstep, leapn, cumuflx, dp, dpinit):
        """transp0 (hycom_transp3d.F90:18): cumuflx restarts, dpinit =
        dp[leapn-1] on the ocean."""
        self._check(["cumuflx", "dp", "dpinit"], (cumuflx, dp, dpinit))
        _chk(self.lib.fimhip_ocn_transp0_r8(self.h, nstep, leapn, _p(cumuflx), _p(dp), _p(dpinit)), "ocn_transp0_r8")

    def transp1(self, nstep, cumuflx, massflx):
        """transp1 (hycom_transp3d.F90:54): cumuflx += massflx*2*batrop."""
        self._check(["cumuflx", "massflx"], (cumuflx, massflx))
        _chk(self.lib.fimhip_ocn_transp1_r8(self.h, nstep, _p(cumuflx), _p(massflx)), "ocn_transp1_r8")

    def transp2(self, nstep, temp, saln, passv_tr, cumuflx, dpinit, dpfinl):
        """transp2 (hycom_transp3d.F90:88) with fct3d and dffusn_lyr: temp,
        saln and passv_tr advected in place (passv_tr None when numtr = 0)."""
        names = ["temp", "saln", "cumuflx", "dpinit", "dpfinl"] + (["passv"] if self.numtr else [])
        arrs = (temp, saln, cumuflx, dpinit, dpfinl) + ((passv_tr,) if self.numtr else ())
        self._check(names, arrs)
        pv = _p(passv_tr) if self.numtr else None
        _chk(self.lib.fimhip_ocn_transp2_r8(self.h, nstep, _p(temp), _p(saln), pv, _p(cumuflx), _p(dpinit),
                                            _p(dpfinl)), "ocn_transp2_r8")

    def set_hybgen(self, targt, salmin, dpmin, latij):
        """hycom_constants' theta (targt), salmin, dpmin (kdm) and the latitude
        in degrees (nip) for the grid generator; batrop from set_momtum."""
        t = [np.ascontiguousarray(x, np.float64) for x in (targt, salmin, dpmin, latij)]
        if any(x.shape != (self.kdm,) for x in t[:3]) or t[3].shape != (self.nip,):
            raise ValueError("set_hybgen: tables of kdm values and nip latitudes expected")
        _chk(self.lib.fimhip_ocn_set_hybgen(self.h, *map(_p, t)), "ocn_set_hybgen")

    def hybgn1(self, nstep, leap, targt, dens, dp, pres, temp, saln, uvel, vvel, latij, passv_tr=None):
        """hybgn1 (hycom_hybgn1.F90:9) on host arrays, updated in place
        (passv_tr remapped only when given, as hycom_run passes none)."""
        names = ["dens", "dp", "pres", "temp", "saln", "uclin", "vclin"] + (["passv"] if passv_tr is not None else [])
        arrs = (dens, dp, pres, temp, saln, uvel, vvel) + ((passv_tr,) if passv_tr is not None else ())
        self._check(names, arrs)
        t = np.ascontiguousarray(targt, np.float64)
        la = np.ascontiguousarray(latij, np.float64)
        _chk(self.lib.fimhip_ocn_hybgn1_r8(self.h, nstep, leap, _p(t), _p(dens), _p(dp), _p(pres), _p(temp),
                                           _p(saln), _p(uvel), _p(vvel), _p(la),
                                           _p(passv_tr) if passv_tr is not None else None), "ocn_hybgn1_r8")

    def step_hybgn1(self, leap: int, with_tracers: bool = False):
        """hybgn1 on the device-resident arrays (dens, dp, pres, temp, saln,
        uclin, vclin; passv with_tracers)."""
        _chk(self.lib.fimhip_ocn_hybgn1(self.h, leap, int(with_tracers)), "ocn_hybgn1")

    def sync(self):
        """Wait for the instance's asynchronous calls."""
        _chk(self.lib.fimhip_ocn_sync(self.h), "ocn_sync")

    def step_transp0(self, leapn: int):
        """transp0 on the device-resident arrays, from dp's level leapn."""
        _chk(self.lib.fimhip_ocn_transp0(self.h, leapn), "ocn_transp0")

    def step_transp1(self):
        """transp1 with the device-resident massflx."""
        _chk(self.lib.fimhip_ocn_transp1(self.h), "ocn_transp1")

    def step_transp2(self, leapn: int):
        """transp2 with dpfinl = dp's level leapn (hycom_run.F90:398-400)."""
        _chk(self.lib.fimhip_ocn_transp2(self.h, leapn), "ocn_transp2")

    def upload(self, name: str, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.shape != self._shape(name):
            raise ValueError("%s: shape %s, expected %s" % (name, a.shape, self._shape(name)))
        _chk(self.lib.fimhip_ocn_upload(self.h, OCN_FIELDS.index(name), _p(a)), "ocn_upload")

    def download(self, name: str) -> np.ndarray:
        a = np.empty(self._shape(name), np.float64)
        _chk(self.lib.fimhip_ocn_download(self.h, OCN_FIELDS.index(name), _p(a)), "ocn_download")
        return a

    def step(self, nstep: int, leapm: int, leapn: int):
        """cnuity on the device-resident arrays (upload them first)."""
        self._step_rc(self.lib.fimhip_ocn_cnuity(self.h, nstep, leapm, leapn), "ocn_cnuity")

    @property
    def mass(self):
        v = np.zeros(3)
        _chk(self.lib.fimhip_ocn_mass(self.h, *(v[j:].ctypes.data_as(_PD) for j in range(3))), "ocn_mass")
        return tuple(float(x) for x in v)

    @mass.setter
    def mass(self, m):
        _chk(self.lib.fimhip_ocn_set_mass(self.h, *map(float, m)), "ocn_set_mass")

    def set_dtleap(self, dtleap: float):
        """hycom_constants' dtleap for the calls that follow (hycom_run.F90:136)."""
        _chk(self.lib.fimhip_ocn_set_dtleap(self.h, float(dtleap)), "ocn_set_dtleap")

    def profile(self, on: bool = True):
        _chk(self.lib.fimhip_ocn_profile(self.h, int(on)), "ocn_profile")

    def profile_report(self) -> dict:
        buf = ct.create_string_buffer(1 << 14)
        _chk(self.lib.fimhip_ocn_profile_report(self.h, buf, len(buf)), "ocn_profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            n, c, t = line.split()
            out[n] = (int(c), float(t))
        return out


OceanCnuity = Ocean        # the continuity equation was its first routine

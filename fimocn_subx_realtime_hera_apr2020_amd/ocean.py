"""The ocean instance of the uneven batch (G8 x 32 layers, temperature and
salinity, a land mask), as a seeded state for the batched dynamics step.

It is a stand-in that carries an ocean-shaped workload through FIM's own
dynamics kernels with a land/sea mask, not the reference's HYCOM ocean
(hycom_*.F90): the edge fluxes are gated by the mask as the ocean's
continuity gates them (hycom_cnuity.F90:100-125), everything else is FIM's
horizontal step.  Because FIM's column code (hybgen's inflate, ptop, thktop)
is written for pressures near 1e5 Pa, the ocean column is mapped into that
range: 4000 m of water is 1e5 - 50 Pa of the model's pressure coordinate
(PA_PER_M Pa per metre), the top interface at 50 Pa as in the atmosphere.

* layers (32, level 1 at the bottom as in FIM): each at least its minimum
  thickness of FIMrun/ocn_dpmin.txt (2 m at the surface, 276 m at the
  bottom), all of them stretched by one factor to fill the 4000 m column,
  times (1 + 0.05 U[-1,1]);
* temperature (tracer 1, which plays theta): U[270, 303] K over the column,
  from a smooth surface field spread uniformly over [297, 303] K decaying
  towards 270 K with depth (e-folding depth 800 m), so it rises
  monotonically from the bottom up and varies little along the layers (a
  horizontal contrast of tens of kelvin would drive winds of 100-200 m/s in
  FIM's equations);
* salinity (tracer 2): U[34, 36] psu per cell and layer;
* currents: solid body 1 m/s about an axis tilted 45 degrees plus N(0, 0.1)
  m/s noise;
* land: a smooth random field on the sphere (a sum of 24 seeded Gaussian
  bumps) below its 30th percentile.
hybgen's target temperatures rise linearly from 270.5 K (bottom) to 302.5 K
(top); its minimum thicknesses are the ocn_dpmin layers in the model's
pressure units.
"""
from __future__ import annotations

import numpy as np

from . import constants as C

DEPTH_M = 4000.0
P_TOP = 50.0
PA_PER_M = (1.0e5 - P_TOP) / DEPTH_M
NVL = 32
NTRA = 2


def vertical_tables(nvl: int = NVL):
    """hybgen's target temperature and minimum thickness (Pa of the model's
    pressure coordinate) per layer, bottom layer first."""
    assert nvl == len(C.OCN_DPMIN32)
    thetac = np.linspace(270.5, 302.5, nvl)
    dpsig = np.array(C.OCN_DPMIN32[::-1], np.float64) * PA_PER_M
    return thetac, dpsig


def smooth_field(lat, lon, seed: int) -> np.ndarray:
    """A seeded smooth random field on the sphere: a sum of 24 Gaussian bumps
    of random centre, sign and width."""
    rng = np.random.default_rng(seed)
    lat = np.asarray(lat, np.float64)
    lon = np.asarray(lon, np.float64)
    x = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=1)
    c = rng.normal(size=(24, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    w = rng.uniform(-1.0, 1.0, 24)
    kappa = rng.uniform(2.0, 8.0, 24)
    return (w[None, :] * np.exp(kappa[None, :] * (x @ c.T - 1.0))).sum(axis=1)


def land_mask(lat, lon, seed: int = 20240605, frac: float = 0.30) -> np.ndarray:
    """1 for ocean, 0 for land: a seeded smooth field below its frac quantile
    is land."""
    f = smooth_field(lat, lon, seed)
    return (f >= np.quantile(f, frac)).astype(np.int32)


def make_ocean_state(lat, lon, nvl: int = NVL, seed: int = 20240605) -> dict:
    """Seeded ocean state (arrays as state.make_state: (nip, nvl),
    (nip, nvl+1), tracers (2, nip, nvl)) and its land mask ('wet')."""
    from . import fimhip
    rng = np.random.default_rng(seed + 1)
    lat = np.asarray(lat, np.float64)
    lon = np.asarray(lon, np.float64)
    nip = lat.shape[0]
    dpmin_m = np.array(C.OCN_DPMIN32[::-1], np.float64)          # bottom layer first
    stretch = DEPTH_M / dpmin_m.sum()
    dz = dpmin_m[None, :] * stretch * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, (nip, nvl)))
    dz = np.maximum(dz, dpmin_m[None, :])
    dp = dz * PA_PER_M
    pr = np.empty((nip, nvl + 1))
    pr[:, nvl] = P_TOP
    for k in range(nvl - 1, -1, -1):
        pr[:, k] = pr[:, k + 1] + dp[:, k]
    ex = C.CP * (pr / C.P1000) ** (C.RD / C.CP)
    # depth of each layer's centre below the surface (m)
    ztop = np.cumsum(dz[:, ::-1], axis=1)[:, ::-1]                 # depth of each layer's bottom
    zc = ztop - 0.5 * dz
    # surface temperature: a smooth field, its values mapped by rank onto
    # [297, 303] K (uniformly distributed, spatially smooth); with the decay
    # with depth the column spans 270-303 K
    f = smooth_field(lat, lon, seed + 2)
    tsurf = 297.0 + 6.0 * (np.argsort(np.argsort(f)) + 0.5) / nip
    temp = 270.0 + (tsurf[:, None] - 270.0) * np.exp(-zc / 800.0)
    sal = rng.uniform(34.0, 36.0, (nip, nvl))
    # currents: tilted solid body (as state.make_state's wind) scaled to 1 m/s, plus noise
    alpha = np.pi / 4
    u0 = 1.0 * (np.cos(lat) * np.cos(alpha) + np.sin(lat) * np.cos(lon) * np.sin(alpha))
    v0 = -1.0 * np.sin(lon) * np.sin(alpha)
    us = u0[:, None] + 0.1 * rng.normal(size=(nip, nvl))
    vs = v0[:, None] + 0.1 * rng.normal(size=(nip, nvl))
    ph = np.zeros((nip, nvl + 1))
    tr = np.stack([temp, sal])
    wet = land_mask(lat, lon, seed)
    del fimhip
    return dict(us=np.ascontiguousarray(us), vs=np.ascontiguousarray(vs), dp=np.ascontiguousarray(dp),
                pr=pr, ex=ex, ph=ph, tr=np.ascontiguousarray(tr), wet=wet)


# ---------------------------------------------------------------------------
# Inputs of the HYCOM ocean continuity equation (hycom_cnuity.F90, cnuity):
# HYCOM's own units and orientation (pressure in Pa, 1 m of water = rho*g Pa,
# layer 1 at the surface, pres(1) the surface pressure).
RHO_OCN = 1000.0                  # hycom_constants.F90: rho
ONEM = RHO_OCN * C.GRVITY         # rho * grvity: 1 m of water in Pa


def hycom_cnuity_case(grid, kdm: int = 32, seed: int = 20240607, dtleap: float = 1800.0) -> dict:
    """A seeded input set for cnuity on a glvl.Grid: the land mask 'wet'
    (nip,) int32 (a smooth field below its 30th percentile is land, and so
    is an ocean cell with no ocean neighbour); the
    thickness dp (2, nip, kdm) at the two leapfrog levels and the interface
    pressure pres (nip, kdm+1) (a smooth bathymetry of 2400-4000 m filled by
    the ocn_dpmin layers (the first kdm, surface first; for kdm > 32 the 32
    interpolated to kdm) stretched per column,
    times 1 + 0.3 U[-1,1] so the limiter has work, the second level 1 %
    apart); edge velocities um/vm (mid time) and un/vn (old), (nip, 6, kdm),
    each edge's value the mean of its two cells' solid-body-plus-noise
    currents (0.3 m/s, N(0, 0.1)), so both cells of an edge see the same
    one; the mid-time edge thickness dpm_edg likewise; the mass fluxes
    massflx and mssfx_avg (nip, 6, kdm) zero.  Land cells and edges carry
    the same construction; cnuity never reads them."""
    rng = np.random.default_rng(seed)
    lat = np.asarray(grid.lat, np.float64)
    lon = np.asarray(grid.lon, np.float64)
    nip = lat.shape[0]
    prox = grid.dyn_prox().astype(np.int64) - 1
    wet = land_mask(lat, lon, seed)
    # an ocean cell with no ocean neighbour would make the biharmonic
    # smoothing's neighbour average 0/0: such cells become land
    nwet = (wet[prox] > 0).sum(axis=1)
    wet = np.where(nwet > 0, wet, 0).astype(np.int32)
    f = smooth_field(lat, lon, seed + 3)
    depth = 2400.0 + 1600.0 * (np.argsort(np.argsort(f)) + 0.5) / nip
    dpmin_m = np.array(C.OCN_DPMIN32[:kdm], np.float64)
    if kdm > len(C.OCN_DPMIN32):      # more layers: the same profile, interpolated
        n = len(C.OCN_DPMIN32)
        dpmin_m = np.interp(np.linspace(0.0, n - 1.0, kdm), np.arange(n), np.asarray(C.OCN_DPMIN32, np.float64))
    dz = dpmin_m[None, :] * (depth / dpmin_m.sum())[:, None] * (1.0 + 0.3 * rng.uniform(-1.0, 1.0, (nip, kdm)))
    dp0 = dz * ONEM
    dp1 = dp0 * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, (nip, kdm)))
    dp = np.stack([dp0, dp1])
    pres = np.zeros((nip, kdm + 1))
    for k in range(kdm):
        pres[:, k + 1] = pres[:, k] + dp0[:, k]

    def cell_currents():
        alpha = np.pi / 4
        u0 = 0.3 * (np.cos(lat) * np.cos(alpha) + np.sin(lat) * np.cos(lon) * np.sin(alpha))
        v0 = -0.3 * np.sin(lon) * np.sin(alpha)
        return (u0[:, None] + 0.1 * rng.normal(size=(nip, kdm)), v0[:, None] + 0.1 * rng.normal(size=(nip, kdm)))

    def on_edges(c):
        return np.ascontiguousarray(0.5 * (c[:, None, :] + c[prox]))

    um, vm = cell_currents()
    un, vn = cell_currents()
    z = np.zeros((nip, 6, kdm))
    return dict(kdm=kdm, dtleap=dtleap, wet=wet, dp=np.ascontiguousarray(dp), pres=pres,
                um_edg=on_edges(um), vm_edg=on_edges(vm), un_edg=on_edges(un), vn_edg=on_edges(vn),
                dpm_edg=on_edges(dp[0]), massflx=z.copy(), mssfx_avg=z.copy())


EDGVAR_CASE_SEED = 20240611


def hycom_edgvar_case(grid, kdm: int = 32, seed: int = EDGVAR_CASE_SEED, dtleap: float = 1800.0) -> dict:
    """The inputs of bclin_edgvar (hycom_edgvar.F90:128) on a glvl.Grid, on top
    of hycom_cnuity_case's (whose dp, pres, wet, massflx and mssfx_avg it
    keeps): the baroclinic velocities uclin/vclin (2, nip, kdm) at the two
    leapfrog levels (0.2 m/s tilted solid body, N(0, 0.1) noise), the
    barotropic ones utrop/vtrop (nip, 3) (N(0, 0.05)), the geopotential
    anomaly geop (nip, kdm+1) at the interfaces (decreasing with depth from
    a smooth surface anomaly), the Montgomery potential montg and the
    specific-volume anomaly spcifv (nip, kdm) (layer means of smooth
    profiles with noise), and the edge arrays bclin_edgvar writes, all 7.0
    at the start (um_edg, vm_edg, un_edg, vn_edg, dpm_edg - the dp_edg that
    cnuity reads as its dpm_edg -, geop_edg, mont_edg, uvsq_edg, spcv_edg
    (nip, 6, kdm) and pr_edg (nip, 6, kdm+1)).  Land cells carry the same
    construction; bclin_edgvar reads some of their values with zero weight."""
    case = hycom_cnuity_case(grid, kdm=kdm, seed=seed, dtleap=dtleap)
    rng = np.random.default_rng(seed + 1)
    lat = np.asarray(grid.lat, np.float64)
    lon = np.asarray(grid.lon, np.float64)
    nip = lat.shape[0]
    alpha = np.pi / 4
    u0 = 0.2 * (np.cos(lat) * np.cos(alpha) + np.sin(lat) * np.cos(lon) * np.sin(alpha))
    v0 = -0.2 * np.sin(lon) * np.sin(alpha)
    uc = np.stack([u0[:, None] + 0.1 * rng.normal(size=(nip, kdm)) for _ in range(2)])
    vc = np.stack([v0[:, None] + 0.1 * rng.normal(size=(nip, kdm)) for _ in range(2)])
    ut = 0.05 * rng.normal(size=(nip, 3))
    vt = 0.05 * rng.normal(size=(nip, 3))
    pres = case["pres"]
    depth = pres / ONEM
    surf = 2.0 * smooth_field(lat, lon, seed + 5)
    geop = (surf[:, None] - 0.01 * depth) * C.GRVITY * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, pres.shape))
    zmid = 0.5 * (depth[:, :-1] + depth[:, 1:])
    spcifv = (1.0e-3 / 1.025) * (1.0 + 2.0e-3 * np.exp(-zmid / 800.0) + 1.0e-4 * rng.normal(size=(nip, kdm)))
    montg = 0.5 * (geop[:, :-1] + geop[:, 1:]) + 0.02 * rng.normal(size=(nip, kdm))
    case.update(uclin=np.ascontiguousarray(uc), vclin=np.ascontiguousarray(vc), utrop=ut, vtrop=vt,
                geop=np.ascontiguousarray(geop), montg=np.ascontiguousarray(montg),
                spcifv=np.ascontiguousarray(spcifv))
    for n in ("um_edg", "vm_edg", "un_edg", "vn_edg", "dpm_edg", "geop_edg", "mont_edg", "uvsq_edg", "spcv_edg"):
        case[n] = np.full((nip, 6, kdm), 7.0)
    case["pr_edg"] = np.full((nip, 6, kdm + 1), 7.0)
    return case


def hycom_momtum_case(grid, kdm: int = 32, seed: int = EDGVAR_CASE_SEED, dtleap: float = 1800.0) -> dict:
    """The inputs of momtum (hycom_momtum.F90:16) on top of
    hycom_edgvar_case's: wind stress taux/tauy (nip,) N(0, 0.1) Pa, the
    bottom friction velocity ustarb (nip,) zero, the barotropic time step
    batrop = dtleap / 2 (hycom_run.F90:136 at the first step), and the
    namelist defaults veldff = 0.1 m/s and smagcf = 0.3
    (fimnamelist.F90:253-254).  In a quarter of the columns the three
    bottom layers are thinned to 0.2-5 m (their thickness moved to the layer
    above, the bottom kept) so that the downward momentum mixing into thin
    bottom layers (hycom_momtum.F90:327-335) has work; pres follows the first
    level.  The edge values momtum reads are bclin_edgvar's outputs, to be
    formed from this case first."""
    case = hycom_edgvar_case(grid, kdm=kdm, seed=seed, dtleap=dtleap)
    rng = np.random.default_rng(seed + 2)
    nip = case["wet"].shape[0]
    dp = case["dp"].copy()
    thin = np.flatnonzero(rng.uniform(size=nip) < 0.25)
    for lv in range(2):
        new = rng.uniform(0.2, 5.0, (thin.size, 3)) * ONEM
        moved = dp[lv][thin][:, kdm - 3:].sum(axis=1) - new.sum(axis=1)
        dp[lv][thin, kdm - 3:] = new
        dp[lv][thin, kdm - 4] += moved
    pres = np.zeros_like(case["pres"])
    for k in range(kdm):
        pres[:, k + 1] = pres[:, k] + dp[0][:, k]
    case.update(dp=np.ascontiguousarray(dp), pres=pres)
    case.update(taux=0.1 * rng.normal(size=nip), tauy=0.1 * rng.normal(size=nip), ustarb=np.zeros(nip),
                batrop=0.5 * dtleap, veldff=0.1, smagcf=0.3)
    return case


def hycom_transport_case(grid, kdm: int = 32, numtr: int = 1, seed: int = EDGVAR_CASE_SEED,
                         batrop: float = 900.0) -> dict:
    """The inputs of HYCOM's tracer transport (hycom_transp3d.F90 transp0/1/2)
    on top of hycom_cnuity_case's grid, mask and thickness: the time-integrated
    mass flux cumuflx (nip, 6, kdm), each edge's two sides equal and opposite
    (U(-1,1) x 0.2 x the edge-mean thickness x the mean cell area, so layers
    gain or lose up to tens of percent and the vertical displacements reach
    several layers), 0 on edges to land, -1e33 in land cells; a mass flux
    massflx for transp1 built the same way at 1e-3 of that size; the
    thickness at the start (dpinit, the first leapfrog level) and end
    (dpfinl, x (1 + 0.05 U[-1,1])) of the interval; temperature (deg C, a
    smooth surface field of 10-28 decaying towards 2 with depth, e-folding
    800 m), salinity U[34, 36] and numtr passive tracers U[0, 1] (kdm per
    column); batrop and temdff = 0.02 (fimnamelist.F90:255)."""
    case = hycom_cnuity_case(grid, kdm=kdm, seed=seed)
    rng = np.random.default_rng(seed + 3)
    lat = np.asarray(grid.lat, np.float64)
    lon = np.asarray(grid.lon, np.float64)
    nip = lat.shape[0]
    wet = case["wet"] > 0
    prox = grid.dyn_prox().astype(np.int64) - 1
    proxs = grid.proxs.astype(np.int64) - 1
    area = np.asarray(grid.area, np.float64)
    dp0 = case["dp"][0]

    def pair_flux(scale):
        f = np.zeros((nip, 6, kdm))
        for i in np.flatnonzero(wet):
            for e in range(int(grid.nprox[i])):
                ix = prox[i, e]
                if wet[ix] and i < ix:
                    v = rng.uniform(-1.0, 1.0, kdm) * scale * 0.5 * (dp0[i] + dp0[ix]) * 0.5 * (area[i] + area[ix])
                    f[i, e] = v
                    f[ix, proxs[i, e]] = -v
        f[~wet] = -1.0e33
        return f

    cumuflx = pair_flux(0.2)
    massflx = pair_flux(2.0e-4)
    dpinit = dp0.copy()
    dpfinl = dp0 * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, dp0.shape))
    depth = case["pres"] / ONEM
    zmid = 0.5 * (depth[:, :-1] + depth[:, 1:])
    sst = 10.0 + 18.0 * (smooth_field(lat, lon, seed + 7) - 0.0)
    sst = np.clip(sst, 10.0, 28.0)
    temp = 2.0 + (sst[:, None] - 2.0) * np.exp(-zmid / 800.0) + 0.1 * rng.normal(size=(nip, kdm))
    saln = rng.uniform(34.0, 36.0, (nip, kdm))
    passv = rng.uniform(0.0, 1.0, (numtr, nip, kdm))
    case.update(cumuflx=cumuflx, massflx_tr=massflx, dpinit=dpinit, dpfinl=dpfinl,
                temp=np.ascontiguousarray(temp), saln=saln, passv=np.ascontiguousarray(passv), numtr=numtr,
                batrop=batrop, temdff=0.02)
    return case


def sigma1(t, s):
    """Potential density sigma_1 (kg/m^3 - 1000) of temperature t (deg C) and
    salinity s (psu): hycom_sigetc.F90's sigocn polynomial (a rational fit at
    1000 dbar), evaluated in numpy to build inputs."""
    t = np.asarray(t, np.float64)
    s = np.asarray(s, np.float64)
    p = 1.0e7 * 1.0e-4
    pn = (9.9984085444849347e2 + t * (7.3471625860981584 + t * (-5.3211231792841769e-2 + 3.6492439109814549e-4 * t))
          + s * (2.588057102399139 - 6.7168282786692355e-3 * t + 1.9203202055760151e-3 * s)
          + p * (1.1798263740430364e-2 + 9.8920219266399117e-8 * t * t + 4.699664277175473e-6 * s
                 + p * (-2.5862187075154352e-8 - 3.2921414007960662e-12 * t * t)))
    pd = (1.0 + t * (7.2815210113327091e-3 - 4.4787265461983921e-5 * t
                     + t * t * (3.3851002965802430e-7 + 1.3651202389758572e-10 * t))
          + s * (1.7632126669040377e-3 - 8.8066583251206474e-6 * t - 1.8832689434804897e-10 * t * t * t
                 + np.sqrt(np.maximum(0.0, s)) * (5.7463776745432097e-6 + 1.4716275472242334e-9 * t * t))
          + p * (6.7103246285651894e-6 + p * t * (-2.4461698007024582e-17 * t * t - 9.1534417604289062e-18 * p)))
    return pn / pd - 1000.0


def salmin_of(targt, temmin: float = -2.0):
    """hycom_init.F90:562-566: the salinity at which water of temperature
    temmin has density targt (hycom_sigetc.F90 sofsig, the quadratic in s of
    the sigma_1 fit's c1..c9)."""
    c1, c2, c3, c4, c5, c6, c7, c8, c9 = (5.019935E+00, 1.470290E-02, 7.904674E-01, -6.861022E-03, -3.051459E-03,
                                          3.099320E-05, 3.050299E-05, 1.046036E-04, 5.287567E-06)
    t = temmin
    aa = c8 + c9 * t
    bb = c3 + c5 * t + c7 * t * t
    cc = c1 + c2 * t + c4 * t * t + c6 * t * t * t - np.asarray(targt, np.float64)
    return (-bb + np.sqrt(bb * bb - 4.0 * aa * cc)) / (2.0 * aa)


def hycom_hybgen_case(grid, kdm: int = 32, seed: int = 20240617, numtr: int = 1, batrop: float = 900.0) -> dict:
    """The inputs of HYCOM's hybrid grid generator (hycom_hybgn1.F90:9) on
    top of hycom_cnuity_case's mask and thickness: in a quarter of the
    columns the three bottom layers thinned to 0.2-5 m and in a tenth the
    bottom two to 0-5 cm (massless layers for the sea-floor absorption), the
    thickness taken from the layer above; temperature a smooth surface field
    of 10-28 deg C decaying to 2 with depth (e-folding 800 m) plus
    N(0, 0.3), salinity U[34, 36], density their sigma_1 (so layers are
    lighter and denser than their targets in turn); velocities (both
    leapfrog levels) N(0, 0.3) m/s; numtr passive tracers U[0, 1]; the
    target densities of FIMrun/ocn_target.txt, salmin from them at -2 deg C,
    the minimum thicknesses of ocn_dpmin.txt (m); latitude in degrees;
    batrop.  Arrays in the reference layout (C order): dens, temp, saln
    (nip, kdm); dp, uvel, vvel (2, nip, kdm); pres (nip, kdm+1); passv
    (numtr, nip, kdm)."""
    case = hycom_cnuity_case(grid, kdm=kdm, seed=seed)
    rng = np.random.default_rng(seed + 11)
    nip = case["wet"].shape[0]
    lat = np.asarray(grid.lat, np.float64)
    lon = np.asarray(grid.lon, np.float64)
    dp = case["dp"].copy()
    thin = np.flatnonzero(rng.uniform(size=nip) < 0.25)
    nothing = np.flatnonzero(rng.uniform(size=nip) < 0.10)
    for lv in range(2):
        new = rng.uniform(0.2, 5.0, (thin.size, 3)) * ONEM
        moved = dp[lv][thin][:, kdm - 3:].sum(axis=1) - new.sum(axis=1)
        dp[lv][thin, kdm - 3:] = new
        dp[lv][thin, kdm - 4] += moved
        new = rng.uniform(0.0, 0.05, (nothing.size, 2)) * ONEM
        moved = dp[lv][nothing][:, kdm - 2:].sum(axis=1) - new.sum(axis=1)
        dp[lv][nothing, kdm - 2:] = new
        dp[lv][nothing, kdm - 3] += moved
    pres = np.zeros((nip, kdm + 1))
    for k in range(kdm):
        pres[:, k + 1] = pres[:, k] + dp[0][:, k]
    depth = pres / ONEM
    zmid = 0.5 * (depth[:, :-1] + depth[:, 1:])
    sst = np.clip(10.0 + 18.0 * smooth_field(lat, lon, seed + 7), 10.0, 28.0)
    temp = 2.0 + (sst[:, None] - 2.0) * np.exp(-zmid / 800.0) + 0.3 * rng.normal(size=(nip, kdm))
    saln = rng.uniform(34.0, 36.0, (nip, kdm))
    dens = sigma1(temp, saln)
    uvel = 0.3 * rng.normal(size=(2, nip, kdm))
    vvel = 0.3 * rng.normal(size=(2, nip, kdm))
    passv = rng.uniform(0.0, 1.0, (numtr, nip, kdm))
    targt = np.array(C.OCN_TARGET32, np.float64)
    dpmin = np.array(C.OCN_DPMIN32, np.float64)
    if kdm != 32:
        x = np.linspace(0.0, 31.0, kdm)
        targt = np.interp(x, np.arange(32), targt)
        dpmin = np.interp(x, np.arange(32), dpmin)
    return dict(kdm=kdm, wet=case["wet"], dp=np.ascontiguousarray(dp), pres=pres,
                temp=np.ascontiguousarray(temp), saln=saln, dens=np.ascontiguousarray(dens),
                uvel=np.ascontiguousarray(uvel), vvel=np.ascontiguousarray(vvel), passv=np.ascontiguousarray(passv),
                targt=targt, salmin=salmin_of(targt), dpmin=dpmin, latij=np.degrees(lat), batrop=batrop,
                numtr=numtr)


def hycom_ocean_case(grid, kdm: int = 32, seed: int = EDGVAR_CASE_SEED, numtr: int = 1) -> dict:
    """The inputs of a whole HYCOM dynamics interval (hycom_run.F90:364-556:
    bclin_edgvar, cnuity and momtum every step, transp1 every second step,
    transp2, hybgn1 and transp0 at the end of the interval): hycom_momtum_case
    plus temperature (a smooth surface field of 10-28 deg C decaying to 2 with
    depth, e-folding 800 m, plus N(0, 0.3)), salinity U[34, 36], their
    sigma_1 as dens, numtr passive tracers U[0, 1], the tracers' time
    integral and start thickness (set by transp0), temdff = 0.02 and the grid
    generator's tables (the targets of FIMrun/ocn_target.txt, salmin from
    them, ocn_dpmin.txt, latitude in degrees)."""
    case = hycom_momtum_case(grid, kdm=kdm, seed=seed)
    rng = np.random.default_rng(seed + 19)
    nip = case["wet"].shape[0]
    lat = np.asarray(grid.lat, np.float64)
    lon = np.asarray(grid.lon, np.float64)
    zmid = 0.5 * (case["pres"][:, :-1] + case["pres"][:, 1:]) / ONEM
    sst = np.clip(10.0 + 18.0 * smooth_field(lat, lon, seed + 7), 10.0, 28.0)
    temp = 2.0 + (sst[:, None] - 2.0) * np.exp(-zmid / 800.0) + 0.3 * rng.normal(size=(nip, kdm))
    saln = rng.uniform(34.0, 36.0, (nip, kdm))
    targt = np.array(C.OCN_TARGET32, np.float64)
    dpmin = np.array(C.OCN_DPMIN32, np.float64)
    if kdm != 32:
        x = np.linspace(0.0, 31.0, kdm)
        targt = np.interp(x, np.arange(32), targt)
        dpmin = np.interp(x, np.arange(32), dpmin)
    case.update(temp=np.ascontiguousarray(temp), saln=saln, dens=np.ascontiguousarray(sigma1(temp, saln)),
                passv=rng.uniform(0.0, 1.0, (numtr, nip, kdm)), numtr=numtr, temdff=0.02,
                cumuflx=np.zeros((nip, 6, kdm)), dpinit=np.zeros((nip, kdm)), targt=targt, salmin=salmin_of(targt),
                dpmin=dpmin, latij=np.degrees(lat))
    return case

"""Seeded synthetic FIM states for the benchmark workloads.

Every random number is a counter-based hash (splitmix64) of
``(seed, member, field, cell, level)``, so a member's state does not depend on
how cells, levels or members are batched or tiled.  Profiles:

* interface pressures ``p_k = 1e5 * (5e-4)**((k-1)/nvl)`` (1000 hPa to 0.5 hPa),
  layer thickness ``dp_k = (p_k - p_{k+1}) * (1 + 0.05 U[-1,1])``;
* potential temperature rising linearly 250 -> 600 K over the layers plus
  U[-0.5, 0.5] K;
* solid-body wind u0 = 38.6 m/s with the rotation axis tilted 45 degrees
  (Williamson et al. 1992, test 1), optionally plus N(0,1) m/s noise;
* tracers 2..T: a cosine bell of radius a/3 centred at (3 pi/2, 0) plus
  U[0, 1e-2] (or a log-normal field times the bell for the chemistry load);
* ensemble members perturb theta by N(0, 0.5 K), u, v by N(0, 1 m/s) and dp by
  (1 + 0.02 U[-1,1]).

Arrays are float64 with the level index fastest: ``(nip, nvl)``,
``(nip, nvl+1)`` and tracers ``(ntra, nip, nvl)``.
"""
from __future__ import annotations

import numpy as np

from . import constants as C

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)

FIELD_DP, FIELD_TH, FIELD_U, FIELD_V, FIELD_TR = 1, 2, 3, 4, 10
FIELD_MEMBER = 100     # member perturbation streams start here


def _mix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        return x ^ (x >> np.uint64(31))


def uniform01(seed: int, member: int, field: int, nip: int, nlev: int) -> np.ndarray:
    """U[0,1) on a (nip, nlev) lattice keyed by (seed, member, field, cell, level)."""
    with np.errstate(over="ignore"):
        key = _mix(np.uint64(seed) * _GOLD + np.uint64(member))
        key = _mix(key ^ (np.uint64(field) * _GOLD))
        cell = np.arange(nip, dtype=np.uint64)[:, None]
        lev = np.arange(nlev, dtype=np.uint64)[None, :]
        x = _mix(key ^ _mix(cell * np.uint64(0x100000) + lev + _GOLD))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, member, field, nip, nlev):
    u1 = uniform01(seed, member, field, nip, nlev)
    u2 = uniform01(seed, member, field + 50, nip, nlev)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def cosine_bell(lat: np.ndarray, lon: np.ndarray) -> np.ndarray:
    """Williamson test-1 bell, radius a/3 around (lon, lat) = (3 pi/2, 0)."""
    lam_c, th_c = 1.5 * np.pi, 0.0
    r = np.arccos(np.clip(np.sin(th_c) * np.sin(lat) + np.cos(th_c) * np.cos(lat) * np.cos(lon - lam_c), -1, 1))
    rb = 1.0 / 3.0
    return np.where(r < rb, 0.5 * (1.0 + np.cos(np.pi * r / rb)), 0.0)


def make_state(lat, lon, nvl: int = 64, ntra: int = 4, *, seed: int = 20240601,
               member: int = 0, wind_noise: bool = False, perturb: bool = False,
               tracer_kind: str = "bell") -> dict:
    """Build one member's state on cells with the given FP32 lat/lon (radians)."""
    lat = np.asarray(lat, np.float64)
    lon = np.asarray(lon, np.float64)
    nip = lat.shape[0]
    kk = np.arange(nvl + 1, dtype=np.float64)
    pint = 1.0e5 * (5.0e-4) ** (kk / nvl)                    # p_1 ... p_{nvl+1}
    dp0 = (pint[:-1] - pint[1:])[None, :]
    dp = dp0 * (1.0 + 0.05 * (2.0 * uniform01(seed, 0, FIELD_DP, nip, nvl) - 1.0))
    if perturb:
        dp = dp * (1.0 + 0.02 * (2.0 * uniform01(seed, member, FIELD_MEMBER + FIELD_DP, nip, nvl) - 1.0))
    pr = np.empty((nip, nvl + 1))
    pr[:, nvl] = pint[nvl]
    for k in range(nvl - 1, -1, -1):
        pr[:, k] = pr[:, k + 1] + dp[:, k]
    ex = C.CP * (pr / C.P1000) ** (C.RD / C.CP)

    lev = np.arange(nvl, dtype=np.float64)[None, :]
    th = 250.0 + 350.0 * lev / max(nvl - 1, 1) + (uniform01(seed, 0, FIELD_TH, nip, nvl) - 0.5)
    if perturb:
        th = th + 0.5 * normal(seed, member, FIELD_MEMBER + FIELD_TH, nip, nvl)

    u0, alpha = 38.6, 0.25 * np.pi
    us = np.broadcast_to((u0 * (np.cos(lat) * np.cos(alpha) + np.sin(lat) * np.cos(lon) * np.sin(alpha)))[:, None], (nip, nvl)).copy()
    vs = np.broadcast_to((-u0 * np.sin(lon) * np.sin(alpha))[:, None], (nip, nvl)).copy()
    if wind_noise:
        us += normal(seed, 0, FIELD_U, nip, nvl)
        vs += normal(seed, 0, FIELD_V, nip, nvl)
    if perturb:
        us += normal(seed, member, FIELD_MEMBER + FIELD_U, nip, nvl)
        vs += normal(seed, member, FIELD_MEMBER + FIELD_V, nip, nvl)

    tr = np.empty((ntra, nip, nvl))
    tr[0] = th
    bell = cosine_bell(lat, lon)[:, None]
    for t in range(1, ntra):
        if tracer_kind == "lognormal":
            z = normal(seed, 0, FIELD_TR + t, nip, nvl)
            tr[t] = np.exp(np.log(1.0e-9) + 1.5 * z) * bell
        else:
            tr[t] = bell + 1.0e-2 * uniform01(seed, 0, FIELD_TR + t, nip, nvl)
    ph = np.zeros((nip, nvl + 1))
    return dict(us=us, vs=vs, dp=dp, pr=pr, ex=ex, ph=ph, tr=tr)


def vertical_tables(nvl: int):
    """Target theta and minimum thickness per layer (FIMrun/theta_coor64.txt,
    FIMrun/dpsig64.txt); other layer counts interpolate the 64-layer tables."""
    if nvl == 64:
        return np.array(C.THETA_COOR64, np.float64), np.array(C.DPSIG64, np.float64)
    x = np.linspace(0, 63, nvl)
    th = np.interp(x, np.arange(64), np.array(C.THETA_COOR64, np.float64))
    ds = np.interp(x, np.arange(64), np.array(C.DPSIG64, np.float64))
    return np.round(th), np.round(ds)

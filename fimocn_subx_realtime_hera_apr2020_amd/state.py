"""Seeded synthetic FIM states for the benchmark workloads.

Every random number is a counter-based hash (splitmix64) of
``(seed, member, field, cell, level)``, so a member's state does not depend on
how cells, levels or members are batched or tiled.  The generator itself is
C++ (csrc/synth.h), shared by the host and the GPU.  Profiles:

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

def make_state(lat, lon, nvl: int = 64, ntra: int = 4, *, seed: int = 20240601, member: int = 0,
               wind_noise: bool = False, perturb: bool = False, tracer_kind: str = "bell") -> dict:
    """One member's state on cells with the given FP32 lat/lon (radians),
    computed by the library's host generator (synth.h); the device
    generator writes the same bytes."""
    from . import fimhip
    flags = (fimhip.SYNTH_WIND_NOISE if wind_noise else 0) | (fimhip.SYNTH_PERTURB if perturb else 0) | \
        (fimhip.SYNTH_LOGNORMAL if tracer_kind == "lognormal" else 0)
    return fimhip.synth_state(lat, lon, nvl, ntra, seed=seed, member=member, flags=flags)


def vertical_tables(nvl: int):
    """Target theta and minimum thickness per layer (FIMrun/theta_coor64.txt,
    FIMrun/dpsig64.txt); other layer counts interpolate the 64-layer tables."""
    if nvl == 64:
        return np.array(C.THETA_COOR64, np.float64), np.array(C.DPSIG64, np.float64)
    x = np.linspace(0, 63, nvl)
    th = np.interp(x, np.arange(64), np.array(C.THETA_COOR64, np.float64))
    ds = np.interp(x, np.arange(64), np.array(C.DPSIG64, np.float64))
    return np.round(th), np.round(ds)

"""Shared helpers of the test suite: golden fixtures and oracle runs."""
import hashlib
import json
import os

import numpy as np

from fimocn_subx_realtime_hera_apr2020_amd import glvl, state

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(v) -> str:
    return hashlib.sha256(np.ascontiguousarray(v, "<f8").tobytes()).hexdigest()


def golden_grid(level: int) -> glvl.Grid:
    return glvl.read_glvl(os.path.join(GOLDEN, "glvl_G%d.dat" % level))


def golden_input():
    z = np.load(os.path.join(GOLDEN, "dyn_G2_input.npz"))
    st = {k: z[k] for k in ("us", "vs", "dp", "pr", "ex", "ph", "tr")}
    return st, z["thetac"], z["dpsig"], float(z["dt"])


def golden_ref():
    with open(os.path.join(GOLDEN, "dyn_G2_ref.json")) as f:
        return json.load(f)


# the HYCOM continuity cases of the golden digests and the GPU tests: name,
# thkdff, biharm, first step, (totmass, massglb0, masscor) at the start; the
# last starts at step 959 so that its third step (day 10 at dtleap 1800 s)
# takes the 10-day mass correction
HYCOM_CASES = (("biharm", 0.05, 0.75, 1, (0.0, 0.0, 0.0)), ("laplace", 0.05, 0.0, 1, (0.0, 0.0, 0.0)),
               ("nosmooth", 0.0, 0.75, 1, (0.0, 0.0, 1e-9)), ("masscor", 0.05, 0.75, 959, (1.119e21, 1.12e21, 2.5e-9)))


# reference record suffix -> oracle / device field name
STATE_FIELDS = {"us": "us", "vs": "vs", "dp": "dp", "pr": "pr", "ex": "ex", "ph": "ph", "mp": "mp", "tr": "tr"}


def seeded_case(level_or_grid, nvl=64, ntra=4, wind_noise=True, seed=20240601, **kw):
    g = golden_grid(level_or_grid) if isinstance(level_or_grid, int) else level_or_grid
    th, ds = state.vertical_tables(nvl)
    st = state.make_state(g.lat, g.lon, nvl, ntra, wind_noise=wind_noise, seed=seed, **kw)
    return g, st, th, ds


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))

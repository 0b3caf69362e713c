"""Write the golden fixtures of tests/golden from the reference FIM build.

Run in a container where /root/reference exists and oracle/build_ref.sh has
built oracle/_ref:
    python tests/golden/make_golden.py

Fixtures (data only):
  glvl_G{1,2,3}.dat     reference grid prep chain output (grid -> ginfo -> getlvl,
                        gtype=2, curve=0), the file dyn_init.F90 reads
  icos_grid_info_G2.dat the reference ginfo output for G2 (icos_grid_info_level.dat,
                        GridGen.F90:437-497), the file getlvl reads
  dyn_G2_input.npz      seeded FP64 input state on the G2 grid (nvl=64, ntra=4)
  dyn_G2_schemes_ref.json  SHA-256 of the final records of the reference FP64
                        build for the alternative schemes on the same input:
                        mass_adv (flux_conserv_schm=2), and class-B tracers
                        (ntrb=2, dtratio=2, fields class_b_fields() of the
                        input) with cnuity and with mass_adv
  dyn_G2_ref.json       SHA-256 of every output record of the reference FP64
                        build (dynref_r8) for 3 full steps (mode 0) and 3
                        transport steps (mode 1), stage dumps of step 1, plus
                        a few sample values for diagnostics
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fimocn_subx_realtime_hera_apr2020_amd import glvl, state  # noqa: E402
from oracle import refrun  # noqa: E402

DT = 240.0


def sha(v: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(v, "<f8").tobytes()).hexdigest()


def main():
    if not refrun.available():
        sys.exit("oracle/_ref not built")
    for lv in (1, 2, 3):
        shutil.copyfile(refrun.ref_glvl(lv), os.path.join(HERE, "glvl_G%d.dat" % lv))
    shutil.copyfile(refrun.ref_file(2, "icos_grid_info_level.dat"), os.path.join(HERE, "icos_grid_info_G2.dat"))
    g = glvl.read_glvl(os.path.join(HERE, "glvl_G2.dat"))
    th, ds = state.vertical_tables(64)
    st = state.make_state(g.lat, g.lon, 64, 4, wind_noise=True)
    np.savez_compressed(os.path.join(HERE, "dyn_G2_input.npz"), thetac=th, dpsig=ds, dt=DT, **st)
    out = {"dt": DT, "nvl": 64, "ntra": 4, "nsteps": 3, "runs": {}}
    for mode in (0, 1):
        r = refrun.run_dynref(refrun.ref_glvl(2), st, th, ds, nsteps=3, dt=DT, mode=mode, dumpstep=1)
        rec = {}
        for k, v in r.items():
            if k.startswith("_"):
                continue
            vs = v if isinstance(v, list) else [v]
            rec[k] = {"sha256": [sha(x) for x in vs], "n": int(vs[0].size),
                      "sample": [float(x[len(x) // 3]) for x in vs]}
        out["runs"]["mode%d" % mode] = rec
    with open(os.path.join(HERE, "dyn_G2_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("golden fixtures written to", HERE)


def class_b_fields(st):
    """Two class-B tracer fields derived from the input state's tracers."""
    return np.ascontiguousarray(np.stack([st["tr"][1], 2.5 * st["tr"][2] + 1.e-3]))


SCHEME_RUNS = {   # name: (mode, flux_scheme, class-B tracers, steps)
    "massadv_mode0": (0, 2, False, 3),
    "classb_mode1": (1, 1, True, 5),
    "classb_massadv_mode0": (0, 2, True, 5),
}


def schemes():
    z = np.load(os.path.join(HERE, "dyn_G2_input.npz"))
    st = {k: z[k] for k in ("us", "vs", "dp", "pr", "ex", "ph", "tr")}
    th, ds, dt = z["thetac"], z["dpsig"], float(z["dt"])
    out = {"dt": dt, "dtratio": 2, "runs": {}}
    for name, (mode, scheme, cb, nsteps) in SCHEME_RUNS.items():
        s2 = dict(st)
        if cb:
            s2["trb"] = class_b_fields(st)
        r = refrun.run_dynref(refrun.ref_glvl(2), s2, th, ds, nsteps=nsteps, dt=dt, mode=mode,
                              init_hybgen=(mode == 0), flux_scheme=scheme, dtratio=2)
        out["runs"][name] = {k: [sha(x) for x in (v if isinstance(v, list) else [v])]
                             for k, v in r.items() if k.startswith("final.")}
    with open(os.path.join(HERE, "dyn_G2_schemes_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
    schemes()

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
  dyn_96_ref.json       the north star's FP64 bar: SHA-256 of the reference FP64
                        build's final state after 96 full steps at dt = 240 s
                        on the G3 grid (start-up hybgen, seeded state), and the
                        relative L2 of the checker with the device pow
                        (fim_pow) against the reference after 96 steps at G3
                        and G4, as measured when the fixture was made
  dyn_G2_r4_ref.npz     the reference built with default real 4 (dynref_r4)
                        after R4_STEPS full steps from dyn_G2_input (start-up
                        hybgen, dt = 240 s): final us, vs, dp, pr, ex, ph, mp,
                        tr as float32, and the relative L2 of that build
                        against the FP64 build (dynref_r8) per field
  hycom_cnuity_ref.json SHA-256 of every output record of the reference HYCOM
                        continuity equation (hycom_cnuity.F90, FP64, hycomref)
                        on the seeded inputs of ocean.hycom_cnuity_case at G2
                        and G3 x 32 layers, 3 leapfrog steps each, for the
                        smoothing branches of HYCOM_CASES
  hycom_edgvar_ref.json SHA-256 of every output record of the reference HYCOM
                        edge interpolation (hycom_edgvar.F90 bclin_edgvar,
                        FP64, hycedgref) and of the continuity equation run
                        on its edge values, step 1, on the seeded inputs of
                        ocean.hycom_edgvar_case at G2 and G3 x 32 layers,
                        with and without janjic_ocn
  hycom_transport_ref.json  SHA-256 of the outputs of the reference HYCOM tracer
                        transport (hycom_transp3d.F90 transp0/1/2 with
                        hycom_fct3d.F90 and hycom_dffusn.F90, FP64, hyctrref) on
                        ocean.hycom_transport_case at G2 and G3 x 32 layers
  hycom_momtum_ref.json SHA-256 of the outputs of the reference HYCOM momentum
                        equation (hycom_momtum.F90 momtum with hycom_dissip.F90,
                        FP64, hycmomref): uclin, vclin, pres, ustarb at step 1
                        on ocean.hycom_momtum_case at G2 and G3 x 32 layers,
                        with and without janjic_ocn, the edge values formed by
                        the checker's bclin_edgvar (momtum_inputs)
  dyn_G2_ref.json       SHA-256 of every output record of the reference FP64
                        build (dynref_r8) for 3 full steps (mode 0) and 3
                        transport steps (mode 1), stage dumps of step 1, the
                        reference globsum totals (qmass, qtrcr) at start-up
                        and after the last step, plus a few sample values for
                        diagnostics
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

sys.path.insert(0, os.path.dirname(HERE))
from helpers import HYCOM_CASES  # noqa: E402

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
    inp = os.path.join(HERE, "dyn_G2_input.npz")
    if os.path.exists(inp):      # the committed input is the fixture: digests are always made from it
        z = np.load(inp)
        st = {k: z[k] for k in ("us", "vs", "dp", "pr", "ex", "ph", "tr")}
    else:
        st = state.make_state(g.lat, g.lon, 64, 4, wind_noise=True)
        np.savez_compressed(inp, thetac=th, dpsig=ds, dt=DT, **st)
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


def fp64_bar_state(g):
    """The seeded state of the 96-step runs: the synthetic generator's
    member 0 with wind noise (seed 20240601)."""
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip
    return fimhip.synth_state(g.lat, g.lon, 64, 4, seed=20240601, member=0, flags=fimhip.SYNTH_WIND_NOISE)


FP64_FIELDS = ("us", "vs", "dp", "pr", "ex", "ph", "mp")


def fp64_bar():
    from oracle import oracle as O
    th, ds = state.vertical_tables(64)
    out = {"dt": DT, "nsteps": 96, "nvl": 64, "ntra": 4, "rel_l2_fim_pow_vs_reference": {}}
    for lv in (3, 4):
        path = refrun.ref_glvl(lv)
        g = glvl.read_glvl(path)
        st = fp64_bar_state(g)
        r = refrun.run_dynref(path, st, th, ds, nsteps=96, dt=DT, mode=0)
        o = O.OracleRun(O.grid_arrays(g, 64, DT, th, ds), 64, 4, DT, pow_kind=1)
        o.load(st)
        o.init(True)
        for its in range(1, 97):
            o.step(its, 0)
        err = {}
        for f in FP64_FIELDS:
            d = o.a[f].ravel() - r["final." + f]
            err[f] = float(np.linalg.norm(d) / np.linalg.norm(r["final." + f]))
        for t in range(4):
            d = o.a["tr"][t].ravel() - r["final.tr"][t]
            err["tr%d" % (t + 1)] = float(np.linalg.norm(d) / np.linalg.norm(r["final.tr"][t]))
        out["rel_l2_fim_pow_vs_reference"]["G%d" % lv] = err
        if lv == 3:
            out["G3_final_sha256"] = {k: [sha(x) for x in (v if isinstance(v, list) else [v])]
                                      for k, v in r.items() if k.startswith("final.")}
    with open(os.path.join(HERE, "dyn_96_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


R4_STEPS = 12
R4_FIELDS = ("us", "vs", "dp", "pr", "ex", "ph", "mp", "tr")


def r4_reference():
    z = np.load(os.path.join(HERE, "dyn_G2_input.npz"))
    st = {k: z[k] for k in ("us", "vs", "dp", "pr", "ex", "ph", "tr")}
    th, ds, dt = z["thetac"], z["dpsig"], float(z["dt"])
    path = refrun.ref_glvl(2)
    r4 = refrun.run_dynref(path, st, th, ds, nsteps=R4_STEPS, dt=dt, mode=0, prec="r4")
    r8 = refrun.run_dynref(path, st, th, ds, nsteps=R4_STEPS, dt=dt, mode=0, prec="r8")
    out, err = {}, {}
    for f in R4_FIELDS:
        a, b = r4["final." + f], r8["final." + f]
        if isinstance(a, list):
            a, b = np.stack(a), np.stack(b)
        shape = st[f].shape if f != "mp" else st["us"].shape
        out[f] = a.reshape(shape).astype(np.float32)
        err[f] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    np.savez_compressed(os.path.join(HERE, "dyn_G2_r4_ref.npz"), steps=R4_STEPS, dt=dt,
                        rel_l2_fields=np.array(R4_FIELDS), rel_l2_r4_vs_r8=np.array([err[f] for f in R4_FIELDS]),
                        **out)


# (name, thkdff, biharm, first step, starting (totmass, massglb0, masscor)):
# biharmonic and Laplacian interface smoothing, none, and the biharmonic one
# across the 10-day mass-correction update (step 960 at dtleap = 1800 s)


def hycom_reference():
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    out = {"kdm": 32, "nsteps": 3, "runs": {}}
    for lv in (2, 3):
        path = os.path.join(HERE, "glvl_G%d.dat" % lv)
        g = glvl.read_glvl(path)
        case = ocean.hycom_cnuity_case(g, kdm=32)
        for name, thkdff, biharm, nstep0, mass in HYCOM_CASES:
            r = refrun.run_hycomref(path, case, nsteps=3, thkdff=thkdff, biharm=biharm, nstep0=nstep0, mass=mass)
            out["runs"]["G%d_%s" % (lv, name)] = {k: [sha(x) for x in v] for k, v in r.items() if not k.startswith("_")}
    with open(os.path.join(HERE, "hycom_cnuity_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def hycom_edgvar_reference():
    """Digests of the reference bclin_edgvar followed by cnuity at step 1
    (hycedgref with chain) at G2 and G3 x 32 layers, with and without
    janjic_ocn, leapm = 2."""
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    out = {"kdm": 32, "nstep": 1, "leapm": 2, "runs": {}}
    for lv in (2, 3):
        path = os.path.join(HERE, "glvl_G%d.dat" % lv)
        g = glvl.read_glvl(path)
        case = ocean.hycom_edgvar_case(g, kdm=32)
        for janjic in (False, True):
            r = refrun.run_hycedgref(path, case, nstep=1, leapm=2, janjic=janjic, chain=True)
            out["runs"]["G%d_janjic%d" % (lv, int(janjic))] = {k: sha(v) for k, v in r.items()
                                                               if not k.startswith("_")}
    with open(os.path.join(HERE, "hycom_edgvar_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def momtum_inputs(O, g, case):
    """The edge values momtum reads, formed from the case by the checker's
    bclin_edgvar (byte-equal to the reference's) with janjic_ocn on, so that
    geop_edg is set too; and a checker instance whose pres is the case's."""
    h = O.HycomCnuity(g, case)
    h.edgvar(g, 2, 1, True)
    edges = {n: h.a[n].copy() for n in ("um_edg", "vm_edg")}
    edges.update({n: h.e[n].copy() for n in ("geop_edg", "mont_edg", "uvsq_edg", "spcv_edg")})
    h.a["pres"][:] = case["pres"]
    return h, edges


def hycom_momtum_reference():
    """Digests of the reference momtum at step 1 (leapm = 2) at G2 and G3 x
    32 layers, with and without janjic_ocn, on ocean.hycom_momtum_case and
    the edge values momtum_inputs forms from it."""
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    from oracle import oracle as O
    out = {"kdm": 32, "nstep": 1, "leapm": 2, "runs": {}}
    for lv in (2, 3):
        path = os.path.join(HERE, "glvl_G%d.dat" % lv)
        g = glvl.read_glvl(path)
        case = ocean.hycom_momtum_case(g, kdm=32)
        _, edges = momtum_inputs(O, g, case)
        for janjic in (False, True):
            r = refrun.run_hycmomref(path, case, edges, nstep=1, leapm=2, janjic=janjic)
            out["runs"]["G%d_janjic%d" % (lv, int(janjic))] = {k: sha(v) for k, v in r.items()
                                                               if not k.startswith("_")}
    with open(os.path.join(HERE, "hycom_momtum_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def hycom_transport_reference():
    """Digests of the reference tracer transport (transp0 on copies, then
    transp1 and transp2) at G2 and G3 x 32 layers, one passive tracer, on
    ocean.hycom_transport_case."""
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    out = {"kdm": 32, "numtr": 1, "leapn": 1, "runs": {}}
    for lv in (2, 3):
        path = os.path.join(HERE, "glvl_G%d.dat" % lv)
        g = glvl.read_glvl(path)
        case = ocean.hycom_transport_case(g, kdm=32)
        r = refrun.run_hyctrref(path, case, leapn=1)
        out["runs"]["G%d" % lv] = {k: sha(v) for k, v in r.items() if not k.startswith("_")}
    with open(os.path.join(HERE, "hycom_transport_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def hycom_hybgen_reference():
    """Digests of the reference hybrid grid generator (hybgn1) at G2 and G3 x
    32 layers on ocean.hycom_hybgen_case: level 1 without tracers (as
    hycom_run calls it) and level 2 with one passive tracer."""
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    out = {"kdm": 32, "runs": {}}
    for lv in (2, 3):
        g = glvl.read_glvl(os.path.join(HERE, "glvl_G%d.dat" % lv))
        case = ocean.hycom_hybgen_case(g, kdm=32)
        for leap, wt in ((1, False), (2, True)):
            r = refrun.run_hychybref(case, leap=leap, with_tracers=wt)
            out["runs"]["G%d_leap%d%s" % (lv, leap, "_tr" if wt else "")] = {
                k: sha(v) for k, v in r.items() if not k.startswith("_")}
    with open(os.path.join(HERE, "hycom_hybgen_ref.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
    schemes()
    fp64_bar()
    r4_reference()
    hycom_reference()
    hycom_edgvar_reference()
    hycom_momtum_reference()
    hycom_transport_reference()
    hycom_hybgen_reference()

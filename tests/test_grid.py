"""The product's icosahedral grid generator (fimhip_icos_grid) against the
reference prep chain's glvl.dat (tests/golden, made by the reference
grid -> ginfo -> getlvl programs) and against the geometry it must satisfy."""
import math

import numpy as np
import pytest

from fimocn_subx_realtime_hera_apr2020_amd import fimhip, glvl
from helpers import golden_grid

EXACT = ["lat", "lon", "nprox", "prox", "proxs", "area", "cs", "sn", "sidevec_c", "sidevec_e", "sideln",
         "rprox_ln", "corner_xy", "inv_perm", "hfcoef"]


@pytest.mark.parametrize("level", [1, 2, 3])
def test_grid_bitexact_vs_reference(level):
    ref = golden_grid(level)
    got = fimhip.icos_grid(level)
    assert got.nip == ref.nip == glvl.nip_of(level)
    for f in EXACT:
        assert np.array_equal(getattr(got, f), getattr(ref, f)), f


def test_grid_live_reference_g5():
    from oracle import refrun
    if not refrun.available():
        pytest.skip("reference build (oracle/_ref) not present")
    ref = glvl.read_glvl(refrun.ref_glvl(5))
    got = fimhip.icos_grid(5)
    for f in EXACT:
        assert np.array_equal(getattr(got, f), getattr(ref, f)), f


def test_pseudo_inverse_vs_reference_sgesvd():
    """getlvl's least-squares pseudo-inverse (the single-precision sgesvd the
    reference vendors) against the reference's own pseudoInverse, built from
    its sources (oracle/ref/lsfitref.F90), bit for bit, on 600 random 5- and
    6-point stencils of radii 1e-3 to 0.5 (stereographic units)."""
    import os
    import struct
    import subprocess
    from oracle import refrun
    exe = os.path.join(refrun.BIN, "lsfitref")
    if not os.path.exists(exe):
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(7)
    cases = []
    for c in range(600):
        n = 5 if c % 4 == 0 else 6
        r = 10 ** rng.uniform(-3, -0.3)
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = r * rng.uniform(0.7, 1.3, n)
        cases.append(np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1).astype(np.float32))
    work = os.path.join(refrun.REF, "runs")
    os.makedirs(work, exist_ok=True)
    fin, fout = os.path.join(work, "ls_in_%d.bin" % os.getpid()), os.path.join(work, "ls_out_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(struct.pack("<i", len(cases)))
        for xy in cases:
            f.write(struct.pack("<i", xy.shape[0]) + xy.astype("<f4").tobytes())
    subprocess.check_call([exe, fin, fout])
    raw = np.fromfile(fout, "<f4")
    os.remove(fin)
    os.remove(fout)
    off = 0
    for xy in cases:
        n = xy.shape[0]
        ref = raw[off:off + 5 * n].reshape(n, 5).T
        off += 5 * n
        got = fimhip.ls_pseudo_inverse(xy)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert off == raw.size


def test_pseudo_inverse_is_a_left_inverse():
    """inv A = I for the observation matrix [x, y, x*x, x*y, y*y] (a full-rank
    5-point fit is exact, a 6-point one is the least-squares solution whose
    normal equations inv A = I also hold), and inv is numpy's FP64
    pseudo-inverse, to single precision at condition numbers of a few hundred
    (measured: residual 3e-3, relative difference 2e-4)."""
    rng = np.random.default_rng(3)
    for n in (5, 6):
        for _ in range(50):
            ang = np.sort(rng.uniform(0, 2 * np.pi, n))
            rad = 0.05 * rng.uniform(0.8, 1.2, n)
            xy = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1).astype(np.float32)
            x, y = xy[:, 0].astype(np.float64), xy[:, 1].astype(np.float64)
            a = np.stack([x, y, x * x, x * y, y * y], 1)
            inv = fimhip.ls_pseudo_inverse(xy).astype(np.float64)
            p = np.linalg.pinv(a)
            assert np.abs(inv @ a - np.eye(5)).max() < 1e-2
            assert np.abs(inv - p).max() < 1e-3 * np.abs(p).max()


@pytest.mark.parametrize("level", [4, 5])
def test_grid_geometry(level):
    g = fimhip.icos_grid(level)
    nip = g.nip
    assert nip == 10 * 4 ** level + 2
    assert (g.nprox == 5).sum() == 12 and (g.nprox == 6).sum() == nip - 12
    # neighbour symmetry: prox(proxs(e,i), prox(e,i)) == i (dyn_init.F90:615-628)
    for i in range(nip):
        for e in range(g.nprox[i]):
            j = g.prox[i, e] - 1
            assert g.prox[j, g.proxs[i, e] - 1] == i + 1
    # the cells tile the sphere: the planar (stereographic) polygon areas
    # exceed the spherical ones by a second-order amount, 0.15 / 4**level
    total = g.area.astype(np.float64).sum()
    excess = total / (4 * math.pi * 6371220.0 ** 2) - 1
    assert 0 < excess < 0.2 / 4 ** level
    # each edge vector seen from the other side is its negative (corner swap of getlvl.F90:129-145)
    for i in range(0, nip, 37):
        for e in range(g.nprox[i]):
            j, eo = g.prox[i, e] - 1, g.proxs[i, e] - 1
            assert np.allclose(g.sidevec_e[i, e], -g.sidevec_e[j, eo], rtol=2e-5, atol=1e-2)
    # the closed polygon: edge vectors from the centre sum to ~0
    s = np.abs(g.sidevec_c.astype(np.float64).sum(axis=1)).max()
    assert s < 1e-3 * g.sideln.max()


def test_grid_info_file_bitexact_vs_reference_fixture(tmp_path):
    """icos_grid_info_level.dat (the ginfo output getlvl reads) written from
    the product's generator is the reference ginfo's file byte for byte, and
    reads back to the same contents."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "icos_grid_info_G2.dat")
    info = fimhip.icos_grid_info(2)
    out = tmp_path / "info.dat"
    glvl.write_grid_info(info, str(out))
    assert out.read_bytes() == open(path, "rb").read()
    back = glvl.read_grid_info(path)
    assert (back.glvl, back.curve, back.nip) == (2, 0, glvl.nip_of(2))
    for f in ("lat", "lon", "prox", "nprox", "icos_edge", "inv_perm"):
        assert np.array_equal(getattr(back, f), getattr(info, f)), f
    # the glvl.dat tables of the same grid agree with the intermediate file
    g = fimhip.icos_grid(2)
    assert np.array_equal(g.lat, info.lat) and np.array_equal(g.nprox, info.nprox)


@pytest.mark.parametrize("level", [3, 4, 5])
def test_grid_info_file_live_reference(tmp_path, level):
    from oracle import refrun
    if not refrun.available():
        pytest.skip("reference build (oracle/_ref) not present")
    out = tmp_path / "info.dat"
    glvl.write_grid_info(fimhip.icos_grid_info(level), str(out))
    assert out.read_bytes() == open(refrun.ref_file(level, "icos_grid_info_level.dat"), "rb").read()

"""The icosahedral grid tables of FIM and the glvl.dat file format.

A :class:`Grid` holds what FIMsrc/cntl/module_constants.F90:45-94 holds for
the dynamics: the neighbour tables ``prox/proxs/nprox`` and the FP32 metrics
written by FIMsrc/prep/grid/getlvl.F90:296-346.  Arrays keep the reference's
Fortran layout with the first Fortran index fastest, i.e. Fortran
``prox(6,nip)`` is the C-order array ``prox[nip, 6]`` and Fortran
``cs(4,6,nip)`` is ``cs[nip, 6, 4]``.  Neighbour indices stay 1-based here, as
in the file; the device copies are made 0-based by the library.

:func:`read_glvl` / :func:`write_glvl` read and write the Fortran unformatted
sequential file (4-byte record markers) with the record order of
getlvl.F90:296-346 (read back by dyn_init.F90:388-538).
"""
from __future__ import annotations

import dataclasses
import struct

import numpy as np

from . import constants as C


@dataclasses.dataclass
class Grid:
    glvl: int
    curve: int
    lat: np.ndarray          # float32 [nip]  (radians)
    lon: np.ndarray          # float32 [nip]
    nprox: np.ndarray        # int32 [nip]
    proxs: np.ndarray        # int32 [nip, 6]  neighbour's index of the shared edge (1-based)
    prox: np.ndarray         # int32 [nip, 6]  neighbour across edge (1-based)
    area: np.ndarray         # float32 [nip]
    cs: np.ndarray           # float32 [nip, 6, 4]
    sn: np.ndarray           # float32 [nip, 6, 4]
    sidevec_c: np.ndarray    # float32 [nip, 6, 2]
    sidevec_e: np.ndarray    # float32 [nip, 6, 2]
    sideln: np.ndarray       # float32 [nip, 6]
    rprox_ln: np.ndarray     # float32 [nip, 6]
    inv_perm: np.ndarray     # int32 [nip]
    corner_xy: np.ndarray    # float32 [nip, 2, 6]
    hfcoef: np.ndarray       # float32 [nip, 6, 6]  hfcoef(j,i,ipn) -> [ipn, i, j]

    @property
    def nip(self) -> int:
        return int(self.lat.shape[0])

    def dyn_prox(self) -> np.ndarray:
        """prox with slot 6 of the 12 pentagons set to slot 5 (dyn_init.F90:546-548)."""
        p = self.prox.copy()
        pent = self.nprox == 5
        p[pent, 5] = p[pent, 4]
        return p

    def derived(self) -> dict:
        """Cell scalars derived at start-up (dyn_init.F90:753-757), in FP64 as the
        FP64 build computes them from the FP32 file values."""
        lat = self.lat.astype(np.float64)
        return dict(
            corio=C.OMEGX2 * np.sin(lat),
            rarea=1.0 / self.area.astype(np.float64),
            grndspd=C.EARTHRAD * 2.0 * C.PI * np.cos(lat) / 86164.0,
        )


def concat(grids) -> Grid:
    """Several grids as one: the cells of grid i follow those of grids 0..i-1
    and their neighbour indices are offset accordingly, so no cell has a
    neighbour in another grid.  A state on the union advances every grid's
    instance in the same launches (instances of different resolution sharing
    the layers and the time step); each instance's cells compute exactly what
    they compute alone."""
    off = 0
    parts = {k: [] for k in ("lat", "lon", "nprox", "proxs", "prox", "area", "cs", "sn", "sidevec_c", "sidevec_e",
                             "sideln", "rprox_ln", "inv_perm", "corner_xy", "hfcoef")}
    for g in grids:
        for k in parts:
            a = getattr(g, k)
            if k == "prox":          # slots past nprox (pentagons) are left as they are
                used = np.arange(6)[None, :] < g.nprox[:, None]
                a = np.where(used, a + off, a).astype(np.int32)
            elif k == "inv_perm":
                a = (a + off).astype(np.int32)
            parts[k].append(a)
        off += g.nip
    return Grid(-1, grids[0].curve, **{k: np.ascontiguousarray(np.concatenate(v)) for k, v in parts.items()})


def nip_of(glvl: int) -> int:
    """Cell count of an all-bisection grid (module_control.F90:209-213)."""
    return 10 * 4 ** glvl + 2


# ---------------------------------------------------------------------------
# Fortran unformatted sequential I/O
# ---------------------------------------------------------------------------

def _records(buf: bytes):
    off = 0
    n = len(buf)
    while off < n:
        (ln,) = struct.unpack_from("<i", buf, off)
        rec = buf[off + 4: off + 4 + ln]
        (ln2,) = struct.unpack_from("<i", buf, off + 4 + ln)
        if ln2 != ln:
            raise ValueError("glvl: corrupt record markers")
        yield rec
        off += ln + 8


def read_glvl(path: str) -> Grid:
    with open(path, "rb") as f:
        buf = f.read()
    recs = _records(buf)
    hg = next(recs).decode()
    hc = next(recs).decode()
    glvl = int(hg[6:8])
    curve = int(hc[7:9])
    f4 = lambda: np.frombuffer(next(recs), dtype="<f4").copy()
    i4 = lambda: np.frombuffer(next(recs), dtype="<i4").copy()
    lat = f4()
    lon = f4()
    nip = lat.shape[0]
    nprox = i4()
    proxs = np.stack([i4() for _ in range(6)], axis=1)
    prox = np.stack([i4() for _ in range(6)], axis=1)
    area = f4()
    cs = np.empty((nip, 6, 4), np.float32)
    for isn in range(6):
        for idx in range(4):
            cs[:, isn, idx] = f4()
    sn = np.empty((nip, 6, 4), np.float32)
    for isn in range(6):
        for idx in range(4):
            sn[:, isn, idx] = f4()
    svc = np.empty((nip, 6, 2), np.float32)
    for isn in range(6):
        for idx in range(2):
            svc[:, isn, idx] = f4()
    sve = np.empty((nip, 6, 2), np.float32)
    for isn in range(6):
        for idx in range(2):
            sve[:, isn, idx] = f4()
    sideln = np.stack([f4() for _ in range(6)], axis=1)
    rprox_ln = np.stack([f4() for _ in range(6)], axis=1)
    inv_perm = i4()
    cxy = np.empty((nip, 2, 6), np.float32)
    for ixy in range(2):
        for isn in range(6):
            cxy[:, ixy, isn] = f4()
    hf = np.empty((nip, 6, 6), np.float32)
    for isn in range(6):
        for idx in range(6):
            hf[:, isn, idx] = f4()
    return Grid(glvl, curve, lat, lon, nprox, proxs, prox, area, cs, sn, svc, sve,
                sideln, rprox_ln, inv_perm, cxy, hf)


def write_glvl(grid: Grid, path: str) -> None:
    out = bytearray()

    def rec(b: bytes):
        out.extend(struct.pack("<i", len(b)))
        out.extend(b)
        out.extend(struct.pack("<i", len(b)))

    rec(("glvl =%2d" % grid.glvl).ljust(16).encode())
    rec(("curve =%2d" % grid.curve).ljust(16).encode())
    f4 = lambda a: rec(np.ascontiguousarray(a, dtype="<f4").tobytes())
    i4 = lambda a: rec(np.ascontiguousarray(a, dtype="<i4").tobytes())
    f4(grid.lat)
    f4(grid.lon)
    i4(grid.nprox)
    for isn in range(6):
        i4(grid.proxs[:, isn])
    for isn in range(6):
        i4(grid.prox[:, isn])
    f4(grid.area)
    for isn in range(6):
        for idx in range(4):
            f4(grid.cs[:, isn, idx])
    for isn in range(6):
        for idx in range(4):
            f4(grid.sn[:, isn, idx])
    for isn in range(6):
        for idx in range(2):
            f4(grid.sidevec_c[:, isn, idx])
    for isn in range(6):
        for idx in range(2):
            f4(grid.sidevec_e[:, isn, idx])
    for isn in range(6):
        f4(grid.sideln[:, isn])
    for isn in range(6):
        f4(grid.rprox_ln[:, isn])
    i4(grid.inv_perm)
    for ixy in range(2):
        for isn in range(6):
            f4(grid.corner_xy[:, ixy, isn])
    for isn in range(6):
        for idx in range(6):
            f4(grid.hfcoef[:, isn, idx])
    with open(path, "wb") as f:
        f.write(bytes(out))


# ---------------------------------------------------------------------------
# icos_grid_info_level.dat: the ginfo output that getlvl reads
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class GridInfo:
    """The contents of icos_grid_info_level.dat (GridGen.F90:437-497, single
    precision), the intermediate file between the prep programs ginfo and
    getlvl."""
    glvl: int
    curve: int
    lat: np.ndarray          # float32 [nip] (radians)
    lon: np.ndarray          # float32 [nip]
    prox: np.ndarray         # int32 [nip, 6]  pattern neighbours (1-based)
    nprox: np.ndarray        # int32 [nip]
    icos_edge: np.ndarray    # float32 [nip, 2, 2, 6]  Fortran icos_edge(edge, end, lat/lon, nip)
    inv_perm: np.ndarray     # int32 [nip]

    @property
    def nip(self) -> int:
        return int(self.lat.shape[0])


def read_grid_info(path: str) -> GridInfo:
    with open(path, "rb") as f:
        buf = f.read()
    recs = _records(buf)
    hg = next(recs).decode()
    hc = next(recs).decode()
    ll = np.frombuffer(next(recs), dtype="<f4")
    nip = ll.shape[0] // 2
    i4 = lambda: np.frombuffer(next(recs), dtype="<i4").copy()
    prox = np.stack([i4() for _ in range(6)], axis=1)
    nprox = i4()
    edge = np.empty((nip, 2, 2, 6), np.float32)
    for i3 in range(2):
        for i2 in range(2):
            for isn in range(6):
                edge[:, i3, i2, isn] = np.frombuffer(next(recs), dtype="<f4")
    inv_perm = i4()
    return GridInfo(int(hg[6:8]), int(hc[7:9]), ll[:nip].copy(), ll[nip:].copy(), prox, nprox, edge, inv_perm)


def write_grid_info(info: GridInfo, path: str) -> None:
    out = bytearray()

    def rec(b: bytes):
        out.extend(struct.pack("<i", len(b)))
        out.extend(b)
        out.extend(struct.pack("<i", len(b)))

    rec(("glvl =%2d" % info.glvl).ljust(16).encode())
    rec(("curve =%2d" % info.curve).ljust(16).encode())
    rec(np.concatenate([np.asarray(info.lat, "<f4"), np.asarray(info.lon, "<f4")]).tobytes())
    for isn in range(6):
        rec(np.ascontiguousarray(info.prox[:, isn], dtype="<i4").tobytes())
    rec(np.ascontiguousarray(info.nprox, dtype="<i4").tobytes())
    for i3 in range(2):
        for i2 in range(2):
            for isn in range(6):
                rec(np.ascontiguousarray(info.icos_edge[:, i3, i2, isn], dtype="<f4").tobytes())
    rec(np.ascontiguousarray(info.inv_perm, dtype="<i4").tobytes())
    with open(path, "wb") as f:
        f.write(bytes(out))

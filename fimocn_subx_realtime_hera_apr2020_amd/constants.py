"""Physical constants and hot-path namelist defaults of FIM.

Values follow FIMsrc/cntl/module_constants.F90:14-24 and the namelist defaults
of FIMsrc/cntl/fimnamelist.F90:95-119; the fixed-size parameters follow
FIMsrc/cntl/module_control.F90:95-104.  Arrays are stored with the level index
fastest, so a Fortran ``(nvl, nip)`` array is a C-order ``(nip, nvl)`` array.
"""

PI = 3.14159265            # module_constants.F90:14 (truncated value, used by grndspd)
RADDEG = 180.0 / PI
EARTHRAD = 6371220.0       # m
OMEGX2 = 1.4584e-4         # 2 x earth rotation rate (1/s)
GRVITY = 9.80665
P1000 = 100000.0           # Pa
CP = 1004.6855
RD = 287.0586
QVMIN = 1.0e-10
QWMIN = 1.0e-10

NPP = 6                    # module_control.F90:101
NABL = 3                   # Adams-Bashforth levels
KBL = 2                    # fixed-depth surface layers

# hot-path namelist defaults (fimnamelist.F90:95-119)
INTFC_SMOOTH = 30.0
PTOP = 30.0
THKTOP = 30.0
RLEIGH = 2.0
RLTHRESH = 90.0
RLTAPER = 0.4
SLAK = 0.5
MIGLIM = 1
REMAP_OPTN = 2
VELDFF_BKGND = 0.3
VELDFF_BOOST = 3.0

# the 64-layer tables FIMrun/theta_coor64.txt and FIMrun/dpsig64.txt (first 64 values)
THETA_COOR64 = [
    224, 232, 240, 245, 250, 255, 260, 265, 270, 273,
    276, 279, 282, 285, 288, 291, 293, 295, 297, 299,
    301, 303, 305, 307, 309, 311, 313, 315, 317, 319,
    321, 323, 325, 328, 332, 337, 343, 350, 358, 367,
    377, 388, 400, 413, 427, 443, 462, 485, 513, 547,
    590, 638, 692, 746, 800, 850, 900, 960, 1030, 1100,
    1200, 1350, 1700, 2200]
DPSIG64 = [
    330, 510, 675, 825, 960, 1080, 1185, 1275, 1350, 1410,
    1455, 1485] + [1500] * 52

# minimum ocean layer thickness (m), surface layer first: the first 32 values
# of FIMrun/ocn_dpmin.txt, the 32-layer table hycom_init.F90:169-170 reads
OCN_DPMIN32 = [2.0, 3.0, 4.0, 5.0, 6.0, 8.0, 10.0, 12.0, 15.0, 18.0, 21.0, 25.0, 29.0, 33.0, 38.0, 43.0,
               48.0, 54.0, 60.0, 66.0, 73.0, 81.0, 90.0, 101.0, 114.0, 129.0, 146.0, 166.0, 189.0, 215.0,
               244.0, 276.0]

# target densities (sigma_1, kg/m^3 - 1000), surface layer first: the first
# 32 values of FIMrun/ocn_target.txt, which hycom_init.F90:558-559 reads into
# hycom_constants' theta for kdm = 32
OCN_TARGET32 = [20.820, 21.640, 22.460, 23.280, 24.100, 24.920, 25.740, 26.550, 27.340, 28.100,
                28.820, 29.490, 30.100, 30.640, 31.100, 31.480, 31.780, 32.010, 32.180, 32.300,
                32.380, 32.430, 32.460, 32.480, 32.500, 32.520, 32.540, 32.560, 32.580, 32.600,
                32.620, 32.640]

"""The graphics file (.grp) that viewers of CADMAS-SURF/3D results read.

VF_OG1INI writes the header (Source code/vf_og1ini.f) and VF_OG1TRN one
record group per selected step (vf_og1trn.f): the output box of FILE GRP AREA
(default: the whole interior, vf_ol1ini.f:372-377), cell flags, the three
staggered velocities, pressure and F over that box, and the boundary values in
the graphics order of VF_OGBCNM/OGBCIN/OGBCVL (x faces, then y, then z faces
inside the box, each in boundary-list order).  Like .rsl it is a Fortran
unformatted sequential file (rsl.py): 4-byte length markers around every
record, little endian.  Single rank, no k-epsilon / temperature / scalars,
static porosity, no vorticity (FILE GRP ON VORT is not written).
"""
import struct

import numpy as np

from .rsl import VERSION, _f64, _get, _i32, _put


def _box(shape, area):
    """(I1,J1,K1,I2,J2,K2) of the output box, 1-based with ghosts."""
    numi, numj, numk = shape
    if area is None:
        area = (1, 1, 1, 1, 1, 1)
    a = list(area)
    a[0] = 2 if a[0] <= 1 else a[0]
    a[1] = 2 if a[1] <= 1 else a[1]
    a[2] = 2 if a[2] <= 1 else a[2]
    a[3] = numi - 1 if a[3] <= 1 else a[3]
    a[4] = numj - 1 if a[4] <= 1 else a[4]
    a[5] = numk - 1 if a[5] <= 1 else a[5]
    return tuple(a)


def boundary_order(shape, indb, box):
    """Boundary-list positions in graphics order and the counts (NBX, NBY,
    NBZ): x faces (sides 1, 2) with I1 <= I <= I2+1, then y faces, then z
    faces, each only when the box is not flat along that axis
    (vf_ogbcnm.f, vf_ogbcin.f)."""
    numi, numj, _ = shape
    i1, j1, k1, i2, j2, k2 = box
    ix, jy, kz = i2 + 1, j2 + 1, k2 + 1
    order = ([], [], [])
    indb = np.asarray(indb)
    for l in range(indb.shape[1]):
        ijk, ns = int(indb[0, l]), int(indb[1, l])
        k = (ijk - 1) // (numi * numj) + 1
        r = ijk - numi * numj * (k - 1)
        j = (r - 1) // numi + 1
        i = r - numi * (j - 1)
        if ns in (1, 2):
            if i1 != i2 and i1 <= i <= ix and j1 <= j <= j2 and k1 <= k <= k2:
                order[0].append(l)
        elif ns in (3, 4):
            if j1 != j2 and i1 <= i <= i2 and j1 <= j <= jy and k1 <= k <= k2:
                order[1].append(l)
        else:
            if k1 != k2 and i1 <= i <= i2 and j1 <= j <= j2 and k1 <= k <= kz:
                order[2].append(l)
    return order[0] + order[1] + order[2], tuple(len(o) for o in order)


class Writer:
    def __init__(self, path, xx, yy, zz, indb, ggv, area=None):
        ggv = np.asarray(ggv)
        self.shape = ggv.shape
        numi, numj, numk = self.shape
        self.box = _box(self.shape, area)
        i1, j1, k1, i2, j2, k2 = self.box
        self.order, (nbx, nby, nbz) = boundary_order(self.shape, indb, self.box)
        self.f = open(path, "wb")
        f = self.f
        _put(f, _i32(VERSION))
        _put(f, _i32((numi - 1, numj - 1, numk - 1)))
        xx, yy, zz = (np.asarray(a) for a in (xx, yy, zz))
        _put(f, _f64((xx[0, 1], yy[0, 1], zz[0, 1])))
        _put(f, _f64((xx[0, numi - 1], yy[0, numj - 1], zz[0, numk - 1])))
        _put(f, _i32(self.box))
        _put(f, _i32((nbx, nby, nbz)))
        # LN,LV,LP,LF,LK,LT,LS,LG,L1,L2
        _put(f, _i32((1, 1, 1, 1, 0, 0, 0, 0, 0, 0)))
        _put(f, _f64(xx[0, 1:numi]))
        _put(f, _f64(yy[0, 1:numj]))
        _put(f, _f64(zz[0, 1:numk]))
        if self.order:
            _put(f, _i32(np.asarray(indb)[0, self.order]))
        _put(f, _f64(self._cut(ggv)))

    def _cut(self, a, di=0, dj=0, dk=0):
        i1, j1, k1, i2, j2, k2 = self.box
        return np.asarray(a)[i1 - 1:i2 + di, j1 - 1:j2 + dj, k1 - 1:k2 + dk]

    def write_step(self, nnow, tnow, state):
        """vf_og1trn.f: NNOW/TNOW, NF, U, V, W (+ boundary values), P, F."""
        f = self.f
        _put(f, struct.pack("<id", int(nnow), float(tnow)))
        _put(f, _i32(self._cut(state["NF"])))
        _put(f, _f64(self._cut(state["UU"], di=1)))
        _put(f, _f64(self._cut(state["VV"], dj=1)))
        _put(f, _f64(self._cut(state["WW"], dk=1)))
        bc = lambda k: _f64(np.asarray(state[k])[self.order])  # noqa: E731
        if self.order:
            for k in ("BCU", "BCV", "BCW"):
                _put(f, bc(k))
        _put(f, _f64(self._cut(state["PP"])))
        if self.order:
            _put(f, bc("BCP"))
        _put(f, _f64(self._cut(state["FF"])))
        if self.order:
            _put(f, bc("BCF"))

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read(path):
    """Parse a .grp file written by VF_OG1INI / VF_OG1TRN (or Writer): returns
    (header, steps).  header holds the file's version, cell counts, extent,
    output box (I1,J1,K1,I2,J2,K2), boundary counts (NBX,NBY,NBZ), the ten
    output switches, the grid coordinates, the boundary cell numbers and GGV
    over the box; each step holds NNOW, TNOW, NF, UU, VV, WW, PP, FF over the
    box (the velocities with their extra face) and, when the box has
    boundary faces, BCU, BCV, BCW, BCP, BCF in graphics order."""
    with open(path, "rb") as f:
        rec = []
        while True:
            r = _get(f)
            if r is None:
                break
            rec.append(r)
    f64 = lambda b: np.frombuffer(b, "<f8")  # noqa: E731
    i32 = lambda b: np.frombuffer(b, "<i4")  # noqa: E731
    h = {"VERSION": i32(rec[0]), "N": i32(rec[1]), "XMIN": f64(rec[2]), "XMAX": f64(rec[3]),
         "BOX": i32(rec[4]), "NB": i32(rec[5]), "SWITCH": i32(rec[6]),
         "X": f64(rec[7]), "Y": f64(rec[8]), "Z": f64(rec[9])}
    i1, j1, k1, i2, j2, k2 = (int(v) for v in h["BOX"])
    nb = int(h["NB"].sum())
    q = 10
    if nb:
        h["IJK"] = i32(rec[q])
        q += 1
    dims = (i2 - i1 + 1, j2 - j1 + 1, k2 - k1 + 1)
    h["GGV"] = f64(rec[q]).reshape(dims, order="F")
    q += 1
    steps = []
    while q < len(rec):
        st = {}
        nnow, tnow = struct.unpack("<id", rec[q])
        st["NNOW"], st["TNOW"] = nnow, tnow
        st["NF"] = i32(rec[q + 1]).reshape(dims, order="F")
        q += 2
        for key, di in (("UU", (1, 0, 0)), ("VV", (0, 1, 0)), ("WW", (0, 0, 1))):
            st[key] = f64(rec[q]).reshape(tuple(a + b for a, b in zip(dims, di)), order="F")
            q += 1
        if nb:
            for key in ("BCU", "BCV", "BCW"):
                st[key] = f64(rec[q])
                q += 1
        st["PP"] = f64(rec[q]).reshape(dims, order="F")
        q += 1
        if nb:
            st["BCP"] = f64(rec[q])
            q += 1
        st["FF"] = f64(rec[q]).reshape(dims, order="F")
        q += 1
        if nb:
            st["BCF"] = f64(rec[q])
            q += 1
        steps.append(st)
    return h, steps

"""ctypes access to the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.

Fields are numpy arrays in the reference's layout: Fortran order, shape
(NUMI, NUMJ, NUMK); metric tables (6, N); INDB (4, NUMB).
"""
import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class OrcState(C.Structure):
    _fields_ = [
        ("numi", C.c_int), ("numj", C.c_int), ("numk", C.c_int), ("numb", C.c_int),
        ("xx", PD), ("yy", PD), ("zz", PD),
        ("uu", PD), ("vv", PD), ("ww", PD), ("pp", PD), ("ff", PD),
        ("anu", PD), ("cd0", PD),
        ("ggv", PD), ("ggx", PD), ("ggy", PD), ("ggz", PD),
        ("glv", PD), ("glx", PD), ("gly", PD), ("glz", PD),
        ("bcu", PD), ("bcv", PD), ("bcw", PD), ("bcp", PD), ("bcf", PD),
        ("bcvi", PD),
        ("nf", PI), ("indx", PI), ("indy", PI), ("indz", PI), ("indc", PI), ("indb", PI),
        ("inds", PI), ("nlim", PI),
        ("nums", C.c_int),
        ("rho0", C.c_double), ("grz0", C.c_double), ("anu0", C.c_double), ("dtnow", C.c_double),
        ("scmvp", C.c_double), ("vvmax", C.c_double), ("flower", C.c_double),
        ("fupper", C.c_double), ("plower", C.c_double),
        ("cgpara", C.c_double), ("cgepsa", C.c_double), ("cgepsr", C.c_double),
        ("cgdiv", C.c_double),
        ("icgtyp", C.c_int), ("icgmax", C.c_int), ("ibsuw0", C.c_int), ("loops", C.c_int),
        ("cgbnrm", C.c_double), ("cgxnrm", C.c_double), ("fsum", C.c_double),
        ("fcut", C.c_double), ("icgitr", C.c_int),
        ("tnow", C.c_double), ("wbub", C.c_double), ("idrop", C.c_int),
        ("tbub", PD), ("droptx", PD), ("dropty", PD), ("droptz", PD),
        ("dropuu", PD), ("dropvv", PD), ("dropww", PD),
        ("wvlvl", C.c_double),
        ("wv_on", C.c_int * 4), ("wv_l1", C.c_int * 4), ("wv_l2", C.c_int * 4),
        ("wv_d", C.c_double * 4), ("wv_t", C.c_double * 4), ("wv_a", C.c_double * 4),
        ("wv_t0", C.c_double * 4), ("wv_level", C.c_double * 4),
        ("mtbzz", C.c_int), ("mtbtt", C.c_int), ("mtbnow", C.c_int),
        ("mtb_zz", PD), ("mtb_tt", PD), ("mtb_hh", PD), ("mtb_un", PD), ("mtb_ut", PD),
        ("op_on", C.c_int * 4), ("op_d", C.c_double * 4), ("op_t", C.c_double * 4),
        ("idrgn", C.c_int), ("drgyu", C.c_double), ("drgdr", C.c_double * 10),
        ("drgap", C.c_double * 10), ("drgbt", C.c_double * 10),
        ("dm_on", C.c_int * 4), ("dm_deg", C.c_int * 4), ("dm_par", (C.c_double * 4) * 4),
        ("glxmin", C.c_double), ("glxmax", C.c_double), ("glymin", C.c_double),
        ("glymax", C.c_double), ("mtb_off", C.c_int), ("mtb_only", C.c_int),
        ("mtbzz2", C.c_int), ("mtbtt2", C.c_int), ("mtbnow2", C.c_int),
        ("mtb_zz2", PD), ("mtb_tt2", PD), ("mtb_hh2", PD), ("mtb_un2", PD), ("mtb_ut2", PD),
        ("mtb_off2", C.c_int), ("mtb_only2", C.c_int), ("wv_tab", C.c_int * 4),
    ]


class OrcMatrix(C.Structure):
    _fields_ = [(n, PD) for n in ("ad", "ali", "alj", "alk", "aui", "auj", "auk", "bb")]


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} missing: run make -C oracle")
        L = C.CDLL(str(path))
        L.orc_predictor.argtypes = [C.POINTER(OrcState)]
        L.orc_vpcoef.argtypes = [C.POINTER(OrcState), C.POINTER(OrcMatrix)]
        L.orc_m1bcgs.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcMatrix), PD, PI,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 PD, PD, PI]
        L.orc_mzdcmp.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcMatrix), PD, C.c_double,
                                 PI, C.c_int]
        L.orc_mzminv.argtypes = [C.c_int, C.c_int, C.c_int, PD, C.POINTER(OrcMatrix), PD, PD, PI]
        L.orc_mzax.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcMatrix), PD, PD, PI]
        L.orc_mzip.argtypes = [C.c_int, C.c_int, C.c_int, PD, PD, PI]
        L.orc_mzip.restype = C.c_double
        L.orc_set_dot_order.argtypes = [C.c_int]
        L.orc_set_dot_order.restype = C.c_int
        L.orc_vmodif.argtypes = [C.POINTER(OrcState), PD]
        L.orc_post_bc.argtypes = [C.POINTER(OrcState)]
        L.orc_v1cal.argtypes = [C.POINTER(OrcState), PD]
        L.orc_f1cal.argtypes = [C.POINTER(OrcState)]
        L.orc_fbubup.argtypes = [C.POINTER(OrcState)]
        L.orc_fbubup.restype = None
        L.orc_fdropf.argtypes = [C.POINTER(OrcState)]
        L.orc_cdiv00.argtypes = [C.POINTER(OrcState)]
        L.orc_cdiv00.restype = C.c_double
        L.orc_step.argtypes = [C.POINTER(OrcState), PD]
        L.orc_v1cal_loop.argtypes = [C.POINTER(OrcState), PD, C.c_int]
        L.orc_cdtcal.argtypes = [C.POINTER(OrcState), C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_double]
        L.orc_cdtcal.restype = C.c_double
        L.orc_init_state.argtypes = [C.POINTER(OrcState), C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int]
        _LIB = L
    return _LIB


def _pd(a):
    assert a.dtype == np.float64 and a.flags.f_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(PD)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.f_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(PI)


F64 = ("XX", "YY", "ZZ", "UU", "VV", "WW", "PP", "FF", "ANU", "CD0", "GGV", "GGX", "GGY", "GGZ",
       "GLV", "GLX", "GLY", "GLZ", "BCU", "BCV", "BCW", "BCP", "BCF", "BCVI",
       "TBUB", "DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW")
DROP = ("DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW")
I32 = ("NF", "INDX", "INDY", "INDZ", "INDC", "INDB")

DEFAULT_PARAMS = dict(rho0=1000.0, grz0=9.8, anu0=1.0e-6, dtnow=1.0e-3, scmvp=1.0, vvmax=30.0,
                      flower=1.0e-6, fupper=1.0 - 1.0e-6, plower=1.0e-4, cgpara=0.95,
                      cgepsa=1.0e-15, cgepsr=1.0e-12, cgdiv=1.0e30, icgtyp=1, icgmax=500,
                      ibsuw0=0, loops=1, tnow=0.0, wbub=0.2, idrop=1)


class Fields:
    """The state of one step in the reference layout, owned by numpy arrays."""

    def __init__(self, arrays, params, inds=None):
        self.a = {}
        for k in F64:
            if k in arrays:
                self.a[k] = np.asfortranarray(np.array(arrays[k], dtype=np.float64))
        for k in I32:
            if k in arrays:
                self.a[k] = np.asfortranarray(np.array(arrays[k], dtype=np.int32))
        numi, numj, numk = self.a["UU"].shape
        self.shape = (numi, numj, numk)
        self.numb = self.a["INDB"].shape[1]
        if "BCF" not in self.a:
            self.a["BCF"] = np.zeros(self.numb, order="F")
        if "BCVI" not in self.a:
            self.a["BCVI"] = np.zeros(self.numb, order="F")
        if "CD0" not in self.a:
            self.a["CD0"] = np.zeros(self.shape, order="F")
        # TimerDoor state at its VF_CINIT value unless given (vf_cinit.f:412-421)
        if "TBUB" not in self.a:
            self.a["TBUB"] = np.zeros(numk)
        for k in DROP:
            if k not in self.a:
                self.a[k] = np.full(self.shape, -1.0e10 if k == "DROPTX" else 0.0, order="F")
        self.a["INDS"] = np.zeros(numi * numj * numk, dtype=np.int32)
        nums = 0
        if inds is not None:
            nums = len(inds)
            self.a["INDS"][:nums] = inds
        self.a["NLIM"] = np.zeros(self.shape, dtype=np.int32, order="F")
        self.params = dict(DEFAULT_PARAMS)
        self.params.update(params)
        self.st = OrcState()
        st = self.st
        st.numi, st.numj, st.numk = numi, numj, numk
        st.numb = self.numb
        for k in F64:
            setattr(st, k.lower(), _pd(self.a[k]))
        for k in I32:
            setattr(st, k.lower(), _pi(self.a[k]))
        st.inds = _pi(self.a["INDS"])
        st.nlim = _pi(self.a["NLIM"])
        st.nums = nums
        for k, v in self.params.items():
            setattr(st, k, v)

    def set_waves(self, waves, table, wvlvl, table2=None):
        """Table wave-makers: waves = deck.waves (side -> DEPTH, PERIOD, AMPL,
        AREA; FUNC MATRIX reads table, MATRIX2 table2, each an mtb.Table)."""
        st = self.st
        st.wvlvl = wvlvl
        numi, numj, _ = self.shape
        for jd, wv in waves.items():
            q = jd - 1
            st.wv_on[q] = 1
            st.wv_tab[q] = 1 if wv.get("func") == "MATRIX2" else 0
            lim = numj if jd <= 2 else numi
            l1, l2 = wv["area"] if wv["area"] else (2, lim - 1)
            st.wv_l1[q], st.wv_l2[q] = l1, l2
            st.wv_d[q], st.wv_t[q], st.wv_a[q] = wv["depth"], wv["period"], wv["ampl"]
            st.wv_t0[q] = (table2 if st.wv_tab[q] else table).t0
        self.tabs = []
        for tab, sfx in ((table, ""), (table2, "2")):
            if tab is None:
                continue
            arrs = {k: np.asfortranarray(np.array(getattr(tab, k), dtype=np.float64))
                    for k in ("zz", "tt", "hh", "un", "ut")}
            self.tabs.append(arrs)
            setattr(st, "mtbzz" + sfx, len(tab.zz))
            setattr(st, "mtbtt" + sfx, len(tab.tt))
            setattr(st, "mtbnow" + sfx, 0)
            setattr(st, "mtb_off" + sfx, 1 if tab.kind == "LEVEL-OFF" else 0)
            setattr(st, "mtb_only" + sfx, 1 if tab.kind == "LEVEL-ONLY" else 0)
            for k in ("zz", "tt", "hh", "un", "ut"):
                setattr(st, f"mtb_{k}{sfx}", _pd(arrs[k]))

    def set_opens(self, opens):
        """Radiation sides: opens = deck.opens (side -> depth, period)."""
        for jd, op in opens.items():
            q = jd - 1
            self.st.op_on[q] = 1
            self.st.op_d[q], self.st.op_t[q] = op["depth"], op["period"]

    def set_drag_df(self, drag):
        """OPTION DRAG-DF: drag = (DRGYU, [(DRGDR, DRGAP, DRGBT), ...]) or
        None for the CD form."""
        st = self.st
        if drag is None:
            st.idrgn = 0
            return
        yu, rows = drag
        st.idrgn, st.drgyu = len(rows), yu
        for l, (dr, ap, bt) in enumerate(rows):
            st.drgdr[l], st.drgap[l], st.drgbt[l] = dr, ap, bt

    def set_damps(self, damps):
        """Damping zones: damps = deck.damps (side -> degree, pxy, pz, width,
        depth); the grid's extent from the metric tables (the first and last
        coordinates, XX(1,2) and XX(1,NUMI))."""
        st = self.st
        xx, yy = self.a["XX"], self.a["YY"]
        st.glxmin, st.glxmax = xx[0, 1], xx[0, -1]
        st.glymin, st.glymax = yy[0, 1], yy[0, -1]
        for jd, dm in damps.items():
            q = jd - 1
            st.dm_on[q], st.dm_deg[q] = 1, dm["degree"]
            for m, key in enumerate(("pxy", "pz", "width", "depth")):
                st.dm_par[q][m] = dm[key]

    def __getitem__(self, k):
        return self.a[k]

    @property
    def nums(self):
        return self.st.nums

    @property
    def inds(self):
        return self.a["INDS"][: self.st.nums].copy()

    # ---- phases ---------------------------------------------------------
    def predictor(self):
        return lib().orc_predictor(C.byref(self.st))

    def v1cal(self):
        pt = np.zeros(self.shape, order="F")
        rc = lib().orc_v1cal(C.byref(self.st), _pd(pt))
        return rc, pt

    def f1cal(self):
        return lib().orc_f1cal(C.byref(self.st))

    def fbubup(self):
        lib().orc_fbubup(C.byref(self.st))

    def fdropf(self):
        return lib().orc_fdropf(C.byref(self.st))

    def v1cal_loop(self, iloop):
        pt = np.zeros(self.shape, order="F")
        rc = lib().orc_v1cal_loop(C.byref(self.st), _pd(pt), iloop)
        return rc, pt

    def cdtcal(self, nnow, dtinit, dtsafe, dtmin, dtmax):
        return lib().orc_cdtcal(C.byref(self.st), nnow, dtinit, dtsafe, dtmin, dtmax)

    def step(self):
        pt = np.zeros(self.shape, order="F")
        rc = lib().orc_step(C.byref(self.st), _pd(pt))
        return rc, pt

    def init_state(self, wvlvl=0.0, uvw=(0.0, 0.0, 0.0), flags=3):
        return lib().orc_init_state(C.byref(self.st), wvlvl, uvw[0], uvw[1], uvw[2], flags)

    def cdiv00(self):
        return lib().orc_cdiv00(C.byref(self.st))

    def vpcoef(self):
        m = {k: np.zeros(self.shape, order="F") for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")}
        mat = OrcMatrix(*[_pd(m[k]) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")])
        rc = lib().orc_vpcoef(C.byref(self.st), C.byref(mat))
        return rc, m


def m1bcgs(mat, indc, x0=None, cgpara=0.95, cgepsa=1e-15, cgepsr=1e-12, cgdiv=1e30, icgtyp=1,
           icgmax=500):
    """VF_M1BCGS on a matrix given as dict AD..BB (Fortran-order float64 arrays)."""
    shape = mat["AD"].shape
    ms = {k: np.asfortranarray(np.array(mat[k], dtype=np.float64)) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")}
    m = OrcMatrix(*[_pd(ms[k]) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")])
    x = np.zeros(shape, order="F") if x0 is None else np.asfortranarray(np.array(x0, dtype=np.float64))
    ic = np.asfortranarray(np.array(indc, dtype=np.int32))
    bn, xn, it = C.c_double(), C.c_double(), C.c_int()
    lib().orc_m1bcgs(shape[0], shape[1], shape[2], C.byref(m), _pd(x), _pi(ic), cgpara, cgepsa,
                     cgepsr, cgdiv, icgtyp, icgmax, C.byref(bn), C.byref(xn), C.byref(it))
    return x, bn.value, xn.value, it.value


DOT_SERIAL, DOT_EXTENDED, DOT_BLOCKED = 0, 1, 2


def set_dot_order(order):
    """Summation order of the solver's dot products (orc_set_dot_order):
    DOT_SERIAL is the reference's; returns the previous order."""
    prev = lib().orc_set_dot_order(int(order))
    if prev < 0:
        raise ValueError(f"unknown dot order {order}")
    return prev


def mzdcmp(mat, indc, cgpara=0.95, icgtyp=1):
    shape = mat["AD"].shape
    ms = {k: np.asfortranarray(np.array(mat[k], dtype=np.float64)) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")}
    m = OrcMatrix(*[_pd(ms[k]) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")])
    dd = np.zeros(shape, order="F")
    ic = np.asfortranarray(np.array(indc, dtype=np.int32))
    ok = lib().orc_mzdcmp(shape[0], shape[1], shape[2], C.byref(m), _pd(dd), cgpara, _pi(ic), icgtyp)
    return dd, ok


def mzminv(mat, dd, y, indc):
    shape = mat["AD"].shape
    ms = {k: np.asfortranarray(np.array(mat.get(k, np.zeros(shape)), dtype=np.float64)) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")}
    m = OrcMatrix(*[_pd(ms[k]) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")])
    x = np.zeros(shape, order="F")
    yy = np.asfortranarray(np.array(y, dtype=np.float64))
    d = np.asfortranarray(np.array(dd, dtype=np.float64))
    ic = np.asfortranarray(np.array(indc, dtype=np.int32))
    lib().orc_mzminv(shape[0], shape[1], shape[2], _pd(d), C.byref(m), _pd(x), _pd(yy), _pi(ic))
    return x


def mzax(mat, x, indc):
    shape = mat["AD"].shape
    ms = {k: np.asfortranarray(np.array(mat.get(k, np.zeros(shape)), dtype=np.float64)) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")}
    m = OrcMatrix(*[_pd(ms[k]) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")])
    xx = np.asfortranarray(np.array(x, dtype=np.float64))
    y = np.zeros(shape, order="F")
    ic = np.asfortranarray(np.array(indc, dtype=np.int32))
    lib().orc_mzax(shape[0], shape[1], shape[2], C.byref(m), _pd(xx), _pd(y), _pi(ic))
    return y

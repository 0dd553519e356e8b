"""ctypes mirror of include/cadmas_hip.h (libcadmas_hip.so).

The library is the product path: there is no CPU fallback.  Loading fails
loudly when libcadmas_hip.so has not been built, and every call raises
CadError with the library's message when it returns a negative code.

Arrays follow the reference layout: numpy float64 / int32 in Fortran order,
fields (NUMI, NUMJ, NUMK), metric tables (6, N), INDB (4, NUMB).
"""
import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcadmas_hip.so"

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)

PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_LEARNED = 2

ERRORS = {-1: "ARG", -2: "HIP", -3: "STATE", -4: "UNSUPPORTED", -30: "NOT CONVERGED",
          -31: "INSTABILITY", -32: "CAN NOT DECOMPOSITION", -40: "ORIENTATION", -50: "WAVE TABLE"}


class CadError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("rho0", C.c_double), ("grz0", C.c_double), ("anu0", C.c_double),
                ("dtnow", C.c_double), ("scmvp", C.c_double), ("vvmax", C.c_double),
                ("flower", C.c_double), ("fupper", C.c_double), ("plower", C.c_double),
                ("cgpara", C.c_double), ("cgepsa", C.c_double), ("cgepsr", C.c_double),
                ("cgdiv", C.c_double), ("icgtyp", C.c_int), ("icgmax", C.c_int),
                ("ibsuw0", C.c_int), ("loops", C.c_int), ("precision", C.c_int),
                ("idrop", C.c_int), ("wbub", C.c_double), ("tnow", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepInfo(C.Structure):
    _fields_ = [("cgbnrm", C.c_double), ("cgxnrm", C.c_double), ("fsum", C.c_double),
                ("fcut", C.c_double), ("icgitr", C.c_int), ("nums", C.c_int),
                ("ms_predictor", C.c_float), ("ms_poisson", C.c_float),
                ("ms_correction", C.c_float), ("ms_vof", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


DEFAULTS = dict(rho0=1.0e3, grz0=9.8, anu0=1.0e-6, dtnow=1.0e-3, scmvp=1.0, vvmax=30.0,
                flower=1.0e-6, fupper=1.0 - 1.0e-6, plower=1.0e-4, cgpara=0.95, cgepsa=1.0e-15,
                cgepsr=1.0e-12, cgdiv=1.0e30, icgtyp=1, icgmax=500, ibsuw0=0, loops=1,
                precision=PRECISION_FP64, idrop=1, wbub=0.2, tnow=0.0)

_lib = None


def lib():
    """Load libcadmas_hip.so (raises if it is missing: the HIP path is mandatory)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(str(LIB_PATH))
    L.cad_last_error.restype = C.c_char_p
    L.cad_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.cad_destroy.argtypes = [C.c_void_p]
    L.cad_destroy.restype = None
    L.cad_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
    L.cad_get_params.argtypes = [C.c_void_p, C.POINTER(Params)]
    L.cad_upload_tdoor.argtypes = [C.c_void_p] + [PD] * 7
    L.cad_cdtcal.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, PD]
    L.cad_download_tdoor.argtypes = [C.c_void_p] + [PD] * 7
    L.cad_set_wave_table.argtypes = [C.c_void_p, C.c_int, C.c_int] + [PD] * 5 + [C.c_double]
    L.cad_set_wave_table2.argtypes = [C.c_void_p, C.c_int, C.c_int] + [PD] * 5 + [C.c_double]
    L.cad_set_wave.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_double] * 4
    L.cad_set_wave2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_double] * 4
    L.cad_get_wave_level.argtypes = [C.c_void_p, C.c_int, PD]
    L.cad_set_open.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
    L.cad_set_wave_theory.argtypes = [C.c_void_p] + [C.c_int] * 4 + [C.c_double] * 8
    L.cad_wave_theory_eval.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.POINTER(C.c_float), C.c_double, C.c_int, PD, PD, PD, PD, PD,
                                       PD]
    L.cad_set_drag_df.argtypes = [C.c_void_p, C.c_double, C.c_int, PD, PD, PD]
    L.cad_set_damp.argtypes = [C.c_void_p, C.c_int, C.c_int] + [C.c_double] * 4
    L.cad_poisson_residual.argtypes = [C.c_void_p, PD]
    L.cad_time_precond.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), PD]
    L.cad_upload_static.argtypes = [C.c_void_p] + [PD] * 13 + [PI] * 4 + [PD]
    L.cad_upload_state.argtypes = [C.c_void_p] + [PD] * 10 + [PI] * 2
    L.cad_download_state.argtypes = [C.c_void_p] + [PD] * 10 + [PI] * 2
    for name in ("cad_v1cal", "cad_f1cal", "cad_step", "cad_vpsol"):
        getattr(L, name).argtypes = [C.c_void_p, C.POINTER(StepInfo)]
    for name in ("cad_predictor", "cad_vpcoef", "cad_correction"):
        getattr(L, name).argtypes = [C.c_void_p]
    L.cad_cdiv00.argtypes = [C.c_void_p, PD]
    L.cad_init_state.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.POINTER(StepInfo)]
    L.cad_download_matrix.argtypes = [C.c_void_p] + [PD] * 8
    L.cad_upload_matrix.argtypes = [C.c_void_p] + [PD] * 8 + [PI]
    L.cad_download_pt.argtypes = [C.c_void_p, PD]
    L.cad_time_minv.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float)]
    L.cad_count_active.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.cad_set_sweep_schedule.argtypes = [C.c_void_p, C.c_int]
    L.cad_set_flag_pass_limit.argtypes = [C.c_void_p, C.c_longlong]
    L.cad_set_true_residual_stop.argtypes = [C.c_void_p, C.c_double]
    L.vf_m1bcgs_.argtypes = [PD] * 24 + [PI] * 10
    L.vf_m1bcgs_.restype = None
    L.vf_vpsol_.argtypes = [PD] * 18 + [PI]
    L.vf_vpsol_.restype = None
    L.cad_milu_factor.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.cad_milu_apply.argtypes = [C.c_void_p, PD, PD]
    L.cad_set_krylov_cap.argtypes = [C.c_void_p, C.c_int]
    L.cad_mixed_stats.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.cad_time_vcycle.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float), PD]
    L.cad_time_learned.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), PD]
    L.cad_unet_mac_per_voxel.restype = C.c_double
    L.cad_set_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    L.cad_kernel_timing.argtypes = [C.c_void_p, C.c_int, PD, PD, C.POINTER(C.c_longlong)]
    L.cad_m1bcgs_noalloc.argtypes = ([C.c_void_p] * 19 + [C.c_int] * 3 + [C.c_double] * 4 +
                                     [C.c_int] * 2 + [PD, PD, PI])
    _lib = L
    return L


EXPORTS = ("cad_last_error", "cad_create", "cad_destroy", "cad_set_params", "cad_upload_static",
           "cad_upload_state", "cad_download_state", "cad_v1cal", "cad_f1cal", "cad_step",
           "cad_cdiv00", "cad_predictor", "cad_vpcoef", "cad_vpsol", "cad_correction",
           "cad_download_matrix", "cad_upload_matrix", "cad_download_pt", "vf_m1bcgs_", "vf_vpsol_",
           "cad_init_state", "cad_time_minv", "cad_count_active", "cad_set_sweep_schedule",
           "cad_milu_factor", "cad_milu_apply", "cad_get_params", "cad_upload_tdoor",
           "cad_download_tdoor", "cad_cdtcal", "cad_set_wave_table", "cad_set_wave",
           "cad_get_wave_level", "cad_set_open", "cad_poisson_residual", "cad_time_precond",
           "cad_set_drag_df", "cad_set_damp", "cad_set_wave_theory", "cad_wave_theory_eval",
           "cad_set_wave_table2", "cad_set_wave2", "cad_set_flag_pass_limit",
           "cad_set_true_residual_stop", "cad_m1bcgs_noalloc", "cad_set_krylov_cap",
           "cad_mixed_stats", "cad_time_vcycle", "cad_set_learned_weights", "cad_set_learned_mode", "cad_learned_stats",
           "cad_learned_box", "cad_learned_box_matrix", "cad_learned_sample", "cad_time_learned", "cad_unet_param_count",
           "cad_unet_mac_per_voxel", "cad_unet_init_params", "cad_unet_create", "cad_unet_destroy",
           "cad_unet_set_params", "cad_unet_get_params", "cad_unet_set_input", "cad_unet_forward",
           "cad_unet_get_output", "cad_unet_set_depth", "cad_unet_train_step", "cad_unet_train_step_dy", "cad_unet_gradients_dy",
           "cad_unet_gradients", "cad_set_kernel_timing", "cad_kernel_timing")

# the forms' dominant kernels for Context.kernel_timing
KERNEL_MINV, KERNEL_VCYCLE, KERNEL_UNET = 0, 1, 2


class TheoryWave:
    """A theory wave evaluated on the host (cad_wave_theory_eval), with
    VF_STK0's carried state: each call() is one VF_CWMAK0 of the wave."""

    def __init__(self, func, depth, height, period):
        self.func = WAVE_FUNC[func] if isinstance(func, str) else int(func)
        self.depth, self.height, self.period = float(depth), float(height), float(period)
        self.state = C.c_float(0.0)

    def call(self, wvt=0.0, zc=(), origin=False):
        """(level at wvt, un, ut at heights zc, phase origin or None); the
        wave length of the call is kept as .length."""
        z = np.ascontiguousarray(np.asarray(zc, dtype=np.float64).reshape(-1))
        un, ut = np.zeros_like(z), np.zeros_like(z)
        lv, t0, wl = C.c_double(), C.c_double(), C.c_double()
        _check(lib().cad_wave_theory_eval(self.func, self.depth, self.height, self.period,
                                          C.byref(self.state), float(wvt), len(z), _pd(z), _pd(un),
                                          _pd(ut), C.byref(lv), C.byref(t0) if origin else None,
                                          C.byref(wl)))
        self.length = wl.value
        return lv.value, un, ut, (t0.value if origin else None)


# MODEL WAVE-BC FUNC words of the theory waves -> cad_set_wave_theory's func
# (IBCTYP(2,side), Source code/vf_iimdl.f:52-63); FUNC STREAM n gives n
WAVE_FUNC = {"STK-CND": 0, "CNOIDAL": -1, "STOKES": -2}


def wave_func(wv):
    """cad_set_wave_theory's func for a deck's wave side (deck.waves[side])."""
    if wv["func"] == "STREAM":
        return int(wv["order"])
    return WAVE_FUNC[wv["func"]]

# schedules of the MILU triangular sweeps (bit-identical results)
SWEEP_PIPELINED = 0
SWEEP_HYPERPLANE = 1
SWEEP_PIPELINED_8X8 = 2
SWEEP_PIPELINED_16X4 = 3

INIT_VELOCITY = 1
INIT_PRESSURE = 2


def _check(rc):
    if rc != 0:
        raise CadError(rc, lib().cad_last_error().decode(errors="replace"))


def _f64(a, shape=None):
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _i32(a, shape=None):
    a = np.asfortranarray(np.asarray(a, dtype=np.int32))
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


MATRIX = ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")
STATE_F = ("UU", "VV", "WW", "PP", "FF")
STATE_B = ("BCU", "BCV", "BCW", "BCP", "BCF")
DROP = ("DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW")


class Context:
    """One grid resident in HBM (cad_ctx)."""

    def __init__(self, numi, numj, numk, numb, device=0):
        self.shape = (int(numi), int(numj), int(numk))
        self.numb = int(numb)
        h = C.c_void_p()
        _check(lib().cad_create(numi, numj, numk, numb, device, C.byref(h)))
        self._h = h
        self.params = dict(DEFAULTS)

    def close(self):
        if getattr(self, "_h", None):
            lib().cad_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- set-up
    def set_params(self, **kw):
        """Change scalars; the others keep the context's current values
        (TNOW as cad_step left it)."""
        p = self.get_params()
        for k, v in kw.items():
            if k not in p:
                raise KeyError(k)
            p[k] = v
        st = Params(**p)
        _check(lib().cad_set_params(self._h, C.byref(st)))
        self.params = p

    def get_params(self):
        st = Params()
        _check(lib().cad_get_params(self._h, C.byref(st)))
        return st.as_dict()

    def upload_tdoor(self, a):
        """TimerDoor state: TBUB (NUMK) and DROPTX..DROPWW fields."""
        tb = _f64(a["TBUB"], (self.shape[2],))
        fl = [_f64(a[k], self.shape) for k in DROP]
        _check(lib().cad_upload_tdoor(self._h, _pd(tb), *[_pd(x) for x in fl]))

    def download_tdoor(self):
        out = {"TBUB": np.zeros(self.shape[2])}
        out.update({k: np.zeros(self.shape, order="F") for k in DROP})
        _check(lib().cad_download_tdoor(self._h, _pd(out["TBUB"]), *[_pd(out[k]) for k in DROP]))
        return out

    def set_waves(self, waves, table, wvlvl, table2=None):
        """Wave-makers: waves = deck.waves (side -> func, depth, height,
        period, ampl, area); table = mtb.Table for the sides with FUNC MATRIX
        and table2 for those with MATRIX2 (None when there are none).  Sides
        are set in side order, as the reference's set-up visits them."""
        for tab, fn, which in ((table, "MATRIX", 0), (table2, "MATRIX2", 1)):
            if not any(wv["func"] == fn for wv in waves.values()):
                continue
            zz, tt, hh = (_f64(getattr(tab, k)) for k in ("zz", "tt", "hh"))
            nz, nt = len(zz), len(tt)
            un, ut = _f64(tab.un, (nz, nt)), _f64(tab.ut, (nz, nt))
            # a LEVEL-OFF table goes without its level (a null level array), a
            # LEVEL-ONLY table without depths and profiles (nz = 0)
            hp = None if tab.kind == "LEVEL-OFF" else _pd(hh)
            f = lib().cad_set_wave_table2 if which else lib().cad_set_wave_table
            if tab.kind == "LEVEL-ONLY":
                _check(f(self._h, 0, nt, None, _pd(tt), hp, None, None, float(tab.t0)))
            else:
                _check(f(self._h, nz, nt, _pd(zz), _pd(tt), hp, _pd(un), _pd(ut), float(tab.t0)))
        ni, nj, _ = self.shape
        for side, wv in sorted(waves.items()):
            lim = nj if side <= 2 else ni
            l1, l2 = wv["area"] if wv["area"] else (2, lim - 1)
            if wv["func"] in ("MATRIX", "MATRIX2"):
                f = lib().cad_set_wave2 if wv["func"] == "MATRIX2" else lib().cad_set_wave
                _check(f(self._h, int(side), int(l1), int(l2), float(wv["depth"]),
                         float(wv["period"]), float(wv["ampl"]), float(wvlvl)))
            else:
                _check(lib().cad_set_wave_theory(
                    self._h, int(side), wave_func(wv), int(l1), int(l2), float(wv["depth"]),
                    float(wv["height"]), float(wv["period"]), float(wv["ampl"]), float(wvlvl),
                    float(wv["angle"]), float(wv.get("ax", 0.0)), float(wv.get("ay", 0.0))))

    def set_opens(self, opens):
        """Radiation sides: opens = deck.opens (side -> depth, period)."""
        for side, op in sorted(opens.items()):
            _check(lib().cad_set_open(self._h, int(side), float(op["depth"]), float(op["period"])))

    def set_drag_df(self, drag):
        """OPTION DRAG-DF: drag = (YU, [(DR, AP, BT), ...]) as deck.drag_df,
        or None for the CD drag form."""
        if drag is None:
            _check(lib().cad_set_drag_df(self._h, 1.0, 0, None, None, None))
            return
        yu, rows = drag
        t = np.ascontiguousarray(np.array(rows, dtype=np.float64).T)
        n = t.shape[1]
        _check(lib().cad_set_drag_df(self._h, float(yu), n, _pd(t[0]), _pd(t[1]), _pd(t[2])))

    def set_damps(self, damps):
        """Damping zones: damps = deck.damps (side -> degree, pxy, pz, width,
        depth); sides not named have none."""
        for side in range(1, 5):
            dm = damps.get(side)
            if dm is None:
                _check(lib().cad_set_damp(self._h, side, -1, 0.0, 0.0, 0.0, 0.0))
            else:
                _check(lib().cad_set_damp(self._h, side, int(dm["degree"]), float(dm["pxy"]),
                                          float(dm["pz"]), float(dm["width"]), float(dm["depth"])))

    def time_precond(self, reps=5):
        """(ms, algorithmic bytes) of one preconditioner application of the
        context's form after a solve."""
        ms, by = C.c_float(), C.c_double()
        _check(lib().cad_time_precond(self._h, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def poisson_residual(self):
        """||BB - A PT|| / ||BB|| of the last pressure solve (FP64)."""
        v = C.c_double()
        _check(lib().cad_poisson_residual(self._h, C.byref(v)))
        return v.value

    def wave_level(self, side):
        """The wave-maker level of a side at the last boundary pass."""
        v = C.c_double()
        _check(lib().cad_get_wave_level(self._h, int(side), C.byref(v)))
        return v.value

    def upload_static(self, a):
        s, nb = self.shape, self.numb
        ni, nj, nk = s
        keep = [_f64(a["XX"], (6, ni)), _f64(a["YY"], (6, nj)), _f64(a["ZZ"], (6, nk))]
        keep += [_f64(a.get("ANU", np.zeros(s)), s), _f64(a.get("CD0", np.zeros(s)), s)]
        keep += [_f64(a[k], s) for k in ("GGV", "GGX", "GGY", "GGZ", "GLV", "GLX", "GLY", "GLZ")]
        ints = [_i32(a[k], s) for k in ("INDX", "INDY", "INDZ")]
        indb = _i32(a["INDB"], (4, nb)) if nb else np.zeros((4, 1), np.int32, order="F")
        bcvi = _f64(a.get("BCVI", np.zeros(max(nb, 1)))) if nb else np.zeros(1)
        _check(lib().cad_upload_static(self._h, *[_pd(x) for x in keep], *[_pi(x) for x in ints],
                                       _pi(indb), _pd(bcvi)))

    def upload_state(self, a):
        s, nb = self.shape, max(self.numb, 1)
        fl = [_f64(a[k], s) for k in STATE_F]
        bc = [_f64(a[k]) if self.numb else np.zeros(nb) for k in STATE_B]
        ints = [_i32(a["NF"], s), _i32(a["INDC"], s)]
        _check(lib().cad_upload_state(self._h, *[_pd(x) for x in fl + bc], *[_pi(x) for x in ints]))

    def download_state(self):
        s, nb = self.shape, max(self.numb, 1)
        out = {k: np.zeros(s, order="F") for k in STATE_F}
        out.update({k: np.zeros(nb) for k in STATE_B})
        out["NF"] = np.zeros(s, dtype=np.int32, order="F")
        out["INDC"] = np.zeros(s, dtype=np.int32, order="F")
        _check(lib().cad_download_state(self._h, *[_pd(out[k]) for k in STATE_F + STATE_B],
                                        _pi(out["NF"]), _pi(out["INDC"])))
        if not self.numb:
            for k in STATE_B:
                out[k] = out[k][:0]
        return out

    # ---------------------------------------------------------------- phases
    def _info_call(self, name):
        info = StepInfo()
        _check(getattr(lib(), name)(self._h, C.byref(info)))
        return info.as_dict()

    def init_state(self, wvlvl=0.0, uvw=(0.0, 0.0, 0.0), flags=INIT_VELOCITY | INIT_PRESSURE):
        """Device part of VF_CINIT on the uploaded state (see cad_init_state)."""
        info = StepInfo()
        _check(lib().cad_init_state(self._h, wvlvl, uvw[0], uvw[1], uvw[2], flags, C.byref(info)))
        return info.as_dict()

    def step(self):
        return self._info_call("cad_step")

    def cdtcal(self, nnow, dtinit, dtsafe, dtmin, dtmax):
        """VF_CDTCAL: the next step size from the resident fields."""
        dt = C.c_double()
        _check(lib().cad_cdtcal(self._h, int(nnow), dtinit, dtsafe, dtmin, dtmax, C.byref(dt)))
        return dt.value

    def v1cal(self):
        return self._info_call("cad_v1cal")

    def f1cal(self):
        return self._info_call("cad_f1cal")

    def vpsol(self):
        return self._info_call("cad_vpsol")

    def predictor(self):
        _check(lib().cad_predictor(self._h))

    def vpcoef(self):
        _check(lib().cad_vpcoef(self._h))

    def correction(self):
        _check(lib().cad_correction(self._h))

    def cdiv00(self):
        v = C.c_double()
        _check(lib().cad_cdiv00(self._h, C.byref(v)))
        return v.value

    def download_matrix(self):
        out = {k: np.zeros(self.shape, order="F") for k in MATRIX}
        _check(lib().cad_download_matrix(self._h, *[_pd(out[k]) for k in MATRIX]))
        return out

    def upload_matrix(self, m, indc):
        arrs = [_f64(m[k], self.shape) for k in MATRIX]
        ic = _i32(indc, self.shape)
        _check(lib().cad_upload_matrix(self._h, *[_pd(a) for a in arrs], _pi(ic)))

    def time_minv(self, reps=5):
        ms = C.c_float()
        _check(lib().cad_time_minv(self._h, reps, C.byref(ms)))
        return ms.value

    def set_flag_pass_limit(self, passes):
        """Cap on VF_FNFINI's flag-fix passes (0: the default bound)."""
        _check(lib().cad_set_flag_pass_limit(self._h, int(passes)))

    def set_true_residual_stop(self, rel):
        """Plain form: also stop only once ||BB - A PT|| / ||BB|| <= rel (0: the
        reference's stop test alone)."""
        _check(lib().cad_set_true_residual_stop(self._h, float(rel)))

    def set_krylov_cap(self, cap):
        """Iterations of the mixed / learned forms before MILU-BiCGSTAB takes
        over the solve (0: the default, 200)."""
        _check(lib().cad_set_krylov_cap(self._h, int(cap)))

    def mixed_stats(self):
        """The mixed form's fallbacks so far and the last one's iterations."""
        a = (C.c_longlong * 2)()
        _check(lib().cad_mixed_stats(self._h, a))
        return {"fallbacks": int(a[0]), "fallback_iters": int(a[1])}

    def time_vcycle(self, reps=3, fp64=False):
        """(ms, algorithmic bytes) of one multigrid V-cycle, FP32 or FP64."""
        ms, by = C.c_float(), C.c_double()
        _check(lib().cad_time_vcycle(self._h, int(reps), int(bool(fp64)), C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def time_learned(self, reps=3):
        """(ms, FLOP) of one forward pass of the learned form's network."""
        ms, fl = C.c_float(), C.c_double()
        _check(lib().cad_time_learned(self._h, int(reps), C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def set_kernel_timing(self, every):
        """Start (clearing the totals) or stop (every = 0 or False) the in-solve
        timing of the forms' dominant kernels: HIP events around every
        `every`-th launch inside the solves (True: each launch)."""
        _check(lib().cad_set_kernel_timing(self._h, int(every)))

    def kernel_timing(self, which):
        """(average ms, average algorithmic work, timed launches) of kernel
        `which` (KERNEL_MINV, KERNEL_VCYCLE, KERNEL_UNET) inside the solves
        since the timing was turned on; work is bytes for the V-cycle, FLOP
        for the network pass and 0 for M^-1."""
        ms, wk, n = C.c_double(), C.c_double(), C.c_longlong()
        _check(lib().cad_kernel_timing(self._h, int(which), C.byref(ms), C.byref(wk), C.byref(n)))
        return ms.value, wk.value, n.value

    def set_sweep_schedule(self, schedule):
        _check(lib().cad_set_sweep_schedule(self._h, int(schedule)))

    def milu_factor(self):
        """MILU factorisation of the uploaded matrix; returns npos (0: a pivot <= 0)."""
        n = C.c_int()
        _check(lib().cad_milu_factor(self._h, C.byref(n)))
        return n.value

    def milu_apply(self, y):
        """x = M^-1 y on the current factorisation (host arrays)."""
        yy = _f64(y, self.shape)
        x = np.zeros(self.shape, order="F")
        _check(lib().cad_milu_apply(self._h, _pd(yy), _pd(x)))
        return x

    def count_active(self):
        n = C.c_longlong()
        _check(lib().cad_count_active(self._h, C.byref(n)))
        return n.value

    def download_pt(self):
        pt = np.zeros(self.shape, order="F")
        _check(lib().cad_download_pt(self._h, _pd(pt)))
        return pt


class AcomprBlock(C.Structure):
    """COMMON /VF_ACOMPR/ (VF_ACOMPR.h)."""
    _fields_ = [(k, C.c_double) for k in ("cgpara", "cgepsa", "cgepsr", "cgdiv", "cgbnrm", "cgxnrm",
                                          "feps", "flower", "fupper", "fsum", "fcut", "plower",
                                          "scmvp", "scmk", "scmt")] + [("scmc", C.c_double * 10)]


class AcompiBlock(C.Structure):
    """COMMON /VF_ACOMPI/ (VF_ACOMPI.h)."""
    _fields_ = [(k, C.c_int) for k in ("icgtyp", "icgmax", "icgitr", "iscmvp", "iscmff", "iscmk",
                                       "iscmt")] + [("iscmc", C.c_int * 10), ("ibsuw0", C.c_int)]


class AnumbiBlock(C.Structure):
    """COMMON /VF_ANUMBI/ (VF_ANUMBI.h)."""
    _fields_ = [(k, C.c_int) for k in ("numi", "numj", "numk", "numb", "nums", "leqk", "leqt",
                                       "leqc", "numij", "numbx")]


def common_blocks():
    """The library's COMMON blocks /VF_ACOMPR/, /VF_ACOMPI/, /VF_ANUMBI/ (its
    own instances when no Fortran program supplies them)."""
    L = lib()
    return (AcomprBlock.in_dll(L, "vf_acompr_"), AcompiBlock.in_dll(L, "vf_acompi_"),
            AnumbiBlock.in_dll(L, "vf_anumbi_"))


def _work(shape):
    return [np.zeros(shape, order="F") for _ in range(8)]


def m1bcgs(mat, indc, x0=None, cgpara=0.95, cgepsa=1.0e-15, cgepsr=1.0e-12, cgdiv=1.0e30,
           icgtyp=1, icgmax=500, loop=None):
    """VF_M1BCGS through the drop-in vf_m1bcgs_ with the reference's full
    argument list (vf_m1bcgs.f:1-5).  The arrays' shape is (MI,MJ,MK); loop =
    (NI,NJ,NK), the loop range, defaults to the shape.  Returns X0, BNORM,
    XNORM, ITER."""
    L = lib()
    shape = mat["AD"].shape
    nijk = tuple(shape if loop is None else loop)
    ms = [_f64(mat[k], shape) for k in MATRIX]
    x = _f64(np.zeros(shape) if x0 is None else x0, shape).copy(order="F")
    ic = _i32(indc, shape)
    d = lambda v: C.byref(C.c_double(v))  # noqa: E731
    n = lambda v: C.byref(C.c_int(int(v)))  # noqa: E731
    bn, xn, it = C.c_double(), C.c_double(), C.c_int()
    dbuf = np.zeros(1)
    wk = _work(shape)
    L.vf_m1bcgs_(_pd(dbuf), *[_pd(a) for a in ms], _pd(x), *[_pd(w) for w in wk], d(cgpara),
                 d(cgepsa), d(cgepsr), d(cgdiv), C.byref(bn), C.byref(xn), _pi(ic), n(shape[0]),
                 n(shape[1]), n(shape[2]), n(nijk[0]), n(nijk[1]), n(nijk[2]), n(icgtyp),
                 n(icgmax), C.byref(it))
    return x, bn.value, xn.value, it.value


def vpsol(mat, indc, cgpara=0.95, cgepsa=1.0e-15, cgepsr=1.0e-12, cgdiv=1.0e30, icgtyp=1,
          icgmax=500):
    """VF_VPSOL through the drop-in vf_vpsol_ (vf_vpsol.f:1-3): sizes and
    solver parameters go in through the COMMON blocks, CGBNRM, CGXNRM and
    ICGITR come back there.  Returns PT, CGBNRM, CGXNRM, ICGITR."""
    L = lib()
    shape = mat["AD"].shape
    pr, pi, nb = common_blocks()
    nb.numi, nb.numj, nb.numk = shape
    pr.cgpara, pr.cgepsa, pr.cgepsr, pr.cgdiv = cgpara, cgepsa, cgepsr, cgdiv
    pi.icgtyp, pi.icgmax, pi.icgitr = icgtyp, icgmax, 0
    pr.cgbnrm = pr.cgxnrm = 0.0
    ms = [_f64(mat[k], shape) for k in MATRIX]
    pt = np.full(shape, np.nan, order="F")      # VF_VPSOL zeroes it first
    ic = _i32(indc, shape)
    dbuf = np.zeros(1)
    wk = _work(shape)
    L.vf_vpsol_(_pd(dbuf), *[_pd(a) for a in ms], _pd(pt), *[_pd(w) for w in wk], _pi(ic))
    return pt, pr.cgbnrm, pr.cgxnrm, pi.icgitr


NOALLOC_MAX_CELLS = 2097152


def m1bcgs_noalloc(ptrs, shape, stream=None, cgpara=0.95, cgepsa=1.0e-15, cgepsr=1.0e-12,
                   cgdiv=1.0e30, icgtyp=1, icgmax=500):
    """cad_m1bcgs_noalloc: VF_M1BCGS on arrays already on the GPU, allocating
    nothing.  ptrs maps AD, ALI, ALJ, ALK, AUI, AUJ, AUK, BB, X0, INDC and the
    work arrays DD, XW, YY, R0, RR, PP, AP, AE to device addresses of
    (NI,NJ,NK) column-major arrays; shape = (NI,NJ,NK); stream a hipStream_t
    (None: the null stream).  X0 holds the solution on return.  Returns
    (rc, BNORM, XNORM, ITER); rc is CAD_OK or a negative code (not raised, so
    that callers can check the refusals)."""
    L = lib()
    names = list(MATRIX) + ["X0", "DD", "XW", "YY", "R0", "RR", "PP", "AP", "AE", "INDC"]
    bn, xn, it = C.c_double(), C.c_double(), C.c_int()
    rc = L.cad_m1bcgs_noalloc(stream, *[C.c_void_p(int(ptrs[k])) for k in names],
                              int(shape[0]), int(shape[1]), int(shape[2]), cgpara, cgepsa, cgepsr,
                              cgdiv, int(icgtyp), int(icgmax), C.byref(bn), C.byref(xn),
                              C.byref(it))
    return rc, bn.value, xn.value, it.value

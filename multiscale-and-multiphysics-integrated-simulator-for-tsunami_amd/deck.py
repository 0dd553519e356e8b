"""Reader for the CADMAS-SURF/3D input deck (<name>.in), the subset the
single-phase VOF-SMAC path uses.

Grammar: Simulators/CADMAS-SURF-3D/Source code/A11111.in; dispatch
vf_ii1inp.f:174-413; tokenising vf_zgetln.f ('#' starts a comment, words are
separated by blanks).  Keywords outside the implemented path raise
UnsupportedDeck instead of being ignored, so a deck that would make the
reference run physics this build does not have is refused, not mis-run.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple


class DeckError(ValueError):
    pass


class UnsupportedDeck(DeckError):
    pass


def zstoi(text):
    """VF_ZSTOI (Source code/vf_zstoi.f): optional sign, then decimal digits
    up to the first other character."""
    t = text.lstrip(" ")
    sg = 1
    if t[:1] == "+":
        t = t[1:]
    elif t[:1] == "-":
        sg, t = -1, t[1:]
    v = 0
    for ch in t:
        if not ch.isdigit():
            break
        v = 10 * v + int(ch)
    return sg * v


def zstor(text):
    """VF_ZSTOR (Source code/vf_zstor.f), the reference's reading of a real
    number, restated operation by operation: the digits accumulate as
    VAL = 10*VAL + d, the decimals count into PW, VAL = SG*(VAL/PW), then an
    exponent multiplies by 10.0D0**I.  It is not the correctly rounded
    decimal conversion: "2.30101890e-01" is read as fl(2.3010189)*fl(0.1),
    so the deck and table values must be read this way for the reference's
    bits."""
    t = text.lstrip(" ")
    if not t:
        return 0.0
    val, sg, pw = 0.0, 1.0, 1.0
    if t[0] == "+":
        t = t[1:]
    elif t[0] == "-":
        sg, t = -1.0, t[1:]
    point = False
    expo = None
    for n, ch in enumerate(t):
        if ch.isdigit():
            val = 10.0 * val + float(ch)
            if point:
                pw = pw * 10.0
        elif ch == "." and not point:
            point = True
        elif ch in "DdEe":
            expo = t[n + 1:]
            break
        else:
            break
    val = sg * (val / pw)
    if expo:
        val = val * (10.0 ** zstoi(expo))
    return val


def _num(w):
    if not w or not (w.lstrip("+-")[:1].isdigit() or w.lstrip("+-")[:1] == "."):
        raise DeckError(f"not a number: {w!r}")
    return zstor(w)


def _int(w):
    try:
        return int(w)
    except ValueError as e:
        raise DeckError(f"not an integer: {w!r}") from e


@dataclass
class BoundaryLine:
    """One B.C. line: kind D/X/Y/Z, a face box (1-based as in the deck, None
    for D), physics 'VP' or 'F', type code and values (vf_iiboun.f:143-260)."""
    kind: str
    box: Optional[Tuple[int, int, int, int, int, int]]
    phys: str
    code: int
    values: Tuple[float, ...]


@dataclass
class Deck:
    grid: dict = field(default_factory=dict)              # 'X','Y','Z' -> list of coords
    obst: List[Tuple[int, int, int, int, int, int]] = field(default_factory=list)
    porous: List[tuple] = field(default_factory=list)     # (kind, box, values)
    plower: float = 1.0e-4
    bcs: List[BoundaryLine] = field(default_factory=list)
    # TIME (vf_iitime.f, defaults vf_a2dflt.f:223-234): IDTTYP 0 = CONST,
    # 1 = AUTO (the default) with DTINIT, DTSAFE and LIMIT DTMIN, DTMAX
    dt_const: Optional[float] = None
    idttyp: int = 1
    dtinit: float = 1.0e-6
    dtsafe: float = 1.0e-1
    dtmin: float = 1.0e-20
    dtmax: float = 1.0e20
    nend: int = 0
    tend: float = 0.0
    # MATE (vf_iimate.f, defaults vf_a2dflt.f:161-164)
    wvlvl: float = 0.0
    rho0: float = 1.0e3
    anu0: float = 1.0e-6
    grz0: float = 9.8
    uini: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    # COMP (vf_iicomp.f, defaults vf_a2dflt.f:37-60)
    scmvp: float = 1.0
    icgtyp: int = 1
    cgpara: float = 0.95
    icgmax: int = 500
    cgepsa: float = 1.0e-15
    cgepsr: float = 1.0e-12
    # OPTION (vf_iiopt.f, defaults vf_a2dflt.f)
    wbub: float = 0.2
    idrop: int = 1
    ibsuw0: int = 0
    loops: int = 1
    vvmax: float = 30.0
    # FILE RSL STEP n1 n2 dn / TIME t1 t2 dt (vf_iifile.f:209-238) and
    # FILE RES step (:241-244)
    rsl: Optional[tuple] = None
    res_step: Optional[int] = None
    # FILE GRP STEP/TIME and AREA (cell indices +1 like IGRARA)
    grp: Optional[tuple] = None
    grp_area: Optional[tuple] = None
    # MODEL WAVE-BC per side (1 X-, 2 X+, 3 Y-, 4 Y+): func, depth, height,
    # period, ampl, area (+1 like IBCTYP(3..4)), angle
    waves: dict = field(default_factory=dict)
    # MODEL OPEN-BC per side: radiation boundary (FUNC TYPE1) with the
    # depth and period that give its celerity
    opens: dict = field(default_factory=dict)
    # OPTION DRAG-DF (vf_iiopt.f:86-108): (DRGYU, [(DRGDR, DRGAP, DRGBT),
    # ...]) with at most 10 rows; None keeps the CD drag form
    drag_df: Optional[tuple] = None
    # MODEL DAMP per side (1 X-, 2 X+, 3 Y-, 4 Y+): the implicit sponge of
    # VF_VEULER with DEGREE (IDAMP), PARAM-XY, PARAM-Z, WIDTH and DEPTH
    # (DAMP(1..4,side)); defaults of vf_a2dflt.f:150,189-204
    damps: dict = field(default_factory=dict)
    # FILE TRN STEP/TIME and its probe lines (words after FILE TRN)
    trn: Optional[tuple] = None
    trn_probes: List[tuple] = field(default_factory=list)
    # DEBUG F-BOX (vf_iidbg.f; cell indices already +1 like IDBGF)
    fbox: Optional[Tuple[int, int, int, int, int, int, float]] = None

    @property
    def ncells(self):
        return tuple(len(self.grid[a]) - 1 for a in "XYZ")


# output controls not produced here: list-file echo (L/P), observation
# points (OBS)
_IGNORED_FILE = {"L/P", "OBS"}


def parse(text):
    d = Deck()
    lines = []
    for raw in text.splitlines():
        words = raw.split("#", 1)[0].split()
        if words:
            lines.append(words)
    i = 0
    while i < len(lines):
        w = lines[i]
        head = w[0]
        i += 1
        if head == "GRID":
            if len(w) < 2 or w[1] not in ("X", "Y", "Z"):
                raise DeckError("GRID needs X, Y or Z")
            axis = w[1]
            if axis in d.grid:
                raise DeckError(f"GRID {axis} given twice")
            vals, words = [], w[2:]
            while True:
                done = False
                for t in words:
                    if t == "END":
                        done = True
                        break
                    vals.append(_num(t))
                if done:
                    break
                if i >= len(lines):
                    raise DeckError("GRID: END not found")
                words = lines[i]
                i += 1
            if len(vals) < 2:
                raise DeckError("GRID needs at least two coordinates")
            for a, b in zip(vals, vals[1:]):
                if b - a < 1.0e-6:
                    raise DeckError("GRID coordinates must increase by at least 1e-6")
            d.grid[axis] = vals
        elif head == "OBST":
            if len(w) >= 2 and w[1] == "FILE":
                raise UnsupportedDeck("OBST FILE is not supported")
            if len(w) < 7:
                raise DeckError("OBST needs 6 cell indices")
            d.obst.append(tuple(_int(x) for x in w[1:7]))
        elif head == "POROUS":
            kind = w[1] if len(w) > 1 else ""
            if kind == "LIM":
                d.plower = _num(w[2])
            elif kind in ("V", "X", "Y", "Z"):
                if len(w) < 10:
                    raise DeckError(f"POROUS {kind} needs 6 indices and 2 values")
                d.porous.append((kind, tuple(_int(x) for x in w[2:8]), (_num(w[8]), _num(w[9]))))
            elif kind in ("CM", "CD"):
                if len(w) < 9:
                    raise DeckError(f"POROUS {kind} needs 6 indices and a value")
                d.porous.append((kind, tuple(_int(x) for x in w[2:8]), (_num(w[8]),)))
            else:
                raise UnsupportedDeck(f"POROUS {kind} is not supported")
        elif head == "B.C.":
            d.bcs.append(_parse_bc(w))
        elif head == "TIME":
            _parse_time(d, w)
        elif head == "MATE":
            key = w[1]
            if key == "W-LEVEL":
                d.wvlvl = _num(w[2])
            elif key == "DENSITY":
                d.rho0 = _num(w[2])
            elif key == "K-VISC":
                d.anu0 = _num(w[2])
            elif key == "GRAVITY":
                d.grz0 = _num(w[2])
            elif key == "I.C." and len(w) >= 6 and w[2] == "V":
                d.uini = (_num(w[3]), _num(w[4]), _num(w[5]))
            else:
                raise UnsupportedDeck(f"MATE {' '.join(w[1:3])} is not supported")
        elif head == "COMP":
            _parse_comp(d, w)
        elif head == "OPTION" and len(w) >= 2 and w[1] == "DRAG-DF":
            i = _parse_drag_df(d, w, lines, i)
        elif head == "OPTION":
            _parse_option(d, w)
        elif head == "DEBUG":
            if w[1] == "F-BOX":
                b = tuple(_int(x) + 1 for x in w[2:8])
                d.fbox = b + (_num(w[8]),)
            else:
                raise UnsupportedDeck(f"DEBUG {w[1]} is not supported")
        elif head == "FILE":
            _parse_file(d, w)
        elif head == "EQUATION":
            if len(w) >= 3 and w[2] == "NOCALC":
                continue
            if len(w) >= 3 and w[1] == "CONCENTRATION" and _int(w[2]) == 0:
                continue
            raise UnsupportedDeck("k-epsilon, temperature and concentration equations are not implemented")
        elif head == "PARALLEL":
            raise UnsupportedDeck("PARALLEL: one device runs the whole grid")
        elif head == "MODEL":
            _parse_model(d, w)
        else:
            raise DeckError(f"unknown header {head}")
    for a in "XYZ":
        if a not in d.grid:
            raise DeckError(f"GRID {a} missing")
    _validate(d)
    return d


def _parse_bc(w):
    kind = w[1]
    if kind == "D":
        box, ip = None, 2
    elif kind in ("X", "Y", "Z"):
        box, ip = tuple(_int(x) for x in w[2:8]), 8
    else:
        raise DeckError(f"B.C. {kind}")
    phys = w[ip]
    t = w[ip + 1] if len(w) > ip + 1 else ""
    if phys == "VP":
        if t == "SLIP":
            return BoundaryLine(kind, box, "VP", 1, (0.0, 0.0, 0.0, 0.0, 0.0))
        if t == "NON-S":
            return BoundaryLine(kind, box, "VP", 2, (0.0, 0.0, 0.0, 0.0, 0.0))
        if t == "FIX-V":
            bu, bv, bw = (_num(x) for x in w[ip + 2:ip + 5])
            return BoundaryLine(kind, box, "VP", 3, (bu, bv, bw, 0.0, 0.0))
        if t == "FREE":
            return BoundaryLine(kind, box, "VP", 4, (0.0, 0.0, 0.0, 0.0, 0.0))
        raise UnsupportedDeck(f"B.C. VP {t} (log-law walls) is not implemented")
    if phys == "F":
        if t == "FIX":
            bf = _num(w[ip + 2])
            if not 0.0 <= bf <= 1.0:
                raise DeckError("B.C. F FIX value outside [0,1]")
            return BoundaryLine(kind, box, "F", 1, (bf,))
        if t == "FREE":
            return BoundaryLine(kind, box, "F", 2, (0.0,))
        raise DeckError(f"B.C. F {t}")
    if phys in ("K", "E", "T", "C"):
        return BoundaryLine(kind, box, phys, 0, ())   # equations not solved: no effect
    raise DeckError(f"B.C. {phys}")


def _parse_comp(d, w):
    if w[1] == "SCHM":
        if w[2] == "VP-DONOR":
            d.scmvp = _num(w[3])
        elif w[2] == "FF-DN-AC":
            pass
        elif w[2] == "FF-SLOPE":
            raise UnsupportedDeck("COMP SCHM FF-SLOPE is not implemented")
        elif w[2] in ("KE-DONOR", "T-DONOR", "C-DONOR"):
            pass
        else:
            raise DeckError(f"COMP SCHM {w[2]}")
    elif w[1] == "MTRX":
        if w[2] == "ILUBCGSTAB":
            d.icgtyp = 0
        elif w[2] == "M-ILUBCGSTAB":
            d.icgtyp = 1
            d.cgpara = _num(w[3])
        elif w[2] == "MAX-ITR":
            d.icgmax = _int(w[3])
        elif w[2] == "A-ERROR":
            d.cgepsa = _num(w[3])
        elif w[2] == "R-ERROR":
            d.cgepsr = _num(w[3])
        else:
            raise DeckError(f"COMP MTRX {w[2]}")
    else:
        raise DeckError(f"COMP {w[1]}")


_WAVE_SIDE = {"X-": 1, "X+": 2, "Y-": 3, "Y+": 4}


def _parse_open(d, w):
    """MODEL OPEN-BC side FUNC TYPE1 / DEPTH d / PERIOD t
    (Source code/vf_iimdl.f:174-210): a radiation boundary on the side."""
    if len(w) < 3:
        raise DeckError("MODEL OPEN-BC: SYNTAX ERROR.")
    if w[2] not in _WAVE_SIDE:
        raise DeckError(f"MODEL OPEN-BC {w[2]}: UNKNOWN WORD.")
    jd = _WAVE_SIDE[w[2]]
    if jd in d.waves:
        raise DeckError("MODEL OPEN-BC: DEFINED DOUBLY.")
    op = d.opens.setdefault(jd, dict(func="TYPE1", depth=-1.0, period=-1.0))
    if len(w) < 4:
        raise DeckError("MODEL OPEN-BC: SYNTAX ERROR.")
    key = w[3]
    if key == "FUNC":
        if len(w) < 5 or w[4] != "TYPE1":
            raise DeckError("MODEL OPEN-BC FUNC: UNKNOWN WORD.")
    elif key in ("DEPTH", "PERIOD"):
        if len(w) < 5:
            raise DeckError(f"MODEL OPEN-BC {key}: SYNTAX ERROR.")
        v = _num(w[4])
        if v < 1.0e-20:
            raise DeckError(f"MODEL OPEN-BC {key}: INVALID VALUE.")
        op[key.lower()] = v
    else:
        raise DeckError(f"MODEL OPEN-BC {key}: UNKNOWN WORD.")


def _parse_model(d, w):
    """MODEL WAVE-BC (Source code/vf_iimdl.f:33-115) and MODEL OPEN-BC
    (:174-210).  Only the matrix-table wave-maker (FUNC MATRIX with a
    <name>.mtb table) and the radiation boundary are produced here."""
    if len(w) >= 2 and w[1] == "OPEN-BC":
        _parse_open(d, w)
        return
    if len(w) >= 2 and w[1] == "DAMP":
        _parse_damp(d, w)
        return
    if len(w) < 2 or w[1] != "WAVE-BC":
        raise UnsupportedDeck(f"MODEL {w[1] if len(w) > 1 else ''} (wave sources, "
                              "turbulence) is not implemented")
    if len(w) < 4:
        raise DeckError("MODEL WAVE-BC: SYNTAX ERROR.")
    if w[2] not in _WAVE_SIDE:
        raise DeckError(f"MODEL WAVE-BC {w[2]}: UNKNOWN WORD.")
    if _WAVE_SIDE[w[2]] in d.opens:
        raise DeckError("MODEL WAVE-BC: DEFINED DOUBLY.")
    wv = d.waves.setdefault(_WAVE_SIDE[w[2]], dict(func=None, depth=0.0, height=0.0, period=0.0,
                                                   ampl=0.0, area=None, angle=0.0, ax=0.0, ay=0.0))
    key = w[3]
    if key == "FUNC":
        if len(w) < 5:
            raise DeckError("MODEL WAVE-BC FUNC: SYNTAX ERROR.")
        if w[4] not in ("MATRIX", "MATRIX2", "STOKES", "CNOIDAL", "STK-CND", "STREAM"):
            raise DeckError(f"MODEL WAVE-BC FUNC {w[4]}: UNKNOWN WORD.")
        if w[4] == "STREAM":
            # the order n of the stream function (vf_iimdl.f:52-57)
            if len(w) < 6:
                raise DeckError("MODEL WAVE-BC FUNC STREAM: SYNTAX ERROR.")
            n = _int(w[5])
            if n < 1 or n > 22:
                raise DeckError("MODEL WAVE-BC FUNC STREAM: INVALID VALUE.")
            if n == 22:
                raise UnsupportedDeck("MODEL WAVE-BC FUNC STREAM 22: at order 22 the reference's "
                                      "VF_DIFE writes past its SHCN(22) (vf_wsfmb2.f:519,612), so "
                                      "the order has no defined result")
            wv["order"] = n
        wv["func"] = w[4]
    elif key in ("DEPTH", "HEIGHT", "PERIOD", "AMPL"):
        if len(w) < 5:
            raise DeckError(f"MODEL WAVE-BC {key}: SYNTAX ERROR.")
        v = _num(w[4])
        if key != "AMPL" and v < 1.0e-20:
            raise DeckError(f"MODEL WAVE-BC {key}: INVALID VALUE.")
        wv[key.lower()] = v
    elif key == "AREA":
        if len(w) < 6:
            raise DeckError("MODEL WAVE-BC AREA: SYNTAX ERROR.")
        wv["area"] = (_int(w[4]) + 1, _int(w[5]) + 1)
    elif key == "ANGLE":
        if len(w) < 7:
            raise DeckError("MODEL WAVE-BC ANGLE: SYNTAX ERROR.")
        wv["angle"], wv["ax"], wv["ay"] = _num(w[4]), _num(w[5]), _num(w[6])
    else:
        raise DeckError(f"MODEL WAVE-BC {key}: UNKNOWN WORD.")


def _parse_damp(d, w):
    """MODEL DAMP side DEGREE n | PARAM-XY v | PARAM-Z v | WIDTH v | DEPTH v
    (Source code/vf_iimdl.f:216-258).  A side named here without DEGREE keeps
    degree -2, which the set-up check rejects."""
    if len(w) < 3:
        raise DeckError("MODEL DAMP: SYNTAX ERROR.")
    if w[2] not in _WAVE_SIDE:
        raise DeckError(f"MODEL DAMP {w[2]}: UNKNOWN WORD.")
    dm = d.damps.setdefault(_WAVE_SIDE[w[2]], dict(degree=-2, pxy=0.6, pz=0.6, width=0.0, depth=0.0))
    if len(w) < 4:
        raise DeckError("MODEL DAMP: SYNTAX ERROR.")
    key = w[3]
    if key not in ("DEGREE", "PARAM-XY", "PARAM-Z", "WIDTH", "DEPTH"):
        raise DeckError(f"MODEL DAMP {key}: UNKNOWN WORD.")
    if len(w) < 5:
        raise DeckError(f"MODEL DAMP {key}: SYNTAX ERROR.")
    if key == "DEGREE":
        n = _int(w[4])
        if n < 0:
            raise DeckError("MODEL DAMP DEGREE: INVALID VALUE.")
        dm["degree"] = n
    else:
        v = _num(w[4])
        if v < 1.0e-20:
            raise DeckError(f"MODEL DAMP {key}: INVALID VALUE.")
        dm[{"PARAM-XY": "pxy", "PARAM-Z": "pz", "WIDTH": "width", "DEPTH": "depth"}[key]] = v


def _parse_file(d, w):
    """FILE RSL (result file) and FILE RES (restart), Source code/vf_iifile.f:209-244."""
    if len(w) < 2:
        raise DeckError("FILE: SYNTAX ERROR.")
    if w[1] == "RSL":
        if len(w) < 3:
            raise DeckError("FILE RSL: SYNTAX ERROR.")
        if w[2] == "STEP":
            if len(w) < 6:
                raise DeckError("FILE RSL STEP: SYNTAX ERROR.")
            a, b, c = _int(w[3]), _int(w[4]), _int(w[5])
            if a > b or c <= 0:
                raise DeckError("FILE RSL STEP: INVALID VALUE.")
            d.rsl = ("STEP", a, b, c)
        elif w[2] == "TIME":
            if len(w) < 6:
                raise DeckError("FILE RSL TIME: SYNTAX ERROR.")
            a, b, c = _num(w[3]), _num(w[4]), _num(w[5])
            if a > b or c <= 0.0:
                raise DeckError("FILE RSL TIME: INVALID VALUE.")
            d.rsl = ("TIME", a, b, c)
        elif w[2] == "ELAPSE":
            pass    # wall-clock stop with a final result record: not produced
        else:
            raise DeckError(f"FILE RSL {w[2]}: UNKNOWN WORD.")
    elif w[1] == "RES":
        if len(w) < 3 or _int(w[2]) < 0:
            raise DeckError("FILE RES: INVALID VALUE.")
        d.res_step = _int(w[2])
    elif w[1] == "TRN":
        # vf_iifile.f:247-441: STEP / TIME, or one probe (parsed by tran.py
        # once the grid size is known)
        if len(w) < 3:
            raise DeckError("FILE TRN: SYNTAX ERROR.")
        if w[2] in ("STEP", "TIME"):
            if len(w) < 6:
                raise DeckError(f"FILE TRN {w[2]}: SYNTAX ERROR.")
            conv = _int if w[2] == "STEP" else _num
            a, b, c = conv(w[3]), conv(w[4]), conv(w[5])
            if a > b or c <= 0:
                raise DeckError(f"FILE TRN {w[2]}: INVALID VALUE.")
            d.trn = (w[2], a, b, c)
        else:
            d.trn_probes.append(tuple(w[2:]))
    elif w[1] == "GRP":
        # vf_iifile.f:148-207
        if len(w) < 3:
            raise DeckError("FILE GRP: SYNTAX ERROR.")
        if w[2] in ("STEP", "TIME"):
            if len(w) < 6:
                raise DeckError(f"FILE GRP {w[2]}: SYNTAX ERROR.")
            conv = _int if w[2] == "STEP" else _num
            a, b, c = conv(w[3]), conv(w[4]), conv(w[5])
            if a > b or c <= 0:
                raise DeckError(f"FILE GRP {w[2]}: INVALID VALUE.")
            d.grp = (w[2], a, b, c)
        elif w[2] == "AREA":
            if len(w) < 9:
                raise DeckError("FILE GRP AREA: SYNTAX ERROR.")
            d.grp_area = tuple(_int(x) + 1 for x in w[3:9])
        elif w[2] in ("ON", "OFF"):
            if len(w) >= 4 and w[3] == "VORT" and w[2] == "ON":
                raise UnsupportedDeck("FILE GRP ON VORT: vorticity output is not written")
        else:
            raise DeckError(f"FILE GRP {w[2]}: UNKNOWN WORD.")
    elif w[1] not in _IGNORED_FILE:
        raise UnsupportedDeck(f"FILE {w[1]} is not supported")


def _parse_time(d, w):
    """TIME CONST / AUTO / LIMIT / END, Source code/vf_iitime.f:28-60."""
    if len(w) < 2:
        raise DeckError("TIME: SYNTAX ERROR.")
    if w[1] == "CONST":
        if len(w) < 3:
            raise DeckError("TIME CONST: SYNTAX ERROR.")
        d.dt_const = _num(w[2])
        if d.dt_const < 1.0e-20:
            raise DeckError("TIME CONST: INVALID VALUE.")
        d.idttyp = 0
    elif w[1] == "AUTO":
        if len(w) < 4:
            raise DeckError("TIME AUTO: SYNTAX ERROR.")
        d.dtinit, d.dtsafe = _num(w[2]), _num(w[3])
        if d.dtinit < 1.0e-20 or d.dtsafe < 1.0e-20:
            raise DeckError("TIME AUTO: INVALID VALUE.")
        d.idttyp = 1
    elif w[1] == "LIMIT":
        if len(w) < 4:
            raise DeckError("TIME LIMIT: SYNTAX ERROR.")
        d.dtmin, d.dtmax = _num(w[2]), _num(w[3])
        if d.dtmin < 1.0e-20 or d.dtmin > d.dtmax:
            raise DeckError("TIME LIMIT: INVALID VALUE.")
    elif w[1] == "END":
        if len(w) < 4:
            raise DeckError("TIME END: SYNTAX.")
        d.nend = _int(w[2])
        d.tend = _num(w[3])
    else:
        raise DeckError(f"TIME {w[1]}: UNKNOWN WORD.")


def _parse_drag_df(d, w, lines, i):
    """OPTION DRAG-DF yu n, then n lines 'dr ap bt' (Source code/vf_iiopt.f:86-108);
    returns the index of the line after the table."""
    if len(w) < 4:
        raise DeckError("OPTION DRAG-DF: SYNTAX ERROR.")
    yu = _num(w[2])
    if yu <= 0.0:
        raise DeckError("OPTION DRAG-DF: INVALID VALUE.")
    n = _int(w[3])
    if n <= 0:
        raise DeckError("OPTION DRAG-DF: INVALID VALUE.")
    if n > 10:
        raise DeckError("OPTION DRAG-DF: AREA IS FULL.")
    rows = []
    for _ in range(n):
        if i >= len(lines):
            raise DeckError("OPTION DRAG-DF: SYNTAX ERROR.")
        r = lines[i]
        i += 1
        if len(r) != 3:
            raise DeckError("OPTION DRAG-DF: SYNTAX ERROR.")
        dr, ap, bt = (_num(x) for x in r)
        if dr <= 0.0 or ap < 0.0 or bt < 0.0:
            raise DeckError("OPTION DRAG-DF: INVALID VALUE.")
        rows.append((dr, ap, bt))
    d.drag_df = (yu, rows)
    return i


def _parse_option(d, w):
    key = w[1]
    if key == "T-DOOR":
        # Source code/vf_iiopt.f:57-77
        if len(w) < 4:
            raise DeckError("OPTION T-DOOR: SYNTAX ERROR.")
        if w[2] == "BUB":
            d.wbub = 0.0 if w[3] == "OFF" else _num(w[3])
            if d.wbub < 1.0e-20 and w[3] != "OFF":
                raise DeckError("OPTION T-DOOR BUB: INVALID VALUE.")
        elif w[2] == "DROP":
            if w[3] == "OFF":
                d.idrop = 0
            elif w[3] == "FREE-RUNDOWN":
                d.idrop = 1
            else:
                raise DeckError(f"OPTION T-DOOR DROP {w[3]}: UNKNOWN WORD.")
        else:
            raise DeckError(f"OPTION T-DOOR {w[2]}")
    elif key == "S-CELL-VEL":
        d.ibsuw0 = 1 if w[2] == "DU=0" else 0
    elif key == "SUB-LOOP":
        d.loops = _int(w[2])
    elif key == "MAX-VELOCITY":
        d.vvmax = _num(w[2])
    else:
        raise UnsupportedDeck(f"OPTION {key} is not implemented")


def _validate(d):
    if d.loops < 1:
        raise DeckError("OPTION SUB-LOOP: INVALID VALUE.")
    # vf_csetup.f:37-95: a table wave-maker needs DEPTH and PERIOD, no ANGLE
    # vf_csetup.f:60-100: a theory wave also needs HEIGHT; ANGLE is refused
    # for the table and not produced here for a theory wave
    for jd, wv in d.waves.items():
        if wv["func"] is None:
            raise DeckError("MODEL WAVE-BC: FUNC missing")
        if wv["depth"] < 1.0e-20:
            raise DeckError("MODEL WAVE-BC: (WAVE-BC DEPTH) NOT FOUND.")
        if wv["func"] not in ("MATRIX", "MATRIX2") and wv["height"] < 1.0e-20:
            raise DeckError("MODEL WAVE-BC: (WAVE-BC HEIGHT) NOT FOUND.")
        if wv["period"] < 1.0e-20:
            raise DeckError("MODEL WAVE-BC: (WAVE-BC PERIOD) NOT FOUND.")
        if abs(wv["angle"]) > 1.0e-20 and wv["func"] in ("MATRIX", "MATRIX2"):
            raise DeckError("MODEL WAVE-BC: (WAVE-BC ANGLE) CAN NOT USE.")
    # vf_csetup.f:221-234: a damping side needs DEGREE, WIDTH and DEPTH
    for jd, dm in d.damps.items():
        if dm["degree"] < 0:
            raise DeckError("VF_CSETUP: (DAMP DEGREE) NOT FOUND.")
        if dm["width"] < 1.0e-20:
            raise DeckError("VF_CSETUP: (DAMP WIDTH) NOT FOUND.")
        if dm["depth"] < 1.0e-20:
            raise DeckError("VF_CSETUP: (DAMP DEPTH) NOT FOUND.")
    # vf_csetup.f:128-137: a radiation boundary needs DEPTH and PERIOD
    for jd, op in d.opens.items():
        if op["depth"] < 1.0e-20:
            raise DeckError("VF_CSETUP: (OPEN-BC DEPTH) NOT FOUND.")
        if op["period"] < 1.0e-20:
            raise DeckError("VF_CSETUP: (OPEN-BC PERIOD) NOT FOUND.")

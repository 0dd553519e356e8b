"""The time-series file (.tran): probe values at selected steps.

FILE TRN lines of the deck (Source code/vf_iifile.f:247-441) declare the
probes; VF_OT1INI writes their list and the column line
(vf_ot1ini.f), VF_OT1TRN one line per selected step with TNOW and every
probe's value (vf_ot1trn.f, FORMAT 9510).  Probes:
  W-LEVEL i j            water level of column (i,j) above W-LEVEL (VF_CWLVL)
  W-LEVEL ANS side       the wave-maker's level on a side (BCTYP(7,side): the
                         table level times the ramp at the last boundary pass)
  POINT phy i j k        U, V, W (face values), P or F of a cell
  MIN|MAX|AV|INT phy box minimum, maximum, F-weighted mean and integral over
                         the fluid cells of a box (VF_CAREAP)
  FORCE side box         pressure force on the obstacle faces of a box
                         normal to x, y or z, on its - or + side (VF_CFORCE)
Indices in the deck are 1-based interior cell numbers.  Single rank; no
k-epsilon, temperature, scalar or vorticity probes, and no W-LEVEL ANS SRC
(wave sources are not modelled).
"""
from dataclasses import dataclass
from typing import List

import numpy as np

from .rsl import _e12_5  # noqa: F401  (shared Fortran E-format helper)

_KIND = {"POINT": 1, "MIN": 2, "MAX": 3, "AV": 4, "INT": 5, "FORCE": 11}
_KIND_TEXT = {0: "W-LEVEL", 1: "POINT", 2: "MIN", 3: "MAX", 4: "AV", 5: "INT", 11: "FORCE"}
_PHY = {"U": 1, "V": 2, "W": 3, "P": 4, "F": 5}
_SIDE = {"X-": 11, "X+": 12, "Y-": 13, "Y+": 14, "Z-": 15, "Z+": 16}
_ANS_SIDE = {"X-": 1, "X+": 2, "Y-": 3, "Y+": 4}
_PHY_TEXT = {0: "------", 1: "U", 2: "V", 3: "W", 4: "P", 5: "F", 11: "X-", 12: "X+", 13: "Y-",
             14: "Y+", 15: "Z-", 16: "Z+"}


class TrnError(ValueError):
    pass


@dataclass
class Probe:
    kind: int           # ITRPRM(1): 0 W-LEVEL, 1 POINT, 2..5 MIN/MAX/AV/INT, 11 FORCE
    phy: int            # ITRPRM(2)
    idx: tuple          # ITRPRM(3..8): cell indices with ghosts (+1), 0 where unused


def parse(words, shape) -> Probe:
    """One FILE TRN probe line (words after FILE TRN), vf_iifile.f:267-441."""
    numi, numj, numk = shape
    w = list(words)
    if not w:
        raise TrnError("FILE TRN: SYNTAX ERROR.")
    if w[0] == "W-LEVEL":
        if len(w) < 3:
            raise TrnError("FILE TRN W-LEVEL: SYNTAX ERROR.")
        if w[1] == "ANS":
            if w[2] not in _ANS_SIDE:
                raise TrnError(f"FILE TRN W-LEVEL ANS {w[2]}: not produced here (vf_iifile.f:275-288)")
            # ITRPRM(3) = -side, the rest 0 (vf_iifile.f:275-289)
            return Probe(0, 0, (-_ANS_SIDE[w[2]], 0, 0, 0, 0, 0))
        i1, j1 = int(w[1]), int(w[2])
        if not (1 <= i1 <= numi - 2 and 1 <= j1 <= numj - 2):
            raise TrnError("FILE TRN W-LEVEL: INVALID VALUE.")
        return Probe(0, 0, (i1 + 1, j1 + 1, 0, 0, 0, 0))
    if w[0] not in _KIND:
        raise TrnError(f"FILE TRN {w[0]}: UNKNOWN WORD.")
    kind = _KIND[w[0]]
    need = 6 if kind == 1 else 9
    if len(w) < need - 1:
        raise TrnError(f"FILE TRN {w[0]}: SYNTAX ERROR.")
    if kind == 11:
        if w[1] not in _SIDE:
            raise TrnError(f"FILE TRN FORCE {w[1]}: UNKNOWN WORD.")
        phy = _SIDE[w[1]]
    else:
        if w[1] not in _PHY:
            raise TrnError(f"FILE TRN {w[0]} {w[1]}: not produced here")
        phy = _PHY[w[1]]
    i1, j1, k1 = (int(x) for x in w[2:5])
    lim = [numi - 2, numj - 2, numk - 2]
    if kind == 1 and 1 <= phy <= 3:
        lim[phy - 1] += 1          # a face index may reach the last face
    if not (1 <= i1 <= lim[0] and 1 <= j1 <= lim[1] and 1 <= k1 <= lim[2]):
        raise TrnError("FILE TRN: INVALID VALUE.")
    i2 = j2 = k2 = -1
    if kind != 1:
        i2, j2, k2 = (int(x) for x in w[5:8])
        if not (i1 <= i2 <= numi - 2 and j1 <= j2 <= numj - 2 and k1 <= k2 <= numk - 2):
            raise TrnError("FILE TRN: INVALID VALUE.")
    return Probe(kind, phy, (i1 + 1, j1 + 1, k1 + 1, i2 + 1, j2 + 1, k2 + 1))


def _e15_8(x: float) -> str:
    s = f"{x:.8E}"
    mant, ex = s.split("E")
    e = int(ex)
    out = f"{mant}E{e:+03d}" if abs(e) <= 99 else f"{mant}{e:+04d}"
    return out.rjust(15)


def header(probes: List[Probe]) -> str:
    """The probe list and column line of VF_OT1INI (FORMAT 9510, 9530)."""
    lines = []
    sides = {v: k for k, v in _ANS_SIDE.items()}
    for n, p in enumerate(probes, 1):
        t1 = _KIND_TEXT[p.kind].ljust(12)[:10]
        if p.kind == 0 and p.idx[0] < 0:
            # FORMAT 9520: the side in A6 in place of the first index
            nums = "".join(f"{v - 1:6d}" for v in p.idx[1:])
            lines.append(f" {n:6d} {t1} {'ANS'.ljust(12)} {sides[-p.idx[0]].rjust(6)}{nums}")
            continue
        t2 = _PHY_TEXT[p.phy].ljust(12)
        nums = "".join(f"{v - 1:6d}" for v in p.idx)
        lines.append(f" {n:6d} {t1} {t2} {nums}")
    lines.append(" " + "TIME".rjust(15) + "".join(f" {n:15d}" for n in range(1, len(probes) + 1)))
    return "\n".join(lines) + "\n"


def _seqsum(vals):
    """Left-to-right sum from 0.0, as a Fortran accumulation loop."""
    acc = 0.0
    for v in vals:
        acc = acc + v
    return acc


def water_level(probe, st, ff, nf, wvlvl):
    """VF_CWLVL: sum over K of ZZ(2,K) times F (1 in obstacles), minus the
    still-water depth WVLVL - ZZ(1,2)."""
    i, j = probe.idx[0] - 1, probe.idx[1] - 1
    zz = st["ZZ"]
    numk = ff.shape[2]
    sf = 0.0
    for k in range(1, numk - 1):
        sf = sf + (zz[1, k] * 1.0 if nf[i, j, k] == -1 else zz[1, k] * ff[i, j, k])
    return sf - (wvlvl - zz[0, 1])


def _box(probe, shape):
    numi, numj, _ = shape
    i1, j1, k1, i2, j2, k2 = probe.idx
    return max(i1, 2), max(j1, 2), k1, min(i2, numi - 1), min(j2, numj - 1), k2


def area_stats(probe, st, fields):
    """VF_CAREAP: (VMIN, VMAX, VAV, VINT) over the box."""
    xx, yy, zz, ggv = st["XX"], st["YY"], st["ZZ"], st["GGV"]
    nf, ff = fields["NF"], fields["FF"]
    i1, j1, k1, i2, j2, k2 = _box(probe, nf.shape)
    isw = probe.phy
    val = {1: fields.get("UU"), 2: fields.get("VV"), 3: fields.get("WW"), 4: fields.get("PP"),
           5: ff}[isw]
    vals, ars, ints = [], [], []
    for k in range(k1, k2 + 1):
        for j in range(j1, j2 + 1):
            for i in range(i1, i2 + 1):
                n = nf[i - 1, j - 1, k - 1]
                if isw == 5:
                    if n == -1:
                        continue
                    f = ff[i - 1, j - 1, k - 1]
                    vals.append(f)
                    ar = xx[1, i - 1] * yy[1, j - 1] * zz[1, k - 1] * ggv[i - 1, j - 1, k - 1]
                    ars.append(ar)
                    ints.append(ar * f)
                    continue
                if n == -1 or n == 8:
                    continue
                ar = (xx[1, i - 1] * yy[1, j - 1] * zz[1, k - 1] * ggv[i - 1, j - 1, k - 1]
                      * ff[i - 1, j - 1, k - 1])
                if isw <= 3:
                    d = [0, 0, 0]
                    d[isw - 1] = 1
                    p1 = val[i - 1, j - 1, k - 1]
                    p2 = val[i - 1 + d[0], j - 1 + d[1], k - 1 + d[2]]
                    vals += [p1, p2]
                    ints.append(ar * (p1 + p2) * 0.5)
                else:
                    v = val[i - 1, j - 1, k - 1]
                    vals.append(v)
                    ints.append(ar * v)
                ars.append(ar)
    if not ars:
        return 1.0e30, -1.0e30, 0.0, 0.0
    vav, vint = _seqsum(ars), _seqsum(ints)
    if vav > 0.0:
        vav = vint / vav
    return min(vals), max(vals), vav, vint


def force(probe, st, fields):
    """VF_CFORCE: pressure force on the obstacle faces normal to one axis
    inside the box; sides 11/13/15 give the force on faces with the obstacle
    on the - side (FM), 12/14/16 on the + side (FP)."""
    h = (st["XX"][1], st["YY"][1], st["ZZ"][1])
    nf, ff, pp = fields["NF"], fields["FF"], fields["PP"]
    i1, j1, k1, i2, j2, k2 = _box(probe, nf.shape)
    a = (probe.phy - 11) // 2
    lo = [i1, j1, k1]
    lo[a] += 1
    others = [b for b in range(3) if b != a]
    fm, fp = 0.0, 0.0

    def contrib(c, ncode):
        """(DS, P) of a fluid (0) or surface cell c (0-based), or None."""
        if ncode == 0:
            return h[others[0]][c[others[0]]] * h[others[1]][c[others[1]]], pp[c]
        if not (1 <= ncode <= 6):
            return None
        b = (ncode - 1) // 2
        if b == a:
            return None
        rest = [x for x in others if x != b][0]
        dl = ff[c] * h[b][c[b]]
        ds = dl * h[rest][c[rest]]
        nb = list(c)
        nb[b] += -1 if ncode % 2 == 1 else 1
        nb = tuple(nb)
        p = dl / 2.0 / (dl + h[b][nb[b]] / 2.0) * pp[nb]
        return ds, p

    for k in range(lo[2], k2 + 1):
        for j in range(lo[1], j2 + 1):
            for i in range(lo[0], i2 + 1):
                c2 = (i - 1, j - 1, k - 1)
                c1 = list(c2)
                c1[a] -= 1
                c1 = tuple(c1)
                n1, n2 = nf[c1], nf[c2]
                if n1 == -1:
                    r = contrib(c2, n2)
                    if r is not None:
                        fm = fm - r[0] * r[1]
                elif n2 == -1:
                    r = contrib(c1, n1)
                    if r is not None:
                        fp = fp + r[0] * r[1]
    return fm if probe.phy % 2 == 1 else fp


def values(probes, st, fields, wvlvl, wave_levels=None):
    """The probe values of one step; wave_levels maps a wave-maker side to
    its level (BCTYP(7,side)) for W-LEVEL ANS."""
    out = []
    for p in probes:
        if p.kind == 0 and p.idx[0] < 0:
            side = -p.idx[0]
            if not wave_levels or side not in wave_levels:
                raise TrnError(f"W-LEVEL ANS: side {side} has no wave-maker")
            out.append(float(wave_levels[side]))
        elif p.kind == 0:
            out.append(water_level(p, st, fields["FF"], fields["NF"], wvlvl))
        elif p.kind == 1:
            arr = {1: "UU", 2: "VV", 3: "WW", 4: "PP", 5: "FF"}[p.phy]
            i, j, k = (x - 1 for x in p.idx[:3])
            out.append(float(fields[arr][i, j, k]))
        elif p.kind in (2, 3, 4, 5):
            out.append(area_stats(p, st, fields)[p.kind - 2])
        else:
            out.append(force(p, st, fields))
    return out


def line(tnow, vals):
    """One step of VF_OT1TRN (FORMAT 9510)."""
    return " " + _e15_8(tnow) + "".join(" " + _e15_8(v) for v in vals) + "\n"


class Writer:
    def __init__(self, path, probes, static, wvlvl):
        self.probes = list(probes)
        self.static = static
        self.wvlvl = wvlvl
        self.f = open(path, "w")
        self.f.write(header(self.probes))

    def write_step(self, tnow, fields, wave_levels=None):
        self.f.write(line(tnow, values(self.probes, self.static, fields, self.wvlvl, wave_levels)))

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

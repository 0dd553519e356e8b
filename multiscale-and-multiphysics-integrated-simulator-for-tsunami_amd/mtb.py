"""The wave-maker table (<name>.mtb) of MODEL WAVE-BC FUNC MATRIX.

Source code/vf_im1inp.f reads it as free-format words ('#' starts a comment,
vf_zgetim.f): the table kind (LEVEL-ON: water level and velocity profiles;
LEVEL-OFF: profiles only; LEVEL-ONLY: level only), the number of z levels
MTBZZ (not for LEVEL-ONLY), the number of phases MTBTT, the initial phase in
[0,1], the z levels (mean water level at 0, increasing), then per phase
'T', the phase (increasing in [0,1]), the level (not for LEVEL-OFF) and for
every z level the normal and tangential velocity.  A LEVEL-OFF table's level is
0 and its profiles are not scaled by the measured level; a LEVEL-ONLY table
holds the level alone and fixes the water level at its sides.
"""
from dataclasses import dataclass

import numpy as np

from .deck import zstoi, zstor


class MtbError(ValueError):
    pass


@dataclass
class Table:
    kind: str          # LEVEL-ON, LEVEL-OFF or LEVEL-ONLY
    t0: float          # DMTBT0
    zz: np.ndarray     # DMTBZZ (MTBZZ)
    tt: np.ndarray     # DMTBTT (MTBTT)
    hh: np.ndarray     # DMTBHH (MTBTT)
    un: np.ndarray     # DMTBUN (MTBZZ, MTBTT), Fortran order
    ut: np.ndarray     # DMTBUT (MTBZZ, MTBTT)


def _words(text):
    for line in text.splitlines():
        # the reader keeps the first 256 characters of a line (MAXCHR)
        line = line[:256].split("#", 1)[0]
        for w in line.replace(",", " ").split():
            yield w


def parse(text) -> Table:
    it = _words(text)

    def nxt():
        try:
            return next(it)
        except StopIteration:
            raise MtbError("VF_IM1INP: SYNTAX.") from None

    num = lambda: zstor(nxt())  # noqa: E731
    kind = nxt()
    if kind not in ("LEVEL-ON", "LEVEL-OFF", "LEVEL-ONLY"):
        raise MtbError(f"VF_IM1INP: UNKNOWN WORD {kind}.")
    only = kind == "LEVEL-ONLY"
    # a LEVEL-ONLY table has no z levels (vf_im1inp.f:71-80)
    nz = 0 if only else zstoi(nxt())
    nt = zstoi(nxt())
    if (nz <= 0 and not only) or nt <= 0:
        raise MtbError("VF_IM1INP: INVALID VALUE.")
    if (nz < 2 and not only) or nt < 2:
        raise MtbError("a table needs at least two z levels and two phases")
    t0 = num()
    if not 0.0 <= t0 <= 1.0:
        raise MtbError("VF_IM1INP: INVALID VALUE (initial phase).")
    zz = np.array([num() for _ in range(nz)])
    if np.any(np.diff(zz) <= 0.0):
        raise MtbError("VF_IM1INP: INVALID VALUE (z levels not increasing).")
    tt = np.zeros(nt)
    hh = np.zeros(nt)
    un = np.zeros((nz, nt), order="F")
    ut = np.zeros((nz, nt), order="F")
    for i in range(nt):
        if nxt() != "T":
            raise MtbError("VF_IM1INP: UNKNOWN WORD (T expected).")
        tt[i] = num()
        if not 0.0 <= tt[i] <= 1.0 or (i > 0 and tt[i] <= tt[i - 1]):
            raise MtbError("VF_IM1INP: INVALID VALUE (phase).")
        if kind != "LEVEL-OFF":
            hh[i] = num()
        for k in range(nz):
            un[k, i] = num()
            ut[k, i] = num()
    return Table(kind, t0, zz, tt, hh, un, ut)

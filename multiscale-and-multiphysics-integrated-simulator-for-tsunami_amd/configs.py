"""Synthetic workloads of the VOF-SMAC step, written as CADMAS-SURF/3D input decks.

Every case is expressed in the reference's own deck grammar (Source code/A11111.in):
GRID, OBST, POROUS, B.C., TIME, MATE, COMP, OPTION and DEBUG lines.  Decks are
produced as text so that the same case can be fed to the reference solver (to make
golden fixtures) and to this package's deck reader (deck.py).

Initial conditions that the grammar cannot express (a solitary wave) are returned
as a separate function applied to the fields after set-up, the way the reference
injects them through a restart file (Source code/vf_ir1inp.f:184-205).
"""
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np


def _coords(n, d, x0=0.0):
    return [x0 + i * d for i in range(n + 1)]


def _grid_lines(axis, coords, per_line=12):
    """GRID {X|Y|Z} c1 c2 ... END; continuation lines hold bare coordinates
    (Source code/vf_iigrid.f:70-111 reads words until END across lines)."""
    words = [f"{c:.6f}" for c in coords] + ["END"]
    lines = []
    for i in range(0, len(words), per_line):
        lines.append(" ".join(words[i:i + per_line]))
    lines[0] = f"GRID {axis} " + lines[0]
    return lines


@dataclass
class Case:
    """A named workload: deck text, sizes and an optional initial-condition hook."""
    name: str
    deck: str
    ni: int
    nj: int
    nk: int
    steps: int
    dt: float
    ic: Optional[Callable] = None
    meta: dict = field(default_factory=dict)

    @property
    def cells(self):
        return self.ni * self.nj * self.nk


def _common(lines, dt, steps, eps_r=1.0e-10, rho=1000.0, nu=1.0e-6, g=9.81, wlevel=None,
            tdoor=("OFF", "OFF"), time=None, loops=1):
    """Time, material and solver lines.  tdoor = (BUB word, DROP word) of
    OPTION T-DOOR (Source code/vf_iiopt.f:57-77); None leaves a model at the
    reference default (bubbles rise at 0.2 m/s, droplets FREE-RUNDOWN)."""
    # time=None: TIME CONST dt; otherwise the words after TIME of the
    # automatic step lines, e.g. ["AUTO 1e-4 0.2", "LIMIT 1e-6 2e-3"]
    tl = [f"TIME CONST {dt:.6e}"] if time is None else [f"TIME {t}" for t in time]
    lines += tl + [
        f"TIME END {steps} 1.0D9",
        f"MATE DENSITY {rho}",
        f"MATE K-VISC {nu:.6e}",
        f"MATE GRAVITY {g}",
        "COMP MTRX M-ILUBCGSTAB 0.95",
        f"COMP MTRX R-ERROR {eps_r:.6e}",
    ]
    for word, val in zip(("BUB", "DROP"), tdoor):
        if val is not None:
            lines.append(f"OPTION T-DOOR {word} {val}")
    if loops != 1:
        lines.append(f"OPTION SUB-LOOP {loops}")
    if wlevel is not None:
        lines.append(f"MATE W-LEVEL {wlevel:.6f}")


def dam_break(ni=64, nj=16, nk=48, d=0.01, col_x=0.20, col_z=0.30, steps=500, dt=1.0e-4,
              eps_r=1.0e-10, name=None):
    """Closed rectangular tank with a water column in the corner (dam break).

    Faces of the column align with grid lines so the initial F is exactly 0/1.
    """
    ic_i = int(round(col_x / d))
    ic_k = int(round(col_z / d))
    lines = [f"# dam break {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE"]
    _common(lines, dt, steps, eps_r)
    lines.append(f"DEBUG F-BOX 1 1 1 {ic_i} {nj} {ic_k} 1.0")
    return Case(name or f"dam_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def _beach_obst(lines, xs, dz, ni, nj, x0, slope, zcap=None):
    """Plane beach of obstacle cells: column i is solid up to the beach surface."""
    for i in range(ni):
        xc = 0.5 * (xs[i] + xs[i + 1])
        if xc <= x0:
            continue
        zb = (xc - x0) * slope
        if zcap is not None:
            zb = min(zb, zcap)
        kc = int(np.floor(zb / dz + 0.5))
        if kc >= 1:
            lines.append(f"OBST {i + 1} 1 1 {i + 1} {nj} {kc}")


def _porous_block(lines, i1, i2, nj, nk, por, cm, cd):
    lines.append(f"POROUS V {i1} 1 1 {i2} {nj} {nk} {por} 0.0")
    lines.append(f"POROUS X {i1} 1 1 {i2 + 1} {nj} {nk} {por} 0.0")
    lines.append(f"POROUS Y {i1} 1 1 {i2} {nj + 1} {nk} {por} 0.0")
    lines.append(f"POROUS Z {i1} 1 1 {i2} {nj} {nk + 1} {por} 0.0")
    lines.append(f"POROUS CM {i1} 1 1 {i2} {nj} {nk} {cm}")
    lines.append(f"POROUS CD {i1} 1 1 {i2} {nj} {nk} {cd}")


def beach_small(ni=40, nj=4, nk=20, dx=0.02, dz=0.01, steps=50, dt=2.0e-4, eps_r=1.0e-10,
                name=None):
    """Small deck-expressible run-up case exercising every static feature of the path.

    A dam-break column (DEBUG F-BOX) over still water (MATE W-LEVEL) runs through a
    porous rubble block (POROUS V/X/Y/Z/CM/CD) towards a plane beach of obstacle cells
    (OBST); the upstream x- face is an open free boundary (B.C. X ... VP FREE).
    """
    xs = _coords(ni, dx)
    lines = [f"# beach {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", xs)
    lines += _grid_lines("Y", _coords(nj, dx))
    lines += _grid_lines("Z", _coords(nk, dz))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE",
              f"B.C. X 1 1 1 1 {nj} {nk} VP FREE"]
    _beach_obst(lines, xs, dz, ni, nj, x0=0.55 * ni * dx, slope=0.25, zcap=0.6 * nk * dz)
    pi1 = int(0.35 * ni) + 1
    _porous_block(lines, pi1, pi1 + 3, nj, nk, 0.4, 1.2, 1.0)
    _common(lines, dt, steps, eps_r, wlevel=0.40 * nk * dz)
    lines.append(f"DEBUG F-BOX 1 1 1 {int(0.2 * ni)} {nj} {int(0.7 * nk)} 1.0")
    return Case(name or f"beach_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def porous_df(ni=40, nj=4, nk=20, dx=0.02, dz=0.01, steps=60, dt=2.0e-4, eps_r=1.0e-10,
              name=None):
    """The run-up case of beach_small with the Dupuit-Forchheimer drag of
    OPTION DRAG-DF (Source code/vf_iiopt.f:86-108, vf_vgene.f:111-146).

    The rubble block is two stone sizes, POROUS CD read as the diameter: 3 cm
    in its front half, 20 cm in its back half.  The three-row table then
    gives the first row on the block's faces to open water (diameter 1.5 cm),
    the second row inside the 3 cm half and the last row past the table's end
    in the 20 cm half; open water (CD = 0) lies below the first diameter and
    has no drag.
    """
    xs = _coords(ni, dx)
    lines = [f"# porous DF {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", xs)
    lines += _grid_lines("Y", _coords(nj, dx))
    lines += _grid_lines("Z", _coords(nk, dz))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE",
              f"B.C. X 1 1 1 1 {nj} {nk} VP FREE"]
    _beach_obst(lines, xs, dz, ni, nj, x0=0.55 * ni * dx, slope=0.25, zcap=0.6 * nk * dz)
    pi1 = int(0.35 * ni) + 1
    _porous_block(lines, pi1, pi1 + 3, nj, nk, 0.4, 1.2, 0.03)
    lines.append(f"POROUS CD {pi1 + 2} 1 1 {pi1 + 3} {nj} {nk} 0.2")
    _common(lines, dt, steps, eps_r, wlevel=0.40 * nk * dz)
    lines += ["OPTION DRAG-DF 1.0E-6 3",
              "  0.01  1000.0  1.1",
              "  0.02  1200.0  1.5",
              "  0.10  1500.0  2.5"]
    lines.append(f"DEBUG F-BOX 1 1 1 {int(0.2 * ni)} {nj} {int(0.7 * nk)} 1.0")
    return Case(name or f"porous_df_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def solitary_ic(h, H, x0, g=9.81, rho=1000.0):
    """Boussinesq solitary wave eta = H sech^2(sqrt(3H/4h^3)(x-x0)), u = sqrt(g/h) eta.

    Returns a function apply(xx, zz, nf, state) that overwrites state["FF"],
    state["UU"] and state["PP"] of a set-up field layout: F per column from the
    free-surface height, U on x-faces below the surface, P hydrostatic under the
    local surface.
    """
    kk = np.sqrt(3.0 * H / (4.0 * h ** 3))

    def apply(xx, zz, nf, state):
        numi, numj, numk = nf.shape
        xc = 0.5 * (xx[0, :-1] + xx[0, 1:])
        xc = np.append(xc, xc[-1] + xx[1, -1])
        eta_c = H / np.cosh(kk * (xc - x0)) ** 2
        eta_f = H / np.cosh(kk * (xx[0, :] - x0)) ** 2
        zb = zz[0, :]
        zt = np.append(zb[1:], zb[-1] + zz[1, -1])
        surf = h + eta_c
        frac = np.clip((surf[:, None] - zb[None, :]) / (zt - zb)[None, :], 0.0, 1.0)
        F = np.broadcast_to(frac[:, None, :], (numi, numj, numk)).copy()
        F[nf == -1] = 0.0
        F[0, :, :] = F[-1, :, :] = 0.0
        F[:, 0, :] = F[:, -1, :] = 0.0
        F[:, :, 0] = F[:, :, -1] = 0.0
        state["FF"][...] = F
        zc = 0.5 * (zb + zt)
        P = rho * g * (surf[:, None] - zc[None, :])
        state["PP"][...] = np.where(F > 0.0, np.broadcast_to(P[:, None, :], (numi, numj, numk)), 0.0)
        u = np.sqrt(g / h) * eta_f
        wet = zc[None, :] < (h + eta_f)[:, None]
        U = np.where(wet, u[:, None], 0.0)
        uu = state["UU"]
        uu[...] = np.broadcast_to(U[:, None, :], (numi, numj, numk))
        uu[nf == -1] = 0.0
        uu[0, :, :] = uu[1, :, :] = uu[-1, :, :] = 0.0
    return apply


def solitary_runup(ni=512, nj=32, nk=96, dx=0.02, dz=0.01, h=0.5, H_over_h=0.1, steps=2000,
                   dt=2.0e-4, eps_r=1.0e-10, name=None):
    """Solitary-wave run-up on a 1:20 beach from x=6 m behind a porous rubble block."""
    xs = _coords(ni, dx)
    lines = [f"# solitary run-up {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", xs)
    lines += _grid_lines("Y", _coords(nj, dx))
    lines += _grid_lines("Z", _coords(nk, dz))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE", f"B.C. X 1 1 1 1 {nj} {nk} VP FREE"]
    _beach_obst(lines, xs, dz, ni, nj, x0=6.0, slope=1.0 / 20.0, zcap=(nk - 4) * dz)
    i1 = int(round(4.0 / dx)) + 1
    i2 = int(round(4.4 / dx))
    if i2 <= ni:
        _porous_block(lines, i1, i2, nj, nk, 0.4, 1.2, 1.0)
    _common(lines, dt, steps, eps_r, wlevel=h)
    case = Case(name or f"runup_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt,
                ic=solitary_ic(h, H_over_h * h, 2.0))
    return case


def buildings_tank(ni=512, nj=256, nk=128, d=0.02, h=0.8, H_over_h=0.15, nbuild=40, seed=0,
                   steps=200, dt=2.0e-4, eps_r=1.0e-10, name=None):
    """Solitary wave over a 1:20 beach with seeded rectangular buildings on the beach half."""
    rng = np.random.default_rng(seed)
    xs = _coords(ni, d)
    lines = [f"# buildings {ni}x{nj}x{nk} seed {seed}"]
    lines += _grid_lines("X", xs)
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE", f"B.C. X 1 1 1 1 {nj} {nk} VP FREE"]
    # the run-up case's beach toe at x = 6 m on tanks long enough to hold it
    # (the 10.24 m tank of 512 cells); shorter test tanks start it half-way
    x0 = 6.0 if ni * d >= 9.0 else 0.5 * ni * d
    _beach_obst(lines, xs, d, ni, nj, x0=x0, slope=1.0 / 20.0, zcap=(nk - 4) * d)
    for _ in range(nbuild):
        fx, fy = rng.integers(4, 17, size=2)
        ht = int(rng.integers(8, 41))
        ia = int(rng.integers(ni // 2 + 1, ni - fx))
        ja = int(rng.integers(1, nj - fy))
        lines.append(f"OBST {ia} {ja} 1 {ia + fx - 1} {ja + fy - 1} {min(ht, nk - 2)}")
    _common(lines, dt, steps, eps_r, wlevel=h)
    return Case(name or f"buildings_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt,
                ic=solitary_ic(h, H_over_h * h, 2.0))


def corner3d(ni=24, nj=12, nk=16, d=0.01, steps=30, dt=1.0e-4, eps_r=1.0e-10, name=None):
    """Fully three-dimensional closed-tank case: a water column in one corner
    collapses around a solid block and through a porous block, so all three
    velocity components carry flow."""
    lines = [f"# corner column {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE"]
    lines.append(f"OBST {ni * 2 // 3} {nj // 3} 1 {ni * 2 // 3 + 2} {nj // 3 + 3} {nk // 4}")
    pi = ni // 2
    lines.append(f"POROUS V {pi} 1 1 {pi + 1} {nj} {nk // 2} 0.5 0.0")
    lines.append(f"POROUS X {pi} 1 1 {pi + 2} {nj} {nk // 2} 0.5 0.0")
    lines.append(f"POROUS Y {pi} 1 1 {pi + 1} {nj + 1} {nk // 2} 0.5 0.0")
    lines.append(f"POROUS Z {pi} 1 1 {pi + 1} {nj} {nk // 2 + 1} 0.5 0.0")
    lines.append(f"POROUS CM {pi} 1 1 {pi + 1} {nj} {nk // 2} 1.0")
    lines.append(f"POROUS CD {pi} 1 1 {pi + 1} {nj} {nk // 2} 0.8")
    _common(lines, dt, steps, eps_r, wlevel=2.5 * d)
    lines.append(f"DEBUG F-BOX 1 1 1 {ni // 3} {nj // 2} {nk * 5 // 8} 1.0")
    return Case(name or f"corner_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def tdoor_drops(ni=20, nj=5, nk=16, d=0.02, steps=300, dt=1.0e-3, eps_r=1.0e-10, wbub="2.0",
                name=None):
    """TimerDoor bubble and droplet models (Source code/vf_fbubup.f, vf_fdropf.f).

    A one-cell sheet of F = 0.5 hangs in the air (DEBUG F-BOX): every cell of
    it is a droplet (gas-flagged with water in it).  The sheet falls onto two
    blocks of different heights (droplets that land on the lower block next to
    the taller one slide sideways, and so do those next to the y walls) and
    into still water (MATE W-LEVEL), where merged cells feed the bubble model.
    Bubbles rise at WBUB m/s (2.0 fires every 10 steps here); the droplet model
    is left at the reference default.
    """
    lines = [f"# timer door {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE"]
    lines.append(f"OBST 12 1 1 14 {nj} 6")
    lines.append(f"OBST 15 1 1 17 {nj} 8")
    _common(lines, dt, steps, eps_r, wlevel=3 * d, tdoor=(wbub, None))
    lines.append(f"DEBUG F-BOX 10 1 12 17 {nj} 12 0.5")
    return Case(name or f"tdoor_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def tdoor_dam(ni=32, nj=4, nk=24, d=0.02, steps=900, dt=5.0e-4, eps_r=1.0e-10, name=None):
    """Dam break with both TimerDoor models at the reference defaults (no
    OPTION T-DOOR line): the bore runs up the far wall and throws droplets
    upward, and the bubble model fires every ZZ(2,K)/0.2 s."""
    lines = [f"# dam break, timer door defaults {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE"]
    _common(lines, dt, steps, eps_r, tdoor=(None, None))
    lines.append(f"DEBUG F-BOX 1 1 1 {ni * 3 // 8} {nj} {nk * 5 // 8} 1.0")
    return Case(name or f"tdoor_dam_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def auto_dam(ni=32, nj=4, nk=24, d=0.02, steps=400, eps_r=1.0e-10, name=None):
    """Dam break with the automatic time step (TIME AUTO DTINIT DTSAFE and
    TIME LIMIT DTMIN DTMAX, Source code/vf_iitime.f:37-52): every step size
    comes from VF_CDTCAL (vf_cdtcal.f), growing from 1e-4 s by at most 1.2x
    per step up to the flow's limit."""
    lines = [f"# dam break, automatic time step {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE"]
    _common(lines, None, steps, eps_r, time=["AUTO 1.0e-4 0.3", "LIMIT 1.0e-6 1.0e-2"])
    lines.append(f"DEBUG F-BOX 1 1 1 {ni * 3 // 8} {nj} {nk * 5 // 8} 1.0")
    return Case(name or f"auto_dam_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, None)


def subloop(loops=3, steps=40, name=None):
    """The three-dimensional corner case with OPTION SUB-LOOP: each step runs
    VF_V1CAL LOOPS times at DT/LOOPS (boundary values refreshed before the
    second and later passes, vf_v1cal.f:115-118), then VF_F1CAL at DT."""
    c = corner3d(steps=steps, dt=3.0e-4, name=name or f"subloop{loops}")
    deck = c.deck.replace("OPTION T-DOOR BUB OFF", f"OPTION SUB-LOOP {loops}\nOPTION T-DOOR BUB OFF")
    return Case(c.name, deck, c.ni, c.nj, c.nk, steps, c.dt)


def solitary_mtb(h, H, period, t_crest, nz=27, nt=201, ztop=None, g=9.81, kind="LEVEL-ON"):
    """A LEVEL-ON wave-maker table (LEVEL-OFF: without its level column;
    LEVEL-ONLY: the level alone)
    (.mtb, Source code/vf_im1inp.f) of one
    solitary wave.  The reference walks the table phase as frac(T0 - t/T),
    so with T0 = 0 the phase falls from 1 and time t reads phase 1 - t/T.
    Level eta = H sech^2(sqrt(3H/4h^3) c (t - t_crest)), c = sqrt(g(h+H));
    the velocity normal to the boundary is sqrt(g/h) eta below the surface
    and 0 above it, the tangential (vertical) velocity 0."""
    kk = np.sqrt(3.0 * H / (4.0 * h ** 3))
    c = np.sqrt(g * (h + H))
    ztop = H * 1.5 if ztop is None else ztop
    zs = np.linspace(-h, ztop, nz)
    # the reader takes at most 256 characters of a line (MAXCHR): one value
    # or one (normal, tangential) pair per line
    if kind == "LEVEL-ONLY":
        lines = [kind, str(nt), "0.0"]
    else:
        lines = [kind, str(nz), str(nt), "0.0"] + [f"{z:.8e}" for z in zs]
    for i in range(nt):
        ph = i / (nt - 1)
        t = (1.0 - ph) * period
        eta = H / np.cosh(kk * c * (t - t_crest)) ** 2
        u = np.sqrt(g / h) * eta
        lines.append(f"T {ph:.8e}" if kind == "LEVEL-OFF" else f"T {ph:.8e} {eta:.8e}")
        if kind != "LEVEL-ONLY":
            lines += [f"{(u if z <= eta else 0.0):.8e} 0.0" for z in zs]
    return "\n".join(lines) + "\n"


def wave_flume(ni=60, nj=4, nk=20, d=0.02, h=0.2, H=0.04, period=2.0, steps=400, dt=1.0e-3,
               eps_r=1.0e-10, open_end=False, table="LEVEL-ON", name=None):
    """A flume with a table wave-maker on its x- side (MODEL WAVE-BC X- FUNC
    MATRIX with <name>.mtb, vf_bwuwn.f:257-480): a solitary wave enters the
    still water (MATE W-LEVEL) and runs to the far wall, or with open_end
    leaves through a radiation boundary on the x+ side (MODEL OPEN-BC X+
    FUNC TYPE1, vf_bwuwn.f:483-620).  The table goes to
    meta["files"]["mtb"]."""
    lines = [f"# wave flume {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE",
              "MODEL WAVE-BC X- FUNC MATRIX", f"MODEL WAVE-BC X- DEPTH {h}",
              f"MODEL WAVE-BC X- PERIOD {period}"]
    if open_end:
        lines += ["MODEL OPEN-BC X+ FUNC TYPE1", f"MODEL OPEN-BC X+ DEPTH {h}",
                  f"MODEL OPEN-BC X+ PERIOD {period}"]
    _common(lines, dt, steps, eps_r, wlevel=h)
    case = Case(name or f"wave_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)
    case.meta["files"] = {"mtb": solitary_mtb(h, H, period, t_crest=0.2, kind=table)}
    return case


def wave_two(ni=40, nj=4, nk=20, d=0.02, h=0.2, H=0.04, period=2.0, steps=400, dt=1.0e-3,
             eps_r=1.0e-10, name=None):
    """The flume with a wave-maker at each end from two tables: X- FUNC MATRIX
    (<name>.mtb, the solitary wave's LEVEL-ON table) and X+ FUNC MATRIX2
    (<name>.mtb2, a LEVEL-ONLY table holding the far end at a smaller
    solitary rise later in the period; Source code/vf_im2inp.f)."""
    c = wave_flume(ni=ni, nj=nj, nk=nk, d=d, h=h, H=H, period=period, steps=steps, dt=dt,
                   eps_r=eps_r, name=name or f"wave_two_{ni}")
    c.deck += (f"MODEL WAVE-BC X+ FUNC MATRIX2\nMODEL WAVE-BC X+ DEPTH {h}\n"
               f"MODEL WAVE-BC X+ PERIOD {period}\n")
    c.meta["files"]["mtb2"] = solitary_mtb(h, H * 0.5, period, t_crest=0.6, kind="LEVEL-ONLY")
    return c


def theory_flume(func="STOKES", ni=40, nj=4, nk=20, d=0.02, h=0.2, H=0.02, period=1.0, ampl=1.0,
                 steps=600, dt=1.0e-3, eps_r=1.0e-10, angle=None, name=None):
    """A flume driven by a theory wave on its x- side (MODEL WAVE-BC X- FUNC
    STOKES, CNOIDAL or STK-CND; Source code/vf_cwmak0.f, vf_wstk0.f,
    vf_wcnd0.f), ramped in over AMPL periods, running into the far wall;
    angle = (degrees, SX, SY) turns it (MODEL WAVE-BC X- ANGLE)."""
    lines = [f"# {func.lower()} flume {ni}x{nj}x{nk}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE",
              f"MODEL WAVE-BC X- FUNC {func}", f"MODEL WAVE-BC X- DEPTH {h}",
              f"MODEL WAVE-BC X- HEIGHT {H}", f"MODEL WAVE-BC X- PERIOD {period}",
              f"MODEL WAVE-BC X- AMPL {ampl}"]
    if angle is not None:
        lines.append("MODEL WAVE-BC X- ANGLE {} {} {}".format(*angle))
    _common(lines, dt, steps, eps_r, wlevel=h)
    return Case(name or f"{func.lower()}_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def wave_damp(ni=40, steps=600, name=None):
    """The 40-cell wave flume with a damping zone (MODEL DAMP X+, the
    implicit sponge of VF_VEULER, Source code/vf_veuler.f:131-166) over its
    last 0.3 m in place of a wall reflection: DEGREE 4, so the profile
    (1 - d/W)**4 is a product whose order matters, at the default PARAM-XY
    and PARAM-Z 0.6."""
    c = wave_flume(ni=ni, steps=steps, name=name or f"wave_damp_{ni}")
    c.deck += ("MODEL DAMP X+ DEGREE 4\nMODEL DAMP X+ WIDTH 0.3\n"
               "MODEL DAMP X+ DEPTH 0.2\n")
    return c


def damp3d(steps=60, name=None):
    """The three-dimensional corner collapse with two overlapping damping
    zones on the column's sides (MODEL DAMP X- DEGREE 1 over 6 cm and Y-
    DEGREE 3 over 4 cm), PARAM-XY and PARAM-Z different, so the corner takes
    the larger of the two zones and w its own coefficient."""
    c = corner3d(steps=steps, name=name or "damp3d")
    c.deck += ("MODEL DAMP X- DEGREE 1\nMODEL DAMP X- WIDTH 0.06\nMODEL DAMP X- DEPTH 0.05\n"
               "MODEL DAMP X- PARAM-XY 1.5\nMODEL DAMP X- PARAM-Z 0.4\n"
               "MODEL DAMP Y- DEGREE 3\nMODEL DAMP Y- WIDTH 0.04\nMODEL DAMP Y- DEPTH 0.05\n"
               "MODEL DAMP Y- PARAM-XY 2.0\n")
    return c


def bore_tank(ni=1024, nj=512, nk=192, d=0.02, h_up=0.6, h_down=0.1, x_gate=4.0, nbuild=120,
              seed=0, bw=(8.0, 8.6), bw_top=0.8, steps=100, dt=2.0e-4, eps_r=1.0e-10, name=None):
    """A dam-break bore (water 0.6 m deep for x < 4 m over 0.1 m downstream)
    running into a full-width porous breakwater (x in [8, 8.6] m, up to
    0.8 m, porosity 0.45, CM 1.2, CD 1.0) and a seeded array of rectangular
    buildings behind it (footprints 6..24 cells, heights 10..60 cells, on
    the grid's last half), in a closed tank.

    The initial water is deck-expressible: MATE W-LEVEL for the downstream
    depth and DEBUG F-BOX for the upstream column, whose faces lie on grid
    lines (x_gate and h_up are whole cells), so F is exactly 0/1 and the
    reference can run the deck as it is.  The pressure is VF_CINIT's
    RHO0*GRZ0*(W-LEVEL - z) on every active cell (vf_cinit.f:283).
    """
    rng = np.random.default_rng(seed)
    lines = [f"# bore {ni}x{nj}x{nk} seed {seed}"]
    lines += _grid_lines("X", _coords(ni, d))
    lines += _grid_lines("Y", _coords(nj, d))
    lines += _grid_lines("Z", _coords(nk, d))
    lines += ["B.C. D VP SLIP", "B.C. D F FREE"]
    i1 = int(round(bw[0] / d)) + 1
    i2 = int(round(bw[1] / d))
    kt = min(int(round(bw_top / d)), nk)
    lines.append(f"POROUS V {i1} 1 1 {i2} {nj} {kt} 0.45 0.0")
    lines.append(f"POROUS X {i1} 1 1 {i2 + 1} {nj} {kt} 0.45 0.0")
    lines.append(f"POROUS Y {i1} 1 1 {i2} {nj + 1} {kt} 0.45 0.0")
    lines.append(f"POROUS Z {i1} 1 1 {i2} {nj} {kt + 1} 0.45 0.0")
    lines.append(f"POROUS CM {i1} 1 1 {i2} {nj} {kt} 1.2")
    lines.append(f"POROUS CD {i1} 1 1 {i2} {nj} {kt} 1.0")
    ilo = max(i2 + 4, ni // 2 + 1)
    for _ in range(nbuild):
        fx, fy = rng.integers(6, 25, size=2)
        ht = int(rng.integers(10, 61))
        ia = int(rng.integers(ilo, ni - fx))
        ja = int(rng.integers(1, nj - fy))
        lines.append(f"OBST {ia} {ja} 1 {ia + fx - 1} {ja + fy - 1} {min(ht, nk - 2)}")
    _common(lines, dt, steps, eps_r, wlevel=h_down)
    lines.append(f"DEBUG F-BOX 1 1 1 {int(round(x_gate / d))} {nj} {int(round(h_up / d))} 1.0")
    return Case(name or f"bore_{ni}x{nj}x{nk}", "\n".join(lines) + "\n", ni, nj, nk, steps, dt)


def named(name):
    """The workloads by name (sizes as in the benchmark table)."""
    table = {
        "dam_a": lambda: dam_break(),
        "runup_b": lambda: solitary_runup(),
        "buildings_c": lambda: buildings_tank(),
        "bore_e": lambda: bore_tank(),
        "tiny": lambda: dam_break(8, 3, 6, 0.01, 0.03, 0.04, steps=5),
        "beach_small": lambda: beach_small(),
        "corner3d": lambda: corner3d(),
        "tdoor": lambda: tdoor_drops(),
        "tdoor_dam": lambda: tdoor_dam(),
        "auto_dam": lambda: auto_dam(),
        "subloop": lambda: subloop(),
        "wave_flume": lambda: wave_flume(),
        "wave_open": lambda: wave_flume(ni=40, steps=1000, open_end=True),
        "porous_df": lambda: porous_df(),
        "wave_damp": lambda: wave_damp(),
        "wave_off": lambda: wave_flume(ni=40, steps=400, table="LEVEL-OFF"),
        "wave_only": lambda: wave_flume(ni=40, steps=400, table="LEVEL-ONLY"),
        "wave_two": lambda: wave_two(),
        "stokes_flume": lambda: theory_flume("STOKES"),
        "cnoidal_flume": lambda: theory_flume("STK-CND", H=0.03, period=2.0),
        "oblique_basin": lambda: theory_flume("STOKES", ni=30, nj=16, nk=16, h=0.16, period=0.8,
                                              steps=300, angle=(25.0, 0.0, 0.1)),
        "damp3d": lambda: damp3d(),
    }
    return table[name]()

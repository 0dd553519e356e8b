"""The reference's staggered porous field layout, built from a deck on the host.

Reproduces the set-up chain of Simulators/CADMAS-SURF-3D/Source code/:
  VF_CGRID  (metric tables XX/YY/ZZ(6,N), vf_cgrid.f:44-67)
  VF_CNFDFL (ghost layer = obstacle, vf_cnfdfl.f:28-52) and OBST boxes
  VF_CINDX  (face indices and the boundary-face numbering, vf_cindx.f:30-146)
  VF_CINDB  (INDB(1..2,*), vf_cindb.f:30-98) and the B.C. lines (vf_iiboun.f)
  POROUS    (GGV/GGX/GGY/GGZ, CM0, CD0; vf_iiporo.f:52-232)
  VF_CGLV / VF_CGLXYZ (inertia-augmented porosities, vf_cglv.f:34, vf_cglxyz.f:56-98)
  VF_CNU00  (viscosity, vf_cnu00.f:30-52)
  the F initial value of VF_CINIT from W-LEVEL and DEBUG F-BOX (vf_cinit.f:156-189).
The remainder of VF_CINIT (clip, flags, pressure, boundary values) runs on the
device (cad_init_state).  Every array is float64 / int32 in Fortran order,
(NUMI, NUMJ, NUMK) with one ghost layer on each side.
"""
import numpy as np

from . import deck as deckmod


def grid_metrics(coords):
    """XYZ(6,NUML) of vf_cgrid.f:44-67 for one axis (NUML = len(coords)+1)."""
    n = len(coords) + 1
    x = np.zeros((6, n))
    x[0, 1:] = coords
    dx = x[0, 2:n] - x[0, 1:n - 1]            # XYZ(2,I) for I=2..NUML-1
    x[1, 1:n - 1] = dx
    x[3, 1:n - 1] = 1.0 / dx
    x[1, 0] = x[1, 1]
    x[3, 0] = 1.0 / x[1, 0]
    x[0, 0] = x[0, 1] - x[1, 0]
    x[1, n - 1] = x[1, n - 2]
    x[3, n - 1] = 1.0 / x[1, n - 1]
    s = x[1, 1:n] + x[1, 0:n - 1]             # dx(I) + dx(I-1) for I=2..NUML
    x[2, 1:n] = (x[1, 1:n] + x[1, 0:n - 1]) / 2.0
    x[4, 1:n] = 1.0 / x[2, 1:n]
    x[5, 1:n] = 1.0 / s
    x[2, 0] = x[4, 0] = x[5, 0] = 0.0
    return np.asfortranarray(x)


def _f(shape, v=0.0, dtype=np.float64):
    return np.full(shape, v, dtype=dtype, order="F")


class Layout:
    """Set-up arrays and the initial state of one deck."""

    def __init__(self, d):
        self.deck = d
        self.XX = grid_metrics(d.grid["X"])
        self.YY = grid_metrics(d.grid["Y"])
        self.ZZ = grid_metrics(d.grid["Z"])
        ni, nj, nk = self.XX.shape[1], self.YY.shape[1], self.ZZ.shape[1]
        self.shape = (ni, nj, nk)
        self._flags()
        self._faces()
        self._boundaries()
        self._porosity()
        self._initial_f()

    # -------------------------------------------------------------- flags
    def _flags(self):
        ni, nj, nk = self.shape
        nf = _f(self.shape, 0, np.int32)
        nf[:, :, 0] = nf[:, :, -1] = -1
        nf[:, 0, :] = nf[:, -1, :] = -1
        nf[0, :, :] = nf[-1, :, :] = -1
        for (i1, j1, k1, i2, j2, k2) in self.deck.obst:
            if not (1 <= i1 <= i2 <= ni - 2 and 1 <= j1 <= j2 <= nj - 2 and 1 <= k1 <= k2 <= nk - 2):
                raise deckmod.DeckError(f"OBST box out of range: {(i1, j1, k1, i2, j2, k2)}")
            nf[i1:i2 + 1, j1:j2 + 1, k1:k2 + 1] = -1
        self.NF0 = nf

    # ---------------------------------------------------------- face index
    def _faces(self):
        ni, nj, nk = self.shape
        nf = self.NF0
        obs = nf == -1
        indx = _f(self.shape, 0, np.int32)
        indy = _f(self.shape, 0, np.int32)
        indz = _f(self.shape, 0, np.int32)
        indx[0, :, :] = -1
        indx[1:, :, :][obs[:-1, :, :] & obs[1:, :, :]] = -1
        indy[:, 0, :] = -1
        indy[:, 1:, :][obs[:, :-1, :] & obs[:, 1:, :]] = -1
        indz[:, :, 0] = -1
        indz[:, :, 1:][obs[:, :, :-1] & obs[:, :, 1:]] = -1
        numb = 0
        # x faces: K=2..NK-1, J=2..NJ-1, I=2..NI (vf_cindx.f:47-64)
        m = np.zeros(self.shape, bool, order="F")
        n1 = obs[:-1, 1:-1, 1:-1]
        n2 = obs[1:, 1:-1, 1:-1]
        cand = (indx[1:, 1:-1, 1:-1] != -1) & (n1 | n2)
        ii = np.arange(2, ni + 1)[:, None, None]
        skip = ((ii == 2) & n2) | ((ii == ni) & n1)
        m[1:, 1:-1, 1:-1] = cand & ~skip
        numb = self._number(indx, m, numb)
        m = np.zeros(self.shape, bool, order="F")
        n1 = obs[1:-1, :-1, 1:-1]
        n2 = obs[1:-1, 1:, 1:-1]
        cand = (indy[1:-1, 1:, 1:-1] != -1) & (n1 | n2)
        jj = np.arange(2, nj + 1)[None, :, None]
        skip = ((jj == 2) & n2) | ((jj == nj) & n1)
        m[1:-1, 1:, 1:-1] = cand & ~skip
        numb = self._number(indy, m, numb)
        m = np.zeros(self.shape, bool, order="F")
        n1 = obs[1:-1, 1:-1, :-1]
        n2 = obs[1:-1, 1:-1, 1:]
        m[1:-1, 1:-1, 1:] = (indz[1:-1, 1:-1, 1:] != -1) & (n1 | n2)
        numb = self._number(indz, m, numb)
        self.INDX, self.INDY, self.INDZ = indx, indy, indz
        self.numb = numb

    @staticmethod
    def _number(ind, mask, start):
        flat = mask.ravel(order="F")
        pos = np.flatnonzero(flat)
        out = ind.ravel(order="F")
        out[pos] = np.arange(start + 1, start + 1 + pos.size, dtype=np.int32)
        ind[...] = out.reshape(ind.shape, order="F")
        return start + pos.size

    # ------------------------------------------------------------ boundary
    def _boundaries(self):
        ni, nj, nk = self.shape
        nb = self.numb
        indb = np.zeros((4, max(nb, 1)), np.int32, order="F")
        obs = self.NF0 == -1
        for a, ind, (ns_fluid, ns_obst) in ((0, self.INDX, (1, 2)), (1, self.INDY, (3, 4)),
                                            (2, self.INDZ, (5, 6))):
            flat = ind.ravel(order="F")
            pos = np.flatnonzero(flat >= 1)
            ids = flat[pos] - 1
            indb[0, ids] = pos + 1
            n2_obst = obs.ravel(order="F")[pos]
            indb[1, ids] = np.where(n2_obst, ns_obst, ns_fluid)
        self.INDB = indb[:, :nb] if nb else np.zeros((4, 0), np.int32, order="F")
        self.BCU = np.zeros(nb)
        self.BCV = np.zeros(nb)
        self.BCW = np.zeros(nb)
        self.BCP = np.zeros(nb)
        self.BCF = np.zeros(nb)
        self.BCVI = np.zeros(nb)
        for bc in self.deck.bcs:
            if bc.phys not in ("VP", "F"):
                continue
            for a, ind in ((0, self.INDX), (1, self.INDY), (2, self.INDZ)):
                if bc.kind != "D" and "XYZ"[a] != bc.kind:
                    continue
                sl = self._bc_box(bc, a)
                ids = ind[sl].ravel(order="F")
                ids = ids[ids >= 1] - 1
                if bc.phys == "VP":
                    self.INDB[2, ids] = bc.code
                    self.BCU[ids], self.BCV[ids], self.BCW[ids], self.BCP[ids], self.BCVI[ids] = bc.values
                else:
                    self.INDB[3, ids] = bc.code
                    self.BCF[ids] = bc.values[0]
        # wave-maker sides (type 5, in AREA) and radiation sides (type 7, the
        # whole side): the type for velocity and F on every face of the side,
        # values cleared (vf_csetup.f:146-205)
        special = [(jd, 5, wv["area"]) for jd, wv in self.deck.waves.items()]
        special += [(jd, 7, None) for jd in self.deck.opens]
        for jd, code, area in sorted(special):
            a, ind = (0, self.INDX) if jd in (1, 2) else (1, self.INDY)
            pos = 2 if jd in (1, 3) else self.shape[a]
            other = 1 - a
            l1, l2 = area if area else (2, self.shape[other] - 1)
            sl = [slice(None)] * 3
            sl[a] = slice(pos - 1, pos)
            sl[other] = slice(l1 - 1, l2)
            sl[2] = slice(1, self.shape[2] - 1)
            ids = ind[tuple(sl)].ravel(order="F")
            ids = ids[ids >= 1] - 1
            self.INDB[2, ids] = code
            self.INDB[3, ids] = code
            for arr in (self.BCU, self.BCV, self.BCW, self.BCP, self.BCF, self.BCVI):
                arr[ids] = 0.0
        self.INDB = np.asfortranarray(self.INDB)
        if nb and (np.any(self.INDB[2] == 0) or np.any(self.INDB[3] == 0)):
            raise deckmod.DeckError("some boundary faces have no B.C. (add B.C. D lines)")

    def _bc_box(self, bc, a):
        ni, nj, nk = self.shape
        if bc.kind == "D":
            lo, hi = [2, 2, 2], [ni - 1, nj - 1, nk - 1]
            hi[a] = (ni, nj, nk)[a]
        else:
            i1, j1, k1, i2, j2, k2 = bc.box
            lo, hi = [i1 + 1, j1 + 1, k1 + 1], [i2 + 1, j2 + 1, k2 + 1]
        return tuple(slice(lo[q] - 1, hi[q]) for q in range(3))

    # ------------------------------------------------------------ porosity
    def _porosity(self):
        d = self.deck
        s = self.shape
        self.GGV, self.GGX, self.GGY, self.GGZ = _f(s, 1.0), _f(s, 1.0), _f(s, 1.0), _f(s, 1.0)
        self.CM0, self.CD0 = _f(s), _f(s)
        obs = self.NF0 == -1
        for kind, (i1, j1, k1, i2, j2, k2), vals in d.porous:
            sl = (slice(i1, i2 + 1), slice(j1, j2 + 1), slice(k1, k2 + 1))
            if kind in ("V", "X", "Y", "Z"):
                v1, v2 = vals
                if not 0.0 <= v1 <= 1.0:
                    raise deckmod.DeckError("POROUS value outside [0,1]")
                val = max(v2 + (1.0 - v2) * v1, d.plower)
                tgt, mask = {"V": (self.GGV, ~obs), "X": (self.GGX, self.INDX != -1),
                             "Y": (self.GGY, self.INDY != -1), "Z": (self.GGZ, self.INDZ != -1)}[kind]
            else:
                val = vals[0]
                tgt, mask = {"CM": (self.CM0, ~obs), "CD": (self.CD0, ~obs)}[kind]
            sub = tgt[sl]
            sub[mask[sl]] = val
        # VF_CGLV
        self.GLV = np.where(obs, 1.0, self.GGV + (1.0 - self.GGV) * self.CM0)
        self.GLV = np.asfortranarray(self.GLV)
        # VF_CGLXYZ (I, J, K = 1..NUM*; index-1 neighbours exist wherever used)
        self.GLX = self._glxyz(0, self.GGX, self.INDX, obs)
        self.GLY = self._glxyz(1, self.GGY, self.INDY, obs)
        self.GLZ = self._glxyz(2, self.GGZ, self.INDZ, obs)
        # VF_CNU00
        anu = _f(s)
        inner = (slice(1, -1),) * 3
        anu[inner] = np.where(obs[inner], 0.0, d.anu0)
        self.ANU = anu

    def _glxyz(self, a, gg, ind, obs):
        met = (self.XX, self.YY, self.ZZ)[a]
        cm = self.CM0
        cm_m = np.roll(cm, 1, axis=a)             # CM0 at index-1 (wraps only where unused)
        n = met.shape[1]
        shp = [1, 1, 1]
        shp[a] = n
        m6 = met[5].reshape(shp)
        m2 = met[1].reshape(shp)
        m2m = np.roll(met[1], 1).reshape(shp)
        cmi = m6 * (m2 * cm + m2m * cm_m)
        gl = np.where(ind == -1, 1.0,
                      np.where(ind == 0, gg + (1.0 - gg) * cmi,
                               np.where(~obs, gg + (1.0 - gg) * cm, gg + (1.0 - gg) * cm_m)))
        return np.asfortranarray(gl)

    # ---------------------------------------------------------- initial F
    def _initial_f(self):
        d = self.deck
        ni, nj, nk = self.shape
        zz = self.ZZ
        ff = _f(self.shape)
        w = d.wvlvl
        for k in range(2, nk):                    # K = 2..NUMK-1, 1-based
            zb, zt = zz[0, k - 1], zz[0, k]
            if zb > w:
                val = 0.0
            elif zt < w:
                val = 1.0
            else:
                val = (w - zb) / (zt - zb)
            plane = ff[1:-1, 1:-1, k - 1]
            plane[self.NF0[1:-1, 1:-1, k - 1] != -1] = val
        if d.fbox is not None:
            i1, j1, k1, i2, j2, k2, v = d.fbox
            sl = (slice(i1 - 1, i2), slice(j1 - 1, j2), slice(k1 - 1, k2))
            sub = ff[sl]
            sub[self.NF0[sl] != -1] = v
        self.FF0 = ff

    # ------------------------------------------------------------- exports
    def static_arrays(self):
        return dict(XX=self.XX, YY=self.YY, ZZ=self.ZZ, ANU=self.ANU, CD0=self.CD0, CM0=self.CM0,
                    GGV=self.GGV, GGX=self.GGX, GGY=self.GGY, GGZ=self.GGZ, GLV=self.GLV,
                    GLX=self.GLX, GLY=self.GLY, GLZ=self.GLZ, INDX=self.INDX, INDY=self.INDY,
                    INDZ=self.INDZ, INDB=self.INDB, BCVI=self.BCVI)

    def raw_state(self):
        """Fields before the device part of VF_CINIT: F from W-LEVEL/F-BOX,
        zero velocity and pressure, obstacle flags, INDC=-1."""
        s = self.shape
        return dict(UU=_f(s), VV=_f(s), WW=_f(s), PP=_f(s), FF=self.FF0.copy(order="F"),
                    BCU=self.BCU.copy(), BCV=self.BCV.copy(), BCW=self.BCW.copy(),
                    BCP=self.BCP.copy(), BCF=self.BCF.copy(), NF=self.NF0.copy(order="F"),
                    INDC=_f(s, -1, np.int32))

    def params(self, dt=None):
        d = self.deck
        if dt is None:
            dt = d.dt_const if d.idttyp == 0 else d.dtinit
        return dict(rho0=d.rho0, grz0=d.grz0, anu0=d.anu0, dtnow=dt,
                    scmvp=d.scmvp, vvmax=d.vvmax, flower=1.0e-6, fupper=1.0 - 1.0e-6,
                    plower=d.plower, cgpara=d.cgpara, cgepsa=d.cgepsa, cgepsr=d.cgepsr,
                    cgdiv=1.0e30, icgtyp=d.icgtyp, icgmax=d.icgmax, ibsuw0=d.ibsuw0,
                    loops=d.loops, wbub=d.wbub, idrop=d.idrop)


def from_text(text):
    return Layout(deckmod.parse(text))

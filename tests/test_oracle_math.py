"""Closed-form and invariant pins of the CPU restatement (oracle/).

These do not re-type the reference's formulas; they check properties the
mathematics fixes, so that a dropped term, a wrong sign or index, or a
transposed operand in the restatement fails:
  * VF_MZAX equals a dense matrix-vector product;
  * VF_MZDCMP with CGPARA=0 is ILU(0): M=(L D^-1 + I)(D + U) reproduces A on
    the diagonal; with CGPARA=1 (MILU) M preserves the row sums of A;
  * VF_MZMINV returns x with M x = y for the dense M;
  * VF_M1BCGS converges to scipy's direct solution;
  * a hydrostatic tank stays at rest; a closed tank conserves fluid volume.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import fixtures as fx
import pyoracle as po

KEYS = ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")


def _random_system(shape, seed, inactive=0.1, dominance=1.5):
    rng = np.random.default_rng(seed)
    ni, nj, nk = shape
    indc = -np.ones(shape, np.int32, order="F")
    inner = (slice(1, -1),) * 3
    indc[inner] = np.where(rng.random((ni - 2, nj - 2, nk - 2)) < inactive, -1, 0)
    m = {k: np.zeros(shape, order="F") for k in KEYS}
    act = indc == 0
    for k in ("ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK"):
        m[k][act] = -rng.uniform(0.2, 1.0, size=act.sum())
    # no coupling to inactive cells (vf_m1bcgs.f:12-14)
    off = {"ALI": (-1, 0), "AUI": (1, 0), "ALJ": (-1, 1), "AUJ": (1, 1), "ALK": (-1, 2), "AUK": (1, 2)}
    for k, (dlt, ax) in off.items():
        nb = np.roll(indc, -dlt, axis=ax)
        m[k][nb == -1] = 0.0
    s = sum(np.abs(m[k]) for k in off)
    m["AD"][act] = dominance * s[act] + 0.1
    m["AD"][~act] = 1.0
    m["BB"][act] = rng.normal(size=act.sum())
    return m, indc


def _dense(m, indc):
    """Dense A over active cells, in lexicographic order of the reference."""
    shape = indc.shape
    act = np.flatnonzero(indc.ravel(order="F") == 0)
    pos = -np.ones(indc.size, int)
    pos[act] = np.arange(act.size)
    ni, nj, _ = shape
    st = {"ALI": -1, "AUI": 1, "ALJ": -ni, "AUJ": ni, "ALK": -ni * nj, "AUK": ni * nj}
    A = np.zeros((act.size, act.size))
    flat = {k: m[k].ravel(order="F") for k in KEYS}
    for r, n in enumerate(act):
        A[r, r] = flat["AD"][n]
        for k, d in st.items():
            c = pos[n + d]
            if c >= 0:
                A[r, c] = flat[k][n]
    return A, act


def test_mzax_equals_dense_product():
    m, indc = _random_system((7, 6, 5), 1)
    A, act = _dense(m, indc)
    x = np.asfortranarray(np.random.default_rng(2).normal(size=indc.shape))
    x[indc == -1] = 0.0
    y = po.mzax(m, x, indc)
    assert np.allclose(y.ravel(order="F")[act], A @ x.ravel(order="F")[act], rtol=1e-13, atol=1e-13)


def _preconditioner(m, indc, dd):
    A, act = _dense(m, indc)
    D = np.diag(1.0 / dd.ravel(order="F")[act])
    Dinv = np.diag(dd.ravel(order="F")[act])
    L = np.tril(A, -1)
    U = np.triu(A, 1)
    return A, (L @ Dinv + np.eye(len(act))) @ (D + U), act


def test_ilu0_reproduces_diagonal():
    m, indc = _random_system((7, 6, 5), 3)
    dd, ok = po.mzdcmp(m, indc, cgpara=0.0, icgtyp=0)
    assert ok == 1
    A, M, _ = _preconditioner(m, indc, dd)
    assert np.allclose(np.diag(M), np.diag(A), rtol=1e-12)
    pattern = A != 0
    assert np.allclose(M[pattern], A[pattern], rtol=1e-12, atol=1e-12)


def test_milu_preserves_row_sums():
    m, indc = _random_system((7, 6, 5), 4)
    dd, ok = po.mzdcmp(m, indc, cgpara=1.0, icgtyp=1)
    assert ok == 1
    A, M, _ = _preconditioner(m, indc, dd)
    assert np.allclose(M.sum(axis=1), A.sum(axis=1), rtol=1e-11, atol=1e-11)


def test_mzminv_inverts_the_factors():
    m, indc = _random_system((8, 5, 6), 5)
    dd, ok = po.mzdcmp(m, indc, cgpara=0.95, icgtyp=1)
    assert ok == 1
    _, M, act = _preconditioner(m, indc, dd)
    y = np.asfortranarray(np.random.default_rng(6).normal(size=indc.shape))
    x = po.mzminv(m, dd, y, indc)
    xa = x.ravel(order="F")[act]
    assert np.allclose(M @ xa, y.ravel(order="F")[act], rtol=1e-11, atol=1e-11)


def test_milu_rejects_nonpositive_pivot():
    m, indc = _random_system((6, 5, 5), 7)
    m["AD"][indc == 0] = -1.0
    _, ok = po.mzdcmp(m, indc, cgpara=0.95, icgtyp=1)
    assert ok == 0
    x, bn, xn, it = po.m1bcgs(m, indc)
    assert it == -1


@pytest.mark.parametrize("seed", [8, 9])
def test_bicgstab_converges_to_direct_solution(seed):
    m, indc = _random_system((9, 7, 6), seed)
    A, act = _dense(m, indc)
    x, bn, xn, it = po.m1bcgs(m, indc, cgepsr=1e-14, cgepsa=0.0)
    assert 0 < it <= 500
    ref = spla.spsolve(sp.csr_matrix(A), m["BB"].ravel(order="F")[act])
    assert np.allclose(x.ravel(order="F")[act], ref, rtol=1e-10, atol=1e-12)


def test_bicgstab_iteration_cap_and_error_codes():
    m, indc = _random_system((9, 7, 6), 10)
    x, bn, xn, it = po.m1bcgs(m, indc, cgepsr=1e-30, cgepsa=0.0, icgmax=3)
    assert it == 4  # > ICGMAX: VF_VPSOL reports NOT CONVERGED


def _still_tank(pkg, level=0.12):
    text = "\n".join([
        "GRID X " + " ".join(f"{0.02 * i:.3f}" for i in range(13)) + " END",
        "GRID Y " + " ".join(f"{0.02 * i:.3f}" for i in range(6)) + " END",
        "GRID Z " + " ".join(f"{0.02 * i:.3f}" for i in range(11)) + " END",
        "B.C. D VP SLIP", "B.C. D F FREE", "OBST 5 2 1 7 3 3",
        "POROUS V 9 1 1 10 5 10 0.5 0.0", "POROUS CM 9 1 1 10 5 10 1.0",
        "TIME CONST 1.0D-3", "TIME END 10 1.0D9", "MATE DENSITY 1000.0", "MATE GRAVITY 9.81",
        f"MATE W-LEVEL {level}", "COMP MTRX R-ERROR 1.0D-12",
        "OPTION T-DOOR BUB OFF", "OPTION T-DOOR DROP OFF", ""])
    lay = pkg.layout.from_text(text)
    arr = dict(lay.static_arrays())
    arr.update(lay.raw_state())
    F = po.Fields(arr, lay.params())
    assert F.init_state(lay.deck.wvlvl, lay.deck.uini, 3) == 0
    return F


def test_hydrostatic_tank_stays_at_rest(pkg):
    """Flat surface on a cell face, p = rho g (h - z): gravity balances the
    pressure gradient (vf_vgene.f:95,301) and the fluid stays at rest."""
    F = _still_tank(pkg, level=0.12)
    f0 = F["FF"].copy()
    for _ in range(5):
        rc, _ = F.step()
        assert rc == 0
    scale = np.sqrt(9.81 * 0.12)
    for k in ("UU", "VV", "WW"):
        assert np.abs(F[k]).max() <= 1e-12 * scale, k
    assert np.array_equal(F["FF"], f0)


def test_closed_tank_conserves_volume():
    """FSUM changes only by the clipped amount FCUT (vf_fcut01.f:42-48)."""
    case = "corner3d"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    p["dtnow"] = float(r1["v1_in"]["DTNOW"][0])
    arr = fx.state_arrays(st, r1["v1_in"])
    arr["BCF"] = r1["f1_in"]["BCF"]
    F = po.Fields(arr, p, inds=r1["v1_in"]["INDS"][: int(r1["v1_in"]["NUMS"][0])])
    prev = float(r1["v1_in"]["FSUM"][0])
    for _ in range(40):
        rc, _ = F.step()
        assert rc == 0
        assert abs(F.st.fsum + F.st.fcut - prev) <= 1e-12 * prev
        prev = F.st.fsum

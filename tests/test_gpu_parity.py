"""Parity of the gfx950 HIP path with the reference, through the C ABI.

Golden fixtures were recorded from the reference solver itself (compiled from its
Fortran sources, tests/golden/make_golden.py); the CPU restatement in oracle/
is the checker where no recording exists.  Bars (from the plain FP64 form's
contract): stages whose arithmetic involves no reduction are bit-exact; the
Poisson solution, velocities, pressure and F agree to 1e-10 relative per step
(velocity components relative to the velocity scale max(|u|,|v|,|w|), so a
component that is pure round-off in a flow uniform along one axis is judged
against the flow); cell flags are bit-exact; the fluid volume is conserved to
1e-12.
"""
import numpy as np
import pytest

import fixtures as fx
import pyoracle as po
from helpers import assert_close, assert_same_int, assert_vel_close, rel_err

pytestmark = pytest.mark.gpu

TOL = 1.0e-10
CASES = ["tiny", "beach_small", "corner3d"]


def _params(st, rec):
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    p["dtnow"] = float(rec["DTNOW"][0])
    return p


def _capi_params(p):
    keys = ("rho0", "grz0", "anu0", "dtnow", "scmvp", "flower", "fupper", "plower", "cgpara",
            "cgepsa", "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")
    return {k: p[k] for k in keys if k in p}


def _ctx(pkg, case, st, rec):
    numi, numj, numk = st["GGV"].shape
    numb = st["INDB"].shape[1]
    c = pkg.capi.Context(numi, numj, numk, numb)
    c.set_params(**_capi_params(_params(st, rec)))
    c.upload_static(st)
    return c


def _state(rec, bcf):
    a = {k: rec[k] for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC")}
    a["BCF"] = bcf
    return a


@pytest.mark.parametrize("case", CASES)
def test_predictor_bitexact(pkg, case):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        with _ctx(pkg, case, st, r["v1_in"]) as c:
            c.upload_state(_state(r["v1_in"], r["f1_in"]["BCF"]))
            c.predictor()
            out = c.download_state()
        for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS")):
            assert np.array_equal(out[k], r["ps_in"][kk]), f"{case} step {s} {k}: {rel_err(out[k], r['ps_in'][kk]):.3e}"


@pytest.mark.parametrize("case", CASES)
def test_vpcoef_bitexact(pkg, case):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        a = _state(r["v1_in"], r["f1_in"]["BCF"])
        a["UU"], a["VV"], a["WW"] = r["ps_in"]["US"], r["ps_in"]["VS"], r["ps_in"]["WS"]
        with _ctx(pkg, case, st, r["v1_in"]) as c:
            c.upload_state(a)
            c.vpcoef()
            m = c.download_matrix()
        for k in pkg.capi.MATRIX:
            assert np.array_equal(m[k], r["ps_in"][k]), f"{case} step {s} {k}"


@pytest.mark.parametrize("case", CASES)
def test_vpsol_matches_reference(pkg, case):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        with _ctx(pkg, case, st, r["v1_in"]) as c:
            c.upload_state(_state(r["v1_in"], r["f1_in"]["BCF"]))
            c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
            info = c.vpsol()
            pt = c.download_pt()
        assert info["icgitr"] == int(r["ps_out"]["ICGITR"][0])
        assert_close(f"{case} step {s} PT", pt, r["ps_out"]["PT"], TOL)
        assert abs(info["cgbnrm"] - r["ps_out"]["CGBNRM"][0]) <= 1e-12 * r["ps_out"]["CGBNRM"][0]


@pytest.mark.parametrize("case", CASES)
def test_m1bcgs_dropin_entry(pkg, case):
    st = fx.static(case)
    s = fx.steps(case)[0]
    r = fx.step(case, s)
    p = _params(st, r["v1_in"])
    x, bn, xn, it = pkg.capi.m1bcgs(r["ps_in"], r["ps_in"]["INDC"], cgpara=p["cgpara"],
                                    cgepsa=p["cgepsa"], cgepsr=p["cgepsr"], cgdiv=p["cgdiv"],
                                    icgtyp=p["icgtyp"], icgmax=p["icgmax"])
    assert it == int(r["ps_out"]["ICGITR"][0])
    assert_close("PT", x, r["ps_out"]["PT"], TOL)


@pytest.mark.parametrize("case", CASES)
def test_v1cal_matches_reference(pkg, case):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        with _ctx(pkg, case, st, r["v1_in"]) as c:
            c.upload_state(_state(r["v1_in"], r["f1_in"]["BCF"]))
            info = c.v1cal()
            out = c.download_state()
            pt = c.download_pt()
        ref = r["v1_out"]
        assert info["icgitr"] == int(ref["ICGITR"][0])
        assert_vel_close(f"{case} step {s} velocity", out, ref, TOL)
        assert_vel_close(f"{case} step {s} boundary velocity", out, ref, TOL, ("BCU", "BCV", "BCW"))
        for k in ("PP", "BCP"):
            assert_close(f"{case} step {s} {k}", out[k], ref[k], TOL)
        assert_close(f"{case} step {s} PT", pt, ref["PT"], TOL)


@pytest.mark.parametrize("case", CASES)
def test_f1cal_matches_reference(pkg, case):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        fin, ref = r["f1_in"], r["f1_out"]
        with _ctx(pkg, case, st, fin) as c:
            a = {k: fin[k] for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF", "INDC")}
            c.upload_state(a)
            info = c.f1cal()
            out = c.download_state()
        # no reduction feeds F, the velocities or the flags in this phase
        for k in ("FF", "UU", "VV", "WW", "PP", "BCF", "BCU", "BCV", "BCW", "BCP"):
            assert np.array_equal(out[k], ref[k]), f"{case} step {s} {k}: {rel_err(out[k], ref[k]):.3e}"
        assert_same_int(f"{case} step {s} NF", out["NF"], ref["NF"])
        assert_same_int(f"{case} step {s} INDC", out["INDC"], ref["INDC"])
        assert info["nums"] == int(ref["NUMS"][0])
        fsum = ref["FSUM"][0]
        assert abs(info["fsum"] - fsum) <= 1e-12 * fsum
        assert abs(info["fcut"] - ref["FCUT"][0]) <= 1e-12 * fsum


def _oracle_fields(st, rec, bcf, params):
    arr = fx.state_arrays(st, rec)
    arr["BCF"] = bcf
    inds = rec["INDS"][: int(rec["NUMS"][0])]
    return po.Fields(arr, params, inds=inds)


def test_dam_break_per_step_parity_with_oracle(pkg):
    """Each step started from the oracle's state agrees with the oracle's step."""
    case = "dam_a"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    p = _params(st, r1["v1_in"])
    F = _oracle_fields(st, r1["v1_in"], r1["f1_in"]["BCF"], p)
    with _ctx(pkg, case, st, r1["v1_in"]) as c:
        for s in range(1, 31):
            c.upload_state({k: F[k] for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF", "INDC")})
            info = c.step()
            rc, _ = F.step()
            assert rc == 0
            out = c.download_state()
            assert info["icgitr"] == F.st.icgitr, f"step {s}"
            assert_vel_close(f"step {s} velocity", out, F, TOL)
            for k in ("PP", "FF"):
                assert_close(f"step {s} {k}", out[k], F[k], TOL)
            assert_same_int(f"step {s} NF", out["NF"], F["NF"])


@pytest.mark.parametrize("case,last", [("dam_a", 100), ("corner3d", 60)])
def test_free_run_against_reference(pkg, case, last):
    """Free-running GPU steps from the reference's initial state: ITR and the
    divergence norm every step, the recorded fields at the last step, and
    closed-tank volume conservation FSUM_n + FCUT_n = FSUM_(n-1) to 1e-12."""
    st = fx.static(case)
    r1 = fx.step(case, 1)
    traj = fx.trajectory(case)
    with _ctx(pkg, case, st, r1["v1_in"]) as c:
        c.upload_state(_state(r1["v1_in"], r1["f1_in"]["BCF"]))
        fsum_prev = float(r1["v1_in"]["FSUM"][0])
        for s in range(1, last + 1):
            info = c.step()
            assert info["icgitr"] == traj["ITR"][s], f"ITR at step {s}"
            assert abs(info["fsum"] + info["fcut"] - fsum_prev) <= 1e-12 * fsum_prev, f"volume at {s}"
            fsum_prev = info["fsum"]
            dv = c.cdiv00()
            assert abs(dv - traj["DV_FULL"][s]) <= 1e-8 * abs(traj["DV_FULL"][s]) + 1e-20, f"DV at {s}"
        out = c.download_state()
    ref = fx.step(case, last)["f1_out"]
    assert_vel_close(f"step {last} velocity", out, ref, 1e-8)
    for k in ("PP", "FF"):
        assert_close(f"step {last} {k}", out[k], ref[k], 1e-8)
    assert_same_int(f"step {last} NF", out["NF"], ref["NF"])


@pytest.mark.parametrize("case", CASES)
def test_true_residual_stop(pkg, case):
    """With cad_set_true_residual_stop(rel) the plain solve runs the
    reference's iteration until its preconditioned test passes and the FP64
    true residual is at most rel: at least the reference's ITR, the true
    residual within the bar, PT still within 1e-8 of the reference's."""
    st = fx.static(case)
    for s in fx.steps(case)[:2]:
        r = fx.step(case, s)
        with _ctx(pkg, case, st, r["v1_in"]) as c:
            c.upload_state(_state(r["v1_in"], r["f1_in"]["BCF"]))
            c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
            c.set_true_residual_stop(1e-11)
            info = c.vpsol()
            res = c.poisson_residual()
            pt = c.download_pt()
        assert info["icgitr"] >= int(r["ps_out"]["ICGITR"][0])
        assert res <= 1e-11, res
        assert_close(f"{case} step {s} PT", pt, r["ps_out"]["PT"], 1e-8)

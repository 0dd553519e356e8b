"""The Dupuit-Forchheimer porous drag (OPTION DRAG-DF, Source code/vf_iiopt.f:86-108,
applied in VF_VGENE, vf_vgene.f:111-146, 184-222, 260-298) on the GPU.

Fixture (tests/golden/make_golden.py): "porous_df", the 40x4x20 run-up case
with a rubble block of two stone sizes (POROUS CD 3 cm and 20 cm read as
diameters) and a three-row drag table, 60 steps recorded at 1, 2, 3, 30, 60.
The predictor has no reduction, so it is compared bit for bit; whole V1CAL
and free-run steps go through the Poisson solve and are held to the
plain-form bars.
"""
import numpy as np
import pytest

import fixtures as fx
from helpers import assert_close, assert_same_int, assert_vel_close, rel_err

pytestmark = pytest.mark.gpu

CASE = "porous_df"
STATE = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC")
KEYS = ("rho0", "grz0", "anu0", "scmvp", "flower", "fupper", "plower", "cgpara", "cgepsa",
        "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")


def _ctx(pkg, st, dtnow, drag=True):
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    c.set_params(**{k: p[k] for k in KEYS}, dtnow=dtnow)
    c.upload_static(st)
    if drag:
        c.set_drag_df(pkg.deck.parse(str(st["DECK"])).drag_df)
    return c


def _v1_state(r):
    a = {k: r["v1_in"][k] for k in STATE}
    a["BCF"] = r["f1_in"]["BCF"]
    return a


def test_predictor_with_drag_df_bitexact(pkg):
    """VF_VFLXDU..VF_VEULER with the Dupuit-Forchheimer drag: US, VS, WS bit
    for bit at every recorded step; without the table (the CD form) the same
    arrays give different velocities inside the block."""
    st = fx.static(CASE)
    for s in fx.steps(CASE):
        r = fx.step(CASE, s)
        print(f"{CASE} step {s} predictor", flush=True)
        outs = {}
        for drag in (True, False):
            with _ctx(pkg, st, float(r["v1_in"]["DTNOW"][0]), drag) as c:
                c.upload_state(_v1_state(r))
                c.predictor()
                outs[drag] = c.download_state()
        for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS")):
            out = outs[True][k]
            assert np.array_equal(out, r["ps_in"][kk]), f"step {s} {k}: {rel_err(out, r['ps_in'][kk]):.3e}"
        if s > 1:
            assert not np.array_equal(outs[False]["UU"], r["ps_in"]["US"]), f"step {s}"


def test_v1cal_with_drag_df(pkg):
    """Whole VF_V1CAL with the drag: ITR equal, fields to the plain-form bars."""
    st = fx.static(CASE)
    for s in fx.steps(CASE):
        r = fx.step(CASE, s)
        print(f"{CASE} step {s} v1cal", flush=True)
        with _ctx(pkg, st, float(r["v1_in"]["DTNOW"][0])) as c:
            c.upload_state(_v1_state(r))
            info = c.v1cal()
            out = c.download_state()
        ref = r["v1_out"]
        assert info["icgitr"] == int(ref["ICGITR"][0]), f"step {s}"
        assert_vel_close(f"step {s} velocity", out, ref, 1e-10)
        assert_close(f"step {s} PP", out["PP"], ref["PP"], 1e-10)


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_drag_df_free_run(pkg, tmp_path, precision):
    """The deck through driver.run to step 60: in the plain form ITR every
    step as the reference's list file has it and the fields at the recorded
    steps within 1e-8; in the mixed form flags equal and fields within 1e-6
    at the last step."""
    st = fx.static(CASE)
    tr = fx.trajectory(CASE)
    last = int(tr["STEP"][-1])
    deck = str(st["DECK"]) + "\nFILE RSL STEP 30 60 30\n"
    prec = pkg.capi.PRECISION_MIXED if precision == "mixed" else None
    res = pkg.driver.run(deck, "df", tmp_path, precision=prec)
    assert res["nnow"] == last
    recs = [r for r in pkg.rsl.read(tmp_path / "df.rsl")][1:]
    assert [r["NNOW"] for r in recs] == [30, 60]
    tol = 1e-8
    if precision == "plain":
        lines = (tmp_path / "df.list").read_text().splitlines()
        itr = [int(ln.split("ITR =")[1]) for ln in lines if "ITR =" in ln]
        assert itr == [int(x) for x in tr["ITR"][: len(itr)]] and len(itr) == last + 1
    else:
        tol = 1e-6
        recs = recs[-1:]
    for rec in recs:
        s = rec["NNOW"]
        print(f"{CASE} {precision} step {s}", flush=True)
        ref = fx.step(CASE, s)["f1_out"]
        assert_same_int(f"step {s} NF", rec["NF"], ref["NF"])
        assert_vel_close(f"step {s} velocity", rec, ref, tol)
        for k in ("PP", "FF"):
            assert_close(f"step {s} {k}", rec[k], ref[k], tol)


def test_drag_df_errors(pkg):
    """cad_set_drag_df refuses what VF_IIOPT stops on (vf_iiopt.f:89-106):
    YU <= 0, more than 10 rows, DR <= 0, AP or BT < 0; n = 0 restores the
    CD form."""
    st = fx.static(CASE)
    E = pkg.capi.CadError
    with _ctx(pkg, st, 1.0e-4, drag=False) as c:
        for drag in ((0.0, [(0.01, 1000.0, 1.1)]),
                     (1.0e-6, [(0.01, 1000.0, 1.1)] * 11),
                     (1.0e-6, [(0.0, 1000.0, 1.1)]),
                     (1.0e-6, [(0.01, -1.0, 1.1)]),
                     (1.0e-6, [(0.01, 1000.0, -0.5)])):
            with pytest.raises(E) as e:
                c.set_drag_df(drag)
            assert e.value.code == -1
        c.set_drag_df((1.0e-6, [(0.01, 1000.0, 1.1)]))
        c.set_drag_df(None)

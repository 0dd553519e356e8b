"""The implicit damping zones (MODEL DAMP, Source code/vf_iimdl.f:216-258,
applied in VF_VEULER, vf_veuler.f:75-166, 187-221, 243-277) on the GPU.

Fixtures (tests/golden/make_golden.py): "wave_damp", the 40x4x20 wave flume
whose solitary wave runs into a degree-4 zone over the last 0.3 m (600
steps, recorded at 1, 2, 300, 450, 600); "damp3d", the 3-D corner collapse
with two overlapping zones (X- degree 1, Y- degree 3, PARAM-Z different
from PARAM-XY; 60 steps, recorded at 1, 2, 10, 30, 60).  The predictor has
no reduction, so it is compared bit for bit; whole V1CAL and free-run steps
go through the Poisson solve and are held to the plain-form bars.
"""
import numpy as np
import pytest

import fixtures as fx
from helpers import assert_close, assert_same_int, assert_vel_close, rel_err

pytestmark = pytest.mark.gpu

CASES = ["wave_damp", "damp3d"]
RSL = {"wave_damp": (300, 600, 150), "damp3d": (30, 60, 30)}
STATE = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC")
KEYS = ("rho0", "grz0", "anu0", "scmvp", "flower", "fupper", "plower", "cgpara", "cgepsa",
        "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")


def _ctx(pkg, st, r, damps=True):
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    c.set_params(**{k: p[k] for k in KEYS}, dtnow=float(r["DTNOW"][0]), tnow=float(r["TNOW"][0]))
    c.upload_static(st)
    d = pkg.deck.parse(str(st["DECK"]))
    if d.waves:
        c.set_waves(d.waves, pkg.mtb.parse(str(st["FILE_MTB"])), d.wvlvl)
    if damps:
        c.set_damps(d.damps)
    return c


def _v1_state(r):
    a = {k: r["v1_in"][k] for k in STATE}
    a["BCF"] = r["f1_in"]["BCF"]
    return a


@pytest.mark.parametrize("case", CASES)
def test_predictor_with_damping_bitexact(pkg, case):
    """VF_VFLXDU..VF_VEULER with the damping zones: US, VS, WS bit for bit at
    every recorded step; without the zones the same step differs."""
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        print(f"{case} step {s} predictor", flush=True)
        outs = {}
        for damps in (True, False):
            with _ctx(pkg, st, r["v1_in"], damps) as c:
                c.upload_state(_v1_state(r))
                c.predictor()
                outs[damps] = c.download_state()
        for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS")):
            out = outs[True][k]
            assert np.array_equal(out, r["ps_in"][kk]), f"step {s} {k}: {rel_err(out, r['ps_in'][kk]):.3e}"
        if s > 1:
            assert not all(np.array_equal(outs[False][k], r["ps_in"][kk])
                           for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS"))), f"step {s}"


@pytest.mark.parametrize("case", CASES)
def test_v1cal_with_damping(pkg, case):
    """Whole VF_V1CAL with the zones: ITR equal, fields to the plain-form bars."""
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        if "v1_out" not in r:
            continue
        print(f"{case} step {s} v1cal", flush=True)
        with _ctx(pkg, st, r["v1_in"]) as c:
            c.upload_state(_v1_state(r))
            info = c.v1cal()
            out = c.download_state()
        ref = r["v1_out"]
        assert info["icgitr"] == int(ref["ICGITR"][0]), f"step {s}"
        assert_vel_close(f"step {s} velocity", out, ref, 1e-10)
        assert_close(f"step {s} PP", out["PP"], ref["PP"], 1e-10)


@pytest.mark.parametrize("case,precision", [("wave_damp", "plain"), ("damp3d", "plain"),
                                            ("wave_damp", "mixed")])
def test_damping_free_run(pkg, tmp_path, case, precision):
    """The deck through driver.run to its last step: in the plain form ITR
    every step as the reference's list file has it and the fields at the
    recorded steps within 1e-8; in the mixed form flags equal and fields
    within 1e-6 at the last step."""
    st = fx.static(case)
    tr = fx.trajectory(case)
    last = int(tr["STEP"][-1])
    a, b, n = RSL[case]
    deck = str(st["DECK"]) + f"\nFILE RSL STEP {a} {b} {n}\n"
    if "FILE_MTB" in st:
        (tmp_path / "dz.mtb").write_text(str(st["FILE_MTB"]))
    prec = pkg.capi.PRECISION_MIXED if precision == "mixed" else None
    res = pkg.driver.run(deck, "dz", tmp_path, precision=prec)
    assert res["nnow"] == last
    recs = [r for r in pkg.rsl.read(tmp_path / "dz.rsl")][1:]
    tol = 1e-8
    if precision == "plain":
        lines = (tmp_path / "dz.list").read_text().splitlines()
        itr = [int(ln.split("ITR =")[1]) for ln in lines if "ITR =" in ln]
        assert itr == [int(x) for x in tr["ITR"][: len(itr)]] and len(itr) == last + 1
    else:
        tol = 1e-6
        recs = recs[-1:]
    assert [r["NNOW"] for r in recs] == (list(range(a, b + 1, n)) if precision == "plain" else [b])
    for rec in recs:
        s = rec["NNOW"]
        print(f"{case} {precision} step {s}", flush=True)
        ref = fx.step(case, s)["f1_out"]
        assert_same_int(f"step {s} NF", rec["NF"], ref["NF"])
        assert_vel_close(f"step {s} velocity", rec, ref, tol)
        for k in ("PP", "FF"):
            assert_close(f"step {s} {k}", rec[k], ref[k], tol)


def test_damping_errors(pkg):
    """cad_set_damp refuses a side outside 1..4 and what VF_IIMDL and
    VF_CSETUP stop on (a coefficient, width or depth below 1e-20); a negative
    degree removes the zone."""
    case = "damp3d"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    E = pkg.capi.CadError
    lib = pkg.capi.lib()
    with _ctx(pkg, st, r1["v1_in"], damps=False) as c:
        assert lib.cad_set_damp(c._h, 0, 1, 0.6, 0.6, 0.1, 0.1) == -1
        assert lib.cad_set_damp(c._h, 5, 1, 0.6, 0.6, 0.1, 0.1) == -1
        for bad in ((0.0, 0.6, 0.1, 0.1), (0.6, 0.0, 0.1, 0.1), (0.6, 0.6, 0.0, 0.1),
                    (0.6, 0.6, 0.1, 0.0)):
            assert lib.cad_set_damp(c._h, 1, 2, *bad) == -1
        assert lib.cad_set_damp(c._h, 1, 2, 0.6, 0.6, 0.1, 0.1) == 0
        assert lib.cad_set_damp(c._h, 1, -1, 0.0, 0.0, 0.0, 0.0) == 0
        with pytest.raises(E):
            c.set_damps({2: dict(degree=1, pxy=0.6, pz=0.6, width=-1.0, depth=0.1)})

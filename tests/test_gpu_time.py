"""Automatic time step (VF_CDTCAL) and OPTION SUB-LOOP on the GPU.

Fixtures from the reference (tests/golden/make_golden.py): "auto_dam", a dam
break under TIME AUTO 1e-4 0.3 / TIME LIMIT 1e-6 1e-2, whose step size the flow
limits on most steps; "subloop", the corner case with OPTION SUB-LOOP 3.  The
trajectories hold every step's DT and TNOW to full precision.

The device combines the cell limits of VF_CDTCAL by an exact minimum, the
reference folds them in serially; the two differ by the rounding of V/(V/DT)
per cell, so step sizes are compared to 1e-12 relative (the CPU suite pins the
oracle's serial fold bit for bit and measures the drift at about 1e-15).
"""
import pytest

import fixtures as fx
from helpers import assert_close, assert_same_int, assert_vel_close

pytestmark = pytest.mark.gpu

STATE = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF", "INDC")
KEYS = ("rho0", "grz0", "anu0", "scmvp", "flower", "fupper", "plower", "cgpara", "cgepsa",
        "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")


def _ctx(pkg, st, loops=1):
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    c.set_params(**{k: p[k] for k in KEYS}, loops=loops)
    c.upload_static(st)
    return c


def _start_state(r1):
    a = {k: r1["v1_in"][k] for k in STATE if k != "BCF"}
    a["BCF"] = r1["f1_in"]["BCF"]
    return a


def test_cdtcal_on_recorded_states(pkg):
    """VF_CDTCAL from the state after recorded steps equals the reference's
    next step size to 1e-12."""
    case = "auto_dam"
    st = fx.static(case)
    tp = fx.time_params(st["DECK"])
    tr = fx.trajectory(case)
    last = int(tr["STEP"][-1])
    with _ctx(pkg, st) as c:
        for s in fx.steps(case):
            if s >= last:
                continue
            rec = fx.step(case, s)["f1_out"]
            c.upload_state({k: rec[k] for k in STATE})
            c.set_params(dtnow=float(tr["DT_FULL"][s]))
            dt = c.cdtcal(s, tp["dtinit"], tp["dtsafe"], tp["dtmin"], tp["dtmax"])
            ref = tr["DT_FULL"][s + 1]
            assert abs(dt - ref) <= 1e-12 * ref, (s, dt, ref)
        # nnow = 0 takes DTINIT as the previous step: 1.2 * DTINIT here
        r1 = fx.step(case, 1)
        c.upload_state(_start_state(r1))
        dt = c.cdtcal(0, tp["dtinit"], tp["dtsafe"], tp["dtmin"], tp["dtmax"])
        assert dt == tr["DT_FULL"][1]


def test_auto_time_step_free_run(pkg):
    """The deck's time loop (TIME AUTO, 400 steps) from the reference's
    initial state: DT and TNOW to 1e-12, ITR and ||div|| every step, and the
    fields at the recorded steps."""
    case = "auto_dam"
    st = fx.static(case)
    lay = pkg.layout.from_text(str(st["DECK"]))
    tr = fx.trajectory(case)
    kept = set(fx.steps(case))
    r1 = fx.step(case, 1)
    with _ctx(pkg, st) as c:
        c.upload_state(_start_state(r1))
        n = 0
        for rec in pkg.timeloop.run(c, lay.deck):
            s = rec.nnow
            n = s
            assert abs(rec.dtnow - tr["DT_FULL"][s]) <= 1e-12 * tr["DT_FULL"][s], f"DT at {s}"
            assert abs(rec.tnow - tr["TNOW_FULL"][s]) <= 1e-12 * tr["TNOW_FULL"][s], f"TNOW at {s}"
            assert rec.info["icgitr"] == tr["ITR"][s], f"ITR at {s}"
            dv = c.cdiv00()
            assert abs(dv - tr["DV_FULL"][s]) <= 1e-8 * abs(tr["DV_FULL"][s]) + 1e-20, f"DV at {s}"
            if s in kept:
                ref = fx.step(case, s)["f1_out"]
                out = c.download_state()
                assert_vel_close(f"step {s} velocity", out, ref, 1e-8)
                for k in ("PP", "FF"):
                    assert_close(f"step {s} {k}", out[k], ref[k], 1e-8)
                assert_same_int(f"step {s} NF", out["NF"], ref["NF"])
        assert n == int(tr["STEP"][-1])


def test_sub_loop_steps_match_reference(pkg):
    """OPTION SUB-LOOP 3: each recorded step from the reference's state at
    its start (the input of its first V1CAL pass) agrees with the
    reference's step to the plain-form bars."""
    case = "subloop"
    st = fx.static(case)
    tr = fx.trajectory(case)
    with _ctx(pkg, st, loops=3) as c:
        for s in fx.steps(case):
            r = fx.step(case, s)
            a = {k: r["v1_in"][k] for k in STATE if k != "BCF"}
            a["BCF"] = r["f1_in"]["BCF"]
            c.upload_state(a)
            c.set_params(dtnow=float(tr["DT_FULL"][s]), tnow=float(tr["TNOW_FULL"][s - 1]))
            info = c.step()
            p = c.get_params()
            assert p["tnow"] == tr["TNOW_FULL"][s] and p["dtnow"] == tr["DT_FULL"][s]
            ref = r["f1_out"]
            assert info["icgitr"] == int(r["v1_out_l3"]["ICGITR"][0]), f"step {s}"
            out = c.download_state()
            assert_vel_close(f"step {s} velocity", out, ref, 1e-10)
            for k in ("PP", "FF"):
                assert_close(f"step {s} {k}", out[k], ref[k], 1e-10)
            assert_same_int(f"step {s} NF", out["NF"], ref["NF"])


def test_sub_loop_free_run(pkg):
    case = "subloop"
    st = fx.static(case)
    lay = pkg.layout.from_text(str(st["DECK"]))
    tr = fx.trajectory(case)
    r1 = fx.step(case, 1)
    with _ctx(pkg, st, loops=3) as c:
        c.upload_state(_start_state(r1))
        n = 0
        for rec in pkg.timeloop.run(c, lay.deck):
            s = n = rec.nnow
            assert rec.info["icgitr"] == tr["ITR"][s], f"ITR at {s}"
            dv = c.cdiv00()
            assert abs(dv - tr["DV_FULL"][s]) <= 1e-8 * abs(tr["DV_FULL"][s]) + 1e-20, f"DV at {s}"
        assert n == int(tr["STEP"][-1])
        out = c.download_state()
    ref = fx.step(case, n)["f1_out"]
    assert_vel_close("last step velocity", out, ref, 1e-8)
    assert_close("last step FF", out["FF"], ref["FF"], 1e-8)

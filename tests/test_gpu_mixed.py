"""The mixed form (CAD_PRECISION_MIXED): FP64 BiCGSTAB right-preconditioned by
an FP32 aggregation multigrid V-cycle, stopped on the true relative residual.

It is not bit-compatible with the reference (a different preconditioner, so a
different iterate), so it is judged by what the solve must deliver: the true
relative residual ||BB - A PT|| / ||BB|| at CGEPSR, recomputed in FP64, and
fields within 1e-6 relative of the reference's, the bound the mixed form is
held to.
"""
import numpy as np
import pytest

import fixtures as fx
from helpers import assert_close, assert_same_int, assert_vel_close

pytestmark = pytest.mark.gpu

CASES = ["tiny", "beach_small", "corner3d"]
KEYS = ("rho0", "grz0", "anu0", "scmvp", "flower", "fupper", "plower", "cgpara", "cgepsa",
        "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")
STATE = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC")


def _ctx(pkg, st, dtnow, **extra):
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    c.set_params(**{k: p[k] for k in KEYS}, dtnow=dtnow, precision=pkg.capi.PRECISION_MIXED, **extra)
    c.upload_static(st)
    return c, p


@pytest.mark.parametrize("case", CASES)
def test_mixed_solve_on_reference_systems(pkg, case):
    """The reference's own Poisson systems: true residual at CGEPSR in far
    fewer iterations than MILU-BiCGSTAB, PT within 1e-6 of the reference's."""
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        if "ps_in" not in r:
            continue
        print(f"{case} step {s}", flush=True)
        c, p = _ctx(pkg, st, float(r["v1_in"]["DTNOW"][0]))
        with c:
            a = {k: r["v1_in"][k] for k in STATE}
            a["BCF"] = r["f1_in"]["BCF"]
            c.upload_state(a)
            c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
            info = c.vpsol()
            res = c.poisson_residual()
            pt = c.download_pt()
        ref_itr = int(r["ps_out"]["ICGITR"][0])
        print(f"  ITR {info['icgitr']} (reference {ref_itr}) true residual {res:.2e}", flush=True)
        # the stop test: ||r|| <= CGEPSR ||BB|| + CGEPSA (the reference's
        # structure, with the true residual in place of the preconditioned one)
        bn = info["cgbnrm"]
        assert res * bn <= (p["cgepsr"] * bn + p["cgepsa"]) * (1.0 + 1e-6), (s, res, bn)
        assert abs(info["cgxnrm"] - res) <= 1e-3 * res + 1e-30
        assert 0 < info["icgitr"] <= max(ref_itr, 20)
        assert_close(f"{case} step {s} PT", pt, r["ps_out"]["PT"], 1e-6)


def test_mixed_free_run_dam_break(pkg):
    """100 steps of the 64x16x48 dam break in the mixed form from the
    reference's initial state: flags equal and velocity, pressure and F
    within 1e-6 of the reference's step 100."""
    case = "dam_a"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    c, _ = _ctx(pkg, st, float(r1["v1_in"]["DTNOW"][0]))
    with c:
        a = {k: r1["v1_in"][k] for k in STATE}
        a["BCF"] = r1["f1_in"]["BCF"]
        c.upload_state(a)
        worst = 0.0
        for s in range(1, 101):
            info = c.step()
            bn = info["cgbnrm"]
            worst = max(worst, c.poisson_residual() * bn / (1.0e-10 * bn + 1.0e-15))
        out = c.download_state()
    ref = fx.step(case, 100)["f1_out"]
    print(f"mixed dam break: worst residual / tolerance {worst:.3f}, last ITR {info['icgitr']}",
          flush=True)
    assert worst <= 1.0 + 1e-6
    assert_same_int("step 100 NF", out["NF"], ref["NF"])
    assert_vel_close("step 100 velocity", out, ref, 1e-6)
    for k in ("PP", "FF"):
        assert_close(f"step 100 {k}", out[k], ref[k], 1e-6)


@pytest.mark.parametrize("case", ["wave_open", "tdoor"])
def test_mixed_form_with_boundary_and_timer_door_models(pkg, tmp_path, case):
    """The mixed form under the models around the step: the wave flume with
    a radiation end (1000 steps) and the droplet deck with both TimerDoor
    models (300 steps), run from the deck through driver.run: flags equal and
    fields within 1e-6 of the reference's last step."""
    st = fx.static(case)
    tr = fx.trajectory(case)
    last = int(tr["STEP"][-1])
    deck = str(st["DECK"]) + f"\nFILE RSL STEP {last} {last} 1\n"
    if "FILE_MTB" in st:
        (tmp_path / "run.mtb").write_text(str(st["FILE_MTB"]))
    res = pkg.driver.run(deck, "run", tmp_path, precision=pkg.capi.PRECISION_MIXED)
    assert res["nnow"] == last
    rec = [r for r in pkg.rsl.read(tmp_path / "run.rsl")][1:][-1]
    ref = fx.step(case, last)["f1_out"]
    print(f"{case}: mixed run to step {last}", flush=True)
    assert_same_int(f"step {last} NF", rec["NF"], ref["NF"])
    assert_vel_close(f"step {last} velocity", rec, ref, 1e-6)
    for k in ("PP", "FF"):
        assert_close(f"step {last} {k}", rec[k], ref[k], 1e-6)


def test_mixed_falls_back_to_milu_bicgstab(pkg):
    """With the mixed form's own iterations capped at one, its solve cannot
    converge by itself: the reference's MILU-BiCGSTAB finishes it from the
    current iterate, to the true residual at CGEPSR, PT within 1e-6 of the
    reference's."""
    case = "corner3d"
    st = fx.static(case)
    s = fx.steps(case)[-1]
    r = fx.step(case, s)
    c, p = _ctx(pkg, st, float(r["v1_in"]["DTNOW"][0]))
    with c:
        a = {k: r["v1_in"][k] for k in STATE}
        a["BCF"] = r["f1_in"]["BCF"]
        c.upload_state(a)
        c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
        c.set_krylov_cap(1)
        info = c.vpsol()
        res = c.poisson_residual()
        pt = c.download_pt()
        stats = c.mixed_stats()
    print(f"fallback: ITR {info['icgitr']} stats {stats} true residual {res:.2e}", flush=True)
    assert stats["fallbacks"] == 1 and stats["fallback_iters"] > 0
    assert info["icgitr"] == 1 + stats["fallback_iters"]
    bn = info["cgbnrm"]
    assert res * bn <= (p["cgepsr"] * bn + p["cgepsa"]) * (1.0 + 1e-6)
    assert_close(f"{case} step {s} PT", pt, r["ps_out"]["PT"], 1e-6)

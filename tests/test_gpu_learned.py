"""The learned form (CAD_PRECISION_LEARNED): flexible GMRES in FP64
preconditioned by one FP32 V-cycle followed by the BF16 U-Net, with the
reference's MILU-BiCGSTAB as the fallback, stopped on the true relative
residual.

Like the mixed form it is not a restatement of the reference, so it is judged
by what the solve must deliver on the reference's own Poisson systems: the
true relative residual ||BB - A PT|| / ||BB|| at CGEPSR (recomputed in FP64)
and PT within 1e-6 of the reference's.  This holds whatever the network's
weights are (the library's initial ones here; trained ones when the weights
file exists), and also when the preconditioner is useless: with a zero
network alone the first restart cycle gains nothing and the fallback solves.
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


def _solve(pkg, case, s, weights=None, mode=1):
    st = fx.static(case)
    r = fx.step(case, s)
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    with c:
        c.set_params(**{k: p[k] for k in KEYS}, dtnow=float(r["v1_in"]["DTNOW"][0]),
                     precision=pkg.capi.PRECISION_LEARNED)
        c.upload_static(st)
        a = {k: r["v1_in"][k] for k in STATE}
        a["BCF"] = r["f1_in"]["BCF"]
        c.upload_state(a)
        c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
        if weights is not None:
            pkg.learned.set_weights(c, weights)
        pkg.learned.set_mode(c, mode)
        info = c.vpsol()
        res = c.poisson_residual()
        pt = c.download_pt()
        stt = pkg.learned.stats(c)
    return info, res, pt, stt, p, r


def _check(case, s, info, res, pt, p, r):
    bn = info["cgbnrm"]
    assert res * bn <= (p["cgepsr"] * bn + p["cgepsa"]) * (1.0 + 1e-6), (case, s, res, bn)
    assert abs(info["cgxnrm"] - res) <= 1e-3 * res + 1e-30
    assert_close(f"{case} step {s} PT", pt, r["ps_out"]["PT"], 1e-6)
    assert np.all(pt[r["ps_in"]["INDC"] == -1] == 0.0)


@pytest.mark.parametrize("case", CASES)
def test_learned_solve_on_reference_systems(pkg, case):
    for s in fx.steps(case):
        if "ps_in" not in fx.step(case, s):
            continue
        print(f"learned {case} step {s}", flush=True)
        info, res, pt, stt, p, r = _solve(pkg, case, s)
        print(f"  ITR {info['icgitr']} (reference {int(r['ps_out']['ICGITR'][0])}) "
              f"true residual {res:.2e} stats {stt}", flush=True)
        _check(case, s, info, res, pt, p, r)


def test_learned_falls_back_to_milu_bicgstab(pkg):
    """A zero network alone is no preconditioner: the first restart cycle
    cannot halve the residual, so the reference's MILU-BiCGSTAB finishes the
    solve from FGMRES's x, to the same true residual."""
    case = "corner3d"
    s = fx.steps(case)[-1]
    zero = np.zeros(pkg.unet.param_count(), np.float32)
    info, res, pt, stt, p, r = _solve(pkg, case, s, weights=zero, mode=0)
    print(f"fallback: ITR {info['icgitr']} stats {stt}", flush=True)
    assert stt["fallbacks"] == 1 and stt["fallback_iters"] > 0
    _check(case, s, info, res, pt, p, r)


def test_learned_free_run_dam_break(pkg):
    """40 steps of the 64x16x48 dam break in the learned form from the
    reference's initial state: every step's true residual at CGEPSR, and
    flags equal, velocity, pressure and F within 1e-6 of the plain form's
    step 40 (the reference's iterate-for-iterate restatement)."""
    case = "dam_a"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    out = {}
    for prec in (pkg.capi.PRECISION_FP64, pkg.capi.PRECISION_LEARNED):
        c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
        with c:
            c.set_params(**{k: p[k] for k in KEYS}, dtnow=float(r1["v1_in"]["DTNOW"][0]),
                         precision=prec)
            c.upload_static(st)
            a = {k: r1["v1_in"][k] for k in STATE}
            a["BCF"] = r1["f1_in"]["BCF"]
            c.upload_state(a)
            for k in range(40):
                info = c.step()
                if prec == pkg.capi.PRECISION_LEARNED:
                    assert info["cgxnrm"] <= p["cgepsr"] * (1 + 1e-6) + p["cgepsa"] / info["cgbnrm"]
            out[prec] = c.download_state()
    a, b = out[pkg.capi.PRECISION_LEARNED], out[pkg.capi.PRECISION_FP64]
    assert_same_int("NF", a["NF"], b["NF"])
    assert_vel_close("velocity", a, b, 1e-6)
    for k in ("PP", "FF"):
        assert_close(k, a[k], b[k], 1e-6)

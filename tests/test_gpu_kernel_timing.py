"""In-solve timing of the forms' dominant kernels (cad_set_kernel_timing,
cad_kernel_timing), which the benchmark's roofline figures come from.

On a reference Poisson system (beach_small, last step) each form is solved
on fresh contexts with the timing off, on for every launch and on for
every third: the events must change no bit of PT and no iteration count,
every launch (or every third) must be counted (M^-1: two per BiCGSTAB iteration plus the start's two, at most), and
the work per launch must equal what the isolated timing helpers report for
the same solve (the V-cycle's bytes on the solve's K range, the network
pass's FLOP on the solve's box)."""
import numpy as np
import pytest

import fixtures as fx

pytestmark = pytest.mark.gpu

KEYS = ("rho0", "grz0", "anu0", "scmvp", "flower", "fupper", "plower", "cgpara", "cgepsa",
        "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")
STATE = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC")
CASE = "beach_small"


def _solve(pkg, precision, timing):
    s = [k for k in fx.steps(CASE) if "ps_in" in fx.step(CASE, k)][-1]
    st = fx.static(CASE)
    r = fx.step(CASE, s)
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    with c:
        c.set_params(**{k: p[k] for k in KEYS}, dtnow=float(r["v1_in"]["DTNOW"][0]), precision=precision)
        c.upload_static(st)
        a = {k: r["v1_in"][k] for k in STATE}
        a["BCF"] = r["f1_in"]["BCF"]
        c.upload_state(a)
        c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
        if timing:
            c.set_kernel_timing(timing)
        info = c.vpsol()
        pt = c.download_pt()
        out = {"itr": info["icgitr"], "pt": pt}
        if timing:
            c.set_kernel_timing(False)
            for name, which in (("minv", pkg.capi.KERNEL_MINV), ("vcycle", pkg.capi.KERNEL_VCYCLE),
                                ("unet", pkg.capi.KERNEL_UNET)):
                out[name] = c.kernel_timing(which)
            if precision == pkg.capi.PRECISION_MIXED:
                out["iso"] = c.time_precond(2)
            elif precision == pkg.capi.PRECISION_LEARNED:
                out["iso"] = c.time_learned(2)
            # a solve after the timing is turned off adds nothing
            c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
            c.vpsol()
            out["after"] = [c.kernel_timing(w)[2] for w in range(3)]
    return out


@pytest.mark.parametrize("form", ["plain", "mixed", "learned"])
def test_kernel_timing_counts_launches_and_changes_nothing(pkg, form):
    prec = {"plain": None, "mixed": pkg.capi.PRECISION_MIXED, "learned": pkg.capi.PRECISION_LEARNED}[form]
    if prec is None:
        prec = pkg.capi.PRECISION_FP64
    print(f"kernel timing {form} {CASE}", flush=True)
    off = _solve(pkg, prec, 0)
    on = _solve(pkg, prec, 1)
    every3 = _solve(pkg, prec, 3)
    assert on["itr"] == off["itr"] == every3["itr"]
    assert np.array_equal(on["pt"], off["pt"]) and np.array_equal(every3["pt"], off["pt"])
    # every third launch: launches 0, 3, 6, ... are timed
    for k in ("minv", "vcycle", "unet"):
        assert every3[k][2] == (on[k][2] + 2) // 3, (k, every3[k][2], on[k][2])
    counts = {k: on[k][2] for k in ("minv", "vcycle", "unet")}
    print(f"ITR {on['itr']} launches {counts} timings {[on[k][:2] for k in ('minv', 'vcycle', 'unet')]}",
          flush=True)
    assert on["after"] == [counts["minv"], counts["vcycle"], counts["unet"]]
    if form == "plain":
        ms, work, n = on["minv"]
        assert on["itr"] <= n <= 2 * on["itr"] + 2 and ms > 0.0 and work == 0.0
        assert counts["vcycle"] == 0 and counts["unet"] == 0
    elif form == "mixed":
        ms, work, n = on["vcycle"]
        assert n >= on["itr"] and ms > 0.0
        assert work == on["iso"][1]
        assert counts["unet"] == 0
    else:
        ms, work, n = on["unet"]
        assert n >= 1 and ms > 0.0
        assert work == on["iso"][1]
        assert counts["vcycle"] == 0

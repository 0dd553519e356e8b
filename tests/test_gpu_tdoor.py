"""TimerDoor bubble (VF_FBUBUP) and droplet (VF_FDROPF) models on the GPU.

The fixtures come from the reference run on two decks (tests/golden/
make_golden.py): a sheet of droplets falling onto two blocks of different
heights and into still water, with bubbles rising every 10 steps ("tdoor"),
and a dam break with both models at the reference defaults whose bore throws
droplets up the far wall ("tdoor_dam").  VF_F1CAL has no reduction feeding F,
the flags or the droplet state, so those are compared bit for bit; steps that
include the Poisson solve are compared to the plain-form bars (1e-10).
"""
import numpy as np
import pytest

import fixtures as fx
import pyoracle as po
from helpers import assert_close, assert_same_int, assert_vel_close, rel_err

pytestmark = pytest.mark.gpu

CASES = ["tdoor", "tdoor_dam"]
DROP = ("DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW")
STATE = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF", "INDC")
KEYS = ("rho0", "grz0", "anu0", "dtnow", "scmvp", "flower", "fupper", "plower", "cgpara", "cgepsa",
        "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")


def _drops_close(name, td, ref, tol):
    """Droplet state: cells without a droplet exactly clear (DROPTX = -1e10,
    the rest 0) in both; on droplet cells the times relative to the largest
    time and the velocities relative to the velocity scale (a component that
    is round-off in a flow uniform along one axis is judged against the flow)."""
    clear = np.asarray(ref["DROPTX"]) == -1.0e10
    assert np.array_equal(np.asarray(td["DROPTX"]) == -1.0e10, clear), f"{name}: droplet cells differ"
    for k in DROP[1:]:
        assert not np.any(np.asarray(td[k])[clear]), f"{name}: {k} not clear"
    live = ~clear
    if not live.any():
        return
    for k in ("DROPTX", "DROPTY", "DROPTZ"):
        assert_close(f"{name} {k}", np.asarray(td[k])[live], np.asarray(ref[k])[live], tol)
    assert_vel_close(f"{name} droplet velocity", {k: np.asarray(td[k])[live] for k in DROP[3:]},
                     {k: np.asarray(ref[k])[live] for k in DROP[3:]}, tol, DROP[3:])


def _params(st, rec):
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    p["dtnow"] = float(rec["DTNOW"][0])
    return p


def _ctx(pkg, st, p, tnow):
    numi, numj, numk = st["GGV"].shape
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    c.set_params(**{k: p[k] for k in KEYS}, tnow=tnow)
    c.upload_static(st)
    return c


@pytest.mark.parametrize("case", CASES)
def test_f1cal_timer_door_bitexact(pkg, case):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        fin, ref = r["f1_in"], r["f1_out"]
        p = _params(st, fin)
        print(f"{case} step {s}", flush=True)
        with _ctx(pkg, st, p, float(fin["TNOW"][0])) as c:
            c.upload_state({k: fin[k] for k in STATE})
            c.upload_tdoor({k: fin[k] for k in ("TBUB",) + DROP})
            info = c.f1cal()
            out = c.download_state()
            td = c.download_tdoor()
        for k in ("FF", "UU", "VV", "WW", "PP", "BCF", "BCU", "BCV", "BCW", "BCP"):
            assert np.array_equal(out[k], ref[k]), f"{case} step {s} {k}: {rel_err(out[k], ref[k]):.3e}"
        assert_same_int(f"{case} step {s} NF", out["NF"], ref["NF"])
        assert_same_int(f"{case} step {s} INDC", out["INDC"], ref["INDC"])
        for k in ("TBUB",) + DROP:
            assert np.array_equal(td[k], ref[k]), f"{case} step {s} {k}: {rel_err(td[k], ref[k]):.3e}"
        assert info["nums"] == int(ref["NUMS"][0])
        fsum = ref["FSUM"][0]
        assert abs(info["fsum"] - fsum) <= 1e-12 * fsum
        assert abs(info["fcut"] - ref["FCUT"][0]) <= 1e-12 * fsum


def test_timer_door_steps_match_oracle(pkg):
    """900 dam-break steps, each started on the GPU from the oracle's state:
    the GPU step agrees with the oracle's step (ITR exactly, fields 1e-10,
    flags and the bubble timers exactly, droplet state 1e-10)."""
    case = "tdoor_dam"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    p = _params(st, r1["v1_in"])
    arr = fx.state_arrays(st, r1["v1_in"])
    arr["BCF"] = r1["f1_in"]["BCF"]
    for k in ("TBUB",) + DROP:
        arr[k] = r1["f1_in"][k]
    F = po.Fields(arr, p, inds=r1["v1_in"]["INDS"][: int(r1["v1_in"]["NUMS"][0])])
    last = int(fx.trajectory(case)["STEP"][-1])
    with _ctx(pkg, st, p, 0.0) as c:
        for s in range(1, last + 1):
            c.set_params(tnow=F.st.tnow)
            c.upload_state({k: F[k] for k in STATE})
            c.upload_tdoor({k: F[k] for k in ("TBUB",) + DROP})
            info = c.step()
            rc, _ = F.step()
            assert rc == 0
            out = c.download_state()
            td = c.download_tdoor()
            assert info["icgitr"] == F.st.icgitr, f"step {s}"
            assert c.get_params()["tnow"] == F.st.tnow, f"step {s}"
            assert_vel_close(f"step {s} velocity", out, F, 1e-10)
            for k in ("PP", "FF"):
                assert_close(f"step {s} {k}", out[k], F[k], 1e-10)
            assert_same_int(f"step {s} NF", out["NF"], F["NF"])
            assert np.array_equal(td["TBUB"], F["TBUB"]), f"step {s} TBUB"
            _drops_close(f"step {s}", td, F, 1e-10)


@pytest.mark.parametrize("case", CASES)
def test_timer_door_free_run_against_reference(pkg, case):
    """Free-running GPU steps from the reference's initial state: ITR every
    step, ||div|| to 1e-8, F and flags at the recorded steps, and the total
    water volume FSUM + FCUT conserved step to step to 1e-12 (closed tank)."""
    st = fx.static(case)
    r1 = fx.step(case, 1)
    tr = fx.trajectory(case)
    p = _params(st, r1["v1_in"])
    kept = set(fx.steps(case))
    with _ctx(pkg, st, p, 0.0) as c:
        a = {k: r1["v1_in"][k] for k in STATE if k != "BCF"}
        a["BCF"] = r1["f1_in"]["BCF"]
        c.upload_state(a)
        c.upload_tdoor({k: r1["f1_in"][k] for k in ("TBUB",) + DROP})
        fsum_prev = float(r1["v1_in"]["FSUM"][0])
        for s in range(1, int(tr["STEP"][-1]) + 1):
            info = c.step()
            assert info["icgitr"] == tr["ITR"][s], f"ITR at step {s}"
            assert abs(info["fsum"] + info["fcut"] - fsum_prev) <= 1e-12 * fsum_prev, f"volume at {s}"
            fsum_prev = info["fsum"]
            dv = c.cdiv00()
            assert abs(dv - tr["DV_FULL"][s]) <= 1e-8 * abs(tr["DV_FULL"][s]) + 1e-20, f"DV at {s}"
            if s in kept:
                ref = fx.step(case, s)["f1_out"]
                out = c.download_state()
                td = c.download_tdoor()
                assert_close(f"step {s} FF", out["FF"], ref["FF"], 1e-8)
                assert_same_int(f"step {s} NF", out["NF"], ref["NF"])
                assert np.array_equal(td["TBUB"], ref["TBUB"]), f"step {s} TBUB"
                _drops_close(f"step {s}", td, ref, 1e-8)

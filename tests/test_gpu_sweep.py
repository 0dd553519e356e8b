"""The pipelined MILU sweeps (one launch per sweep, wavefront tiles of 8x8
lines) against the hyperplane schedule (one launch per plane I+J+K) and the
reference.

Both schedules evaluate every cell's recurrence on the same operands in the same
order, so the Poisson solution, the iteration count and every field after whole
steps must agree bit for bit between them; against the reference the solver's
usual bar applies (same ITR, PT to 1e-10 relative).  The grids cover one tile,
ragged tiles in J and K, and (buildings tank) many tiles with obstacles and a
free surface, so tiles with no active cell and partial I ranges occur.
"""
import numpy as np
import pytest

import fixtures as fx
import pyoracle as po
from helpers import assert_close

pytestmark = pytest.mark.gpu

TOL = 1.0e-10
CASES = ["tiny", "beach_small", "corner3d"]


def _params(st, rec):
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    p["dtnow"] = float(rec["DTNOW"][0])
    keys = ("rho0", "grz0", "anu0", "dtnow", "scmvp", "flower", "fupper", "plower", "cgpara",
            "cgepsa", "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")
    return {k: p[k] for k in keys if k in p}


def _state(rec, bcf):
    a = {k: rec[k] for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC")}
    a["BCF"] = bcf
    return a


def _solve(pkg, st, r, schedule):
    numi, numj, numk = st["GGV"].shape
    with pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1]) as c:
        c.set_params(**_params(st, r["v1_in"]))
        c.upload_static(st)
        c.upload_state(_state(r["v1_in"], r["f1_in"]["BCF"]))
        c.set_sweep_schedule(schedule)
        c.upload_matrix(r["ps_in"], r["ps_in"]["INDC"])
        info = c.vpsol()
        return info, c.download_pt()


PIPELINED = ["SWEEP_PIPELINED", "SWEEP_PIPELINED_8X8", "SWEEP_PIPELINED_16X4"]


@pytest.mark.parametrize("sched", PIPELINED)
@pytest.mark.parametrize("case", CASES)
def test_schedules_agree_on_reference_systems(pkg, case, sched):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        print(f"case {case} step {s} {sched}", flush=True)
        ip, pp = _solve(pkg, st, r, getattr(pkg.capi, sched))
        ih, ph = _solve(pkg, st, r, pkg.capi.SWEEP_HYPERPLANE)
        for k in ("cgbnrm", "cgxnrm", "icgitr"):
            assert ip[k] == ih[k], (k, ip[k], ih[k])
        assert np.array_equal(pp, ph), f"{case} step {s}: PT differs between schedules"
        assert ip["icgitr"] == int(r["ps_out"]["ICGITR"][0])
        assert_close(f"{case} step {s} PT", pp, r["ps_out"]["PT"], TOL)


def _run_steps(pkg, lay, state, flags, schedule, nsteps):
    ni, nj, nk = lay.shape
    with pkg.capi.Context(ni, nj, nk, lay.numb) as c:
        p = lay.params()
        c.set_params(**{k: v for k, v in p.items() if k in c.params})
        c.upload_static(lay.static_arrays())
        c.upload_state(state)
        c.init_state(lay.deck.wvlvl, lay.deck.uini, flags)
        c.set_sweep_schedule(schedule)
        infos = [c.step() for _ in range(nsteps)]
        return infos, c.download_state(), c.download_pt()


@pytest.mark.parametrize("sched", PIPELINED)
@pytest.mark.parametrize("shape", [(48, 40, 36), (70, 19, 45)])
def test_schedules_agree_over_steps(pkg, shape, sched):
    ni, nj, nk = shape
    case = pkg.configs.buildings_tank(ni, nj, nk, nbuild=6, seed=3, eps_r=1.0e-10,
                                      name=f"buildings_{ni}x{nj}x{nk}")
    lay = pkg.layout.from_text(case.deck)
    state = lay.raw_state()
    case.ic(lay.XX, lay.ZZ, lay.NF0, state)
    print(f"buildings {shape} {sched}", flush=True)
    ip, sp, pp = _run_steps(pkg, lay, dict(state), 0, getattr(pkg.capi, sched), 4)
    ih, sh, ph = _run_steps(pkg, lay, dict(state), 0, pkg.capi.SWEEP_HYPERPLANE, 4)
    for k in ("icgitr", "cgbnrm", "cgxnrm", "fsum", "fcut", "nums"):
        assert [i[k] for i in ip] == [i[k] for i in ih], k
    assert all(i["icgitr"] > 0 for i in ip)
    assert np.array_equal(pp, ph), "PT differs between schedules"
    for k in sp:
        assert np.array_equal(sp[k], sh[k]), f"{k} differs between schedules"


ALL_SCHEDULES = PIPELINED + ["SWEEP_HYPERPLANE"]


def _milu_check(pkg, mat, indc, prm, sched, ys):
    ni, nj, nk = mat["AD"].shape
    with pkg.capi.Context(ni, nj, nk, 0) as c:
        c.set_params(**prm)
        c.set_sweep_schedule(getattr(pkg.capi, sched))
        c.upload_matrix(mat, indc)
        assert c.milu_factor() == 1
        outs = [c.milu_apply(y) for y in ys]
    dd, ok = po.mzdcmp(mat, indc, prm["cgpara"], prm["icgtyp"])
    assert ok == 1
    for y, x in zip(ys, outs):
        ref = po.mzminv(mat, dd, y, indc)
        bad = np.flatnonzero((x != ref).ravel(order="F"))
        assert bad.size == 0, f"{sched}: {bad.size} cells differ, first at {bad[:5]}"


@pytest.mark.parametrize("sched", ALL_SCHEDULES)
@pytest.mark.parametrize("case", CASES)
def test_milu_apply_bitexact_with_oracle(pkg, case, sched):
    """x = M^-1 y on the reference's Poisson matrices equals the serial MILU
    sweeps of the oracle (vf_mzdcmp.f, vf_mzfrwd.f, vf_mzbkwd.f) bit for bit."""
    st = fx.static(case)
    rs = np.random.default_rng(7)
    for s in fx.steps(case)[:2]:
        r = fx.step(case, s)
        print(f"case {case} step {s} {sched}", flush=True)
        mat = r["ps_in"]
        prm = _params(st, r["v1_in"])
        ys = [mat["BB"], rs.standard_normal(mat["AD"].shape)]
        _milu_check(pkg, mat, mat["INDC"], prm, sched, ys)


@pytest.mark.parametrize("sched", ALL_SCHEDULES)
def test_milu_apply_bitexact_many_tiles(pkg, sched):
    """The same on a buildings tank with obstacles and a free surface: many
    tiles, ragged edges, empty tiles and partial I ranges."""
    case = pkg.configs.buildings_tank(70, 38, 45, nbuild=6, seed=3, eps_r=1.0e-10, name="b70")
    lay = pkg.layout.from_text(case.deck)
    state = lay.raw_state()
    case.ic(lay.XX, lay.ZZ, lay.NF0, state)
    ni, nj, nk = lay.shape
    with pkg.capi.Context(ni, nj, nk, lay.numb) as c:
        p = lay.params()
        c.set_params(**{k: v for k, v in p.items() if k in c.params})
        c.upload_static(lay.static_arrays())
        c.upload_state(state)
        c.init_state(lay.deck.wvlvl, lay.deck.uini, 0)
        c.predictor()
        c.vpcoef()
        mat = c.download_matrix()
        indc = c.download_state()["INDC"]
    prm = {k: v for k, v in p.items() if k in ("cgpara", "icgtyp", "cgepsa", "cgepsr", "cgdiv", "icgmax")}
    rs = np.random.default_rng(11)
    _milu_check(pkg, mat, indc, prm, sched, [mat["BB"], rs.standard_normal(mat["AD"].shape)])

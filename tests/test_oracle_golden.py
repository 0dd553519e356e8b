"""The CPU restatement (oracle/) against the reference's recorded phases.

Every function of oracle/cadmas_oracle.c is pinned here by the reference's own
outputs on the same inputs (recorded by tests/golden/make_golden.py from the
reference compiled from its Fortran sources): bit-exact equality is required.
"""
import numpy as np
import pytest

import fixtures as fx
import pyoracle as po

CASES = ["tiny", "beach_small", "corner3d"]


def _params(st, rec):
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    p["dtnow"] = float(rec["DTNOW"][0])
    return p


def _fields(st, rec, p, bcf=None):
    arr = fx.state_arrays(st, rec)
    if bcf is not None:
        arr["BCF"] = bcf
    return po.Fields(arr, p, inds=rec["INDS"][: int(rec["NUMS"][0])])


@pytest.mark.parametrize("case", CASES + ["dam_a"])
def test_init_state_matches_reference(pkg, case):
    """Deck -> host layout -> VF_CINIT restatement == the reference's initial state."""
    st = fx.static(case)
    lay = pkg.layout.from_text(str(st["DECK"]))
    arr = dict(lay.static_arrays())
    arr.update(lay.raw_state())
    F = po.Fields(arr, lay.params())
    assert F.init_state(lay.deck.wvlvl, lay.deck.uini, 3) == 0
    r1 = fx.step(case, 1)
    ref = r1["v1_in"]
    for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC"):
        assert np.array_equal(F[k], ref[k]), k
    assert np.array_equal(F["BCF"], r1["f1_in"]["BCF"])
    assert np.array_equal(F.inds, ref["INDS"][: int(ref["NUMS"][0])])


@pytest.mark.parametrize("case", ["wave_flume", "wave_open", "wave_off", "wave_only", "wave_two"])
def test_init_state_with_special_sides(pkg, case):
    """VF_CINIT's boundary pass with a table wave-maker (at TNOW = 0) and a
    radiation side (the initial pass, ALN = ALT = 1) == the reference's
    initial state."""
    st = fx.static(case)
    lay = pkg.layout.from_text(str(st["DECK"]))
    arr = dict(lay.static_arrays())
    arr.update(lay.raw_state())
    F = po.Fields(arr, lay.params())
    d = lay.deck
    F.set_waves(d.waves, pkg.mtb.parse(str(st["FILE_MTB"])), d.wvlvl, _table2(pkg, st))
    F.set_opens(d.opens)
    F.st.tnow = 0.0
    assert F.init_state(d.wvlvl, d.uini, 3) == 0
    r1 = fx.step(case, 1)
    ref = r1["v1_in"]
    for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC"):
        assert np.array_equal(F[k], ref[k]), k
    assert np.array_equal(F["BCF"], r1["f1_in"]["BCF"])


@pytest.mark.parametrize("case", CASES)
def test_phases_bitexact(case):
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        p = _params(st, r["v1_in"])
        F = _fields(st, r["v1_in"], p, r["f1_in"]["BCF"])
        assert F.predictor() == 0
        for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS")):
            assert np.array_equal(F[k], r["ps_in"][kk]), (s, k)
        rc, m = F.vpcoef()
        assert rc == 0
        for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB"):
            assert np.array_equal(m[k], r["ps_in"][k]), (s, k)
        x, bn, xn, it = po.m1bcgs(r["ps_in"], r["ps_in"]["INDC"], cgpara=p["cgpara"],
                                  cgepsa=p["cgepsa"], cgepsr=p["cgepsr"])
        assert it == int(r["ps_out"]["ICGITR"][0])
        assert np.array_equal(x, r["ps_out"]["PT"])
        assert bn == r["ps_out"]["CGBNRM"][0] and xn == r["ps_out"]["CGXNRM"][0]
        F = _fields(st, r["v1_in"], p, r["f1_in"]["BCF"])
        rc, pt = F.v1cal()
        assert rc == 0
        for k in ("UU", "VV", "WW", "PP", "BCU", "BCV", "BCW", "BCP"):
            assert np.array_equal(F[k], r["v1_out"][k]), (s, k)
        assert np.array_equal(pt, r["v1_out"]["PT"])
        fin, ref = r["f1_in"], r["f1_out"]
        F = _fields(st, fin, p)
        assert F.f1cal() == 0
        for k in ("FF", "UU", "VV", "WW", "PP", "BCF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC", "NLIM"):
            assert np.array_equal(F[k], ref[k]), (s, k)
        assert F.nums == int(ref["NUMS"][0])
        assert np.array_equal(F.inds, ref["INDS"])
        assert F.st.fsum == ref["FSUM"][0] and F.st.fcut == ref["FCUT"][0]


def test_dam_break_500_steps_bitexact():
    """Config A (64x16x48 dam break) for all 500 steps: ITR and ||div|| every step
    and every field at steps 100 and 500 equal the reference's."""
    case = "dam_a"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    p = _params(st, r1["v1_in"])
    F = _fields(st, r1["v1_in"], p, r1["f1_in"]["BCF"])
    tr = fx.trajectory(case)
    for s in range(1, 501):
        rc, _ = F.step()
        assert rc == 0
        assert F.st.icgitr == tr["ITR"][s], s
        assert F.cdiv00() == tr["DV_FULL"][s], s
        if s in (100, 500):
            ref = fx.step(case, s)["f1_out"]
            for k in ("UU", "VV", "WW", "PP", "FF", "NF"):
                assert np.array_equal(F[k], ref[k]), (s, k)


TDOOR = ["tdoor", "tdoor_dam"]
DROP = ("DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW")


@pytest.mark.parametrize("case", TDOOR)
def test_f1cal_timer_door_bitexact(case):
    """VF_F1CAL with the bubble (VF_FBUBUP) and droplet (VF_FDROPF) models on:
    F, flags, the droplet fields and the bubble timers equal the reference's
    at every recorded step (droplets born, falling, rising, sliding off
    obstacles, merging, landing; bubble levels firing)."""
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        fin, ref = r["f1_in"], r["f1_out"]
        p = _params(st, fin)
        p["tnow"] = float(fin["TNOW"][0])
        F = _fields(st, fin, p)
        assert F.f1cal() == 0
        for k in ("FF", "UU", "VV", "WW", "PP", "BCF", "NF", "INDC", "NLIM", "TBUB") + DROP:
            assert np.array_equal(F[k], ref[k]), (case, s, k)
        assert F.nums == int(ref["NUMS"][0])
        assert F.st.fsum == ref["FSUM"][0] and F.st.fcut == ref["FCUT"][0]


@pytest.mark.parametrize("case", TDOOR)
def test_timer_door_free_run_bitexact(case):
    """Every step of the run from the reference's initial state: ITR and
    ||div|| each step, and all fields at the recorded steps, equal the
    reference's (TNOW advanced by the step as the time loop does)."""
    st = fx.static(case)
    r1 = fx.step(case, 1)
    p = _params(st, r1["v1_in"])
    arr = fx.state_arrays(st, r1["v1_in"])
    arr["BCF"] = r1["f1_in"]["BCF"]
    for k in ("TBUB",) + DROP:
        arr[k] = r1["f1_in"][k]
    F = po.Fields(arr, p, inds=r1["v1_in"]["INDS"][: int(r1["v1_in"]["NUMS"][0])])
    tr = fx.trajectory(case)
    kept = set(fx.steps(case))
    for s in range(1, int(tr["STEP"][-1]) + 1):
        rc, _ = F.step()
        assert rc == 0, s
        assert F.st.icgitr == tr["ITR"][s], s
        assert F.cdiv00() == tr["DV_FULL"][s], s
        if s in kept:
            ref = fx.step(case, s)["f1_out"]
            assert F.st.tnow == ref["TNOW"][0], s
            for k in ("UU", "VV", "WW", "PP", "FF", "NF", "TBUB") + DROP:
                assert np.array_equal(F[k], ref[k]), (s, k)


def _free_run_fields(case):
    st = fx.static(case)
    r1 = fx.step(case, 1)
    p = _params(st, r1["v1_in"])
    arr = fx.state_arrays(st, r1["v1_in"])
    arr["BCF"] = r1["f1_in"]["BCF"]
    return st, p, po.Fields(arr, p, inds=r1["v1_in"]["INDS"][: int(r1["v1_in"]["NUMS"][0])])


def test_auto_time_step_free_run_bitexact():
    """TIME AUTO (VF_CDTCAL before every step): every step size equals the
    reference's to the last bit, and so do ITR, ||div|| and the fields."""
    case = "auto_dam"
    st, p, F = _free_run_fields(case)
    tp = fx.time_params(st["DECK"])
    tr = fx.trajectory(case)
    kept = set(fx.steps(case))
    for s in range(1, int(tr["STEP"][-1]) + 1):
        dt = F.cdtcal(s - 1, tp["dtinit"], tp["dtsafe"], tp["dtmin"], tp["dtmax"])
        assert dt == tr["DT_FULL"][s], (s, dt, tr["DT_FULL"][s])
        F.st.dtnow = dt
        rc, _ = F.step()
        assert rc == 0, s
        assert F.st.tnow == tr["TNOW_FULL"][s], s
        assert F.st.icgitr == tr["ITR"][s], s
        assert F.cdiv00() == tr["DV_FULL"][s], s
        if s in kept:
            ref = fx.step(case, s)["f1_out"]
            for k in ("UU", "VV", "WW", "PP", "FF", "NF"):
                assert np.array_equal(F[k], ref[k]), (s, k)


def test_sub_loop_free_run_bitexact():
    """OPTION SUB-LOOP 3: three VF_V1CAL passes at DT/3 per step (boundary
    values refreshed before passes 2 and 3), then VF_F1CAL at DT."""
    case = "subloop"
    st, p, F = _free_run_fields(case)
    assert p["loops"] == 3
    tr = fx.trajectory(case)
    kept = set(fx.steps(case))
    F.st.dtnow = tr["DT_FULL"][1]     # the step's DT (V1CAL's record holds DT/3)
    for s in range(1, int(tr["STEP"][-1]) + 1):
        rc, _ = F.step()
        assert rc == 0, s
        assert F.st.dtnow == tr["DT_FULL"][s] and F.st.tnow == tr["TNOW_FULL"][s], s
        assert F.st.icgitr == tr["ITR"][s], s
        assert F.cdiv00() == tr["DV_FULL"][s], s
        if s in kept:
            ref = fx.step(case, s)["f1_out"]
            for k in ("UU", "VV", "WW", "PP", "FF", "NF"):
                assert np.array_equal(F[k], ref[k]), (s, k)


def test_sub_loop_passes_bitexact():
    """Each recorded V1CAL pass of a sub-looped step from its own input."""
    case = "subloop"
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        for l, sfx in ((1, ""), (2, "_l2"), (3, "_l3")):
            rin, rout = r["v1_in" + sfx], r["v1_out" + sfx]
            p = _params(st, rin)
            F = _fields(st, rin, p, r["f1_in"]["BCF"])
            rc, pt = F.v1cal_loop(l)
            assert rc == 0
            for k in ("UU", "VV", "WW", "PP", "BCU", "BCV", "BCW", "BCP"):
                assert np.array_equal(F[k], rout[k]), (s, l, k)
            assert np.array_equal(pt, rout["PT"]), (s, l)


def _cell_limits(st, rec):
    """Per-cell step limits of VF_CDTCAL evaluated independently (numpy):
    V / max|face flux| and (0.5 / sum 1/dx^2) / nu over non-obstacle cells."""
    ff, nf = rec["FF"], rec["NF"]
    xx, yy, zz = st["XX"], st["YY"], st["ZZ"]
    dx = xx[1][1:-1, None, None]
    dy = yy[1][None, 1:-1, None]
    dz = zz[1][None, None, 1:-1]
    I = (slice(1, -1),) * 3
    def fl(area, g, q, ind, sh):
        lo = np.abs(area * g[I] * q[I])
        up = np.abs(area * np.roll(g, -1, sh)[I] * np.roll(q, -1, sh)[I])
        for a, f in ((lo, ind[I]), (up, np.roll(ind, -1, sh)[I])):
            b = f >= 1
            bcf = np.where(b, rec["BCF"][np.maximum(f, 1) - 1], 0.0)
            a[b & (np.maximum(ff[I], bcf) <= 0.0)] = 0.0
        return lo, up
    fls = []
    for area, g, q, ind, sh in ((dy * dz, st["GGX"], rec["UU"], st["INDX"], 0),
                                (dx * dz, st["GGY"], rec["VV"], st["INDY"], 1),
                                (dx * dy, st["GGZ"], rec["WW"], st["INDZ"], 2)):
        fls += list(fl(area, g, q, ind, sh))
    m = np.max(fls, axis=0)
    v = st["GGV"][I] * dx * dy * dz
    live = nf[I] != -1
    with np.errstate(divide="ignore"):
        c1 = np.where(live & (m > 0), v / np.where(m > 0, m, 1.0), np.inf)
        vd = 0.5 / (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz))
        anu = st["ANU"][I]
        c2 = np.where(live & (anu > 0), vd / np.where(anu > 0, anu, 1.0), np.inf)
    return min(c1.min(), c2.min())


def test_cdtcal_is_the_cell_minimum():
    """What the recurrence computes: the smallest per-cell limit (times
    DTSAFE, capped by 1.2x the last step and DTMIN/DTMAX), up to the
    rounding that V/(V/DT) adds cell by cell: within 1e-12 relative."""
    case = "auto_dam"
    st = fx.static(case)
    tp = fx.time_params(st["DECK"])
    tr = fx.trajectory(case)
    worst = 0.0
    for s in fx.steps(case):
        if s + 1 > int(tr["STEP"][-1]):
            continue
        rec = fx.step(case, s)["f1_out"]
        lim = min(_cell_limits(st, rec), tp["dtmax"] / tp["dtsafe"])
        dt = max(min(lim * tp["dtsafe"], 1.2 * tr["DT_FULL"][s]), tp["dtmin"])
        ref = tr["DT_FULL"][s + 1]
        worst = max(worst, abs(dt - ref) / ref)
    assert worst <= 1e-12, worst


def _table2(pkg, st):
    """The second wave-maker table (FUNC MATRIX2) of a case, if it has one."""
    return pkg.mtb.parse(str(st["FILE_MTB2"])) if "FILE_MTB2" in st else None


def _wave_fields(pkg, case, rec, bcf, damps=True):
    """Oracle fields with the deck's models around the step: wave-maker and
    radiation sides, the drag table and the damping zones."""
    st = fx.static(case)
    p = _params(st, rec)
    F = _fields(st, rec, p, bcf)
    d = pkg.deck.parse(str(st["DECK"]))
    if d.waves:
        F.set_waves(d.waves, pkg.mtb.parse(str(st["FILE_MTB"])), d.wvlvl, _table2(pkg, st))
    F.set_opens(d.opens)
    F.set_drag_df(d.drag_df)
    if damps:
        F.set_damps(d.damps)
    return st, p, F


@pytest.mark.parametrize("case", ["wave_flume", "wave_open", "wave_off", "wave_only", "wave_two"])
def test_wave_maker_phases_bitexact(pkg, case):
    """MODEL WAVE-BC X- FUNC MATRIX, and in wave_open a radiation boundary
    (MODEL OPEN-BC X+ FUNC TYPE1) at the far end: V1CAL (the wave-maker and
    the radiation side set the boundary velocities inside it) and F1CAL at
    every recorded step equal the reference's bit for bit."""
    for s in fx.steps(case):
        r = fx.step(case, s)
        st, p, F = _wave_fields(pkg, case, r["v1_in"], r["f1_in"]["BCF"])
        F.st.tnow = float(r["v1_in"]["TNOW"][0])
        assert F.predictor() == 0
        for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS")):
            assert np.array_equal(F[k], r["ps_in"][kk]), (s, k)
        st, p, F = _wave_fields(pkg, case, r["v1_in"], r["f1_in"]["BCF"])
        F.st.tnow = float(r["v1_in"]["TNOW"][0])
        rc, pt = F.v1cal()
        assert rc == 0
        for k in ("UU", "VV", "WW", "PP", "BCU", "BCV", "BCW", "BCP"):
            assert np.array_equal(F[k], r["v1_out"][k]), (s, k)
        st, p, F = _wave_fields(pkg, case, r["f1_in"], None)
        F.st.tnow = float(r["f1_in"]["TNOW"][0])   # VF_BWFFSF's level (LEVEL-ONLY)
        assert F.f1cal() == 0
        for k in ("FF", "NF", "BCF", "UU"):
            assert np.array_equal(F[k], r["f1_out"][k]), (s, k)


@pytest.mark.parametrize("case", ["wave_flume", "wave_open", "wave_off", "wave_only", "wave_two"])
def test_wave_maker_free_run_bitexact(pkg, case):
    """The wave flume (400 steps) and the flume with the radiation end (1000
    steps) from the reference's initial state: ITR and ||div|| every step and
    the fields at the recorded steps, bit for bit."""
    r1 = fx.step(case, 1)
    st, p, F = _wave_fields(pkg, case, r1["v1_in"], r1["f1_in"]["BCF"])
    tr = fx.trajectory(case)
    kept = set(fx.steps(case))
    for s in range(1, int(tr["STEP"][-1]) + 1):
        rc, _ = F.step()
        assert rc == 0, s
        assert F.st.icgitr == tr["ITR"][s], s
        assert F.cdiv00() == tr["DV_FULL"][s], s
        if s in kept:
            ref = fx.step(case, s)["f1_out"]
            for k in ("UU", "VV", "WW", "PP", "FF", "NF", "BCU", "BCW"):
                assert np.array_equal(F[k], ref[k]), (s, k)


def test_drag_df_predictor_bitexact(pkg):
    """VF_VGENE's Dupuit-Forchheimer branch (OPTION DRAG-DF, vf_vgene.f:111-146):
    the predictor's US/VS/WS equal the reference's at every recorded step of
    the two-stone rubble block, whose faces reach every row of the table
    (none below the first diameter, a middle row, the last row), while the
    CD form of the same arrays differs inside the block."""
    case = "porous_df"
    st = fx.static(case)
    drag = pkg.deck.parse(str(st["DECK"])).drag_df
    assert drag is not None and len(drag[1]) == 3
    for s in fx.steps(case):
        r = fx.step(case, s)
        p = _params(st, r["v1_in"])
        F = _fields(st, r["v1_in"], p, r["f1_in"]["BCF"])
        F.set_drag_df(drag)
        assert F.predictor() == 0
        for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS")):
            assert np.array_equal(F[k], r["ps_in"][kk]), (s, k)
        if s > 1:
            G = _fields(st, r["v1_in"], p, r["f1_in"]["BCF"])
            assert G.predictor() == 0
            assert not np.array_equal(G["UU"], r["ps_in"]["US"]), s


def test_drag_df_free_run_bitexact(pkg):
    """The 60 steps of the run with the Dupuit-Forchheimer drag from the
    reference's initial state: ITR each step and every field at the recorded
    steps equal the reference's."""
    case = "porous_df"
    st = fx.static(case)
    r1 = fx.step(case, 1)
    p = _params(st, r1["v1_in"])
    F = _fields(st, r1["v1_in"], p, r1["f1_in"]["BCF"])
    F.set_drag_df(pkg.deck.parse(str(st["DECK"])).drag_df)
    tr = fx.trajectory(case)
    for s in range(1, int(tr["STEP"][-1]) + 1):
        rc, _ = F.step()
        assert rc == 0
        assert F.st.icgitr == tr["ITR"][s], s
        if s in fx.steps(case):
            ref = fx.step(case, s)["f1_out"]
            for k in ("UU", "VV", "WW", "PP", "FF", "NF"):
                assert np.array_equal(F[k], ref[k]), (s, k)


DAMP = ["wave_damp", "damp3d"]


@pytest.mark.parametrize("case", DAMP)
def test_damping_predictor_bitexact(pkg, case):
    """VF_VEULER's implicit damping (MODEL DAMP, vf_veuler.f:131-166, 187-221,
    243-277): the predictor's US/VS/WS equal the reference's at every recorded
    step -- a degree-4 zone at the flume's end (the power evaluated by
    squaring), and two overlapping zones of degrees 1 and 3 with their own
    PARAM-Z in the 3-D tank -- while the same step without the zones
    differs."""
    for s in fx.steps(case):
        r = fx.step(case, s)
        for damps in (True, False):
            st, p, F = _wave_fields(pkg, case, r["v1_in"], r["f1_in"]["BCF"], damps)
            F.st.tnow = float(r["v1_in"]["TNOW"][0])
            assert F.predictor() == 0
            same = [np.array_equal(F[k], r["ps_in"][kk]) for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS"))]
            if damps:
                assert all(same), (s, same)
            elif s > 1:
                assert not all(same), s


@pytest.mark.parametrize("case", DAMP)
def test_damping_free_run_bitexact(pkg, case):
    """Both damping decks from the reference's initial state to their last
    step (600 and 60): ITR and ||div|| every step and the fields at the
    recorded steps, bit for bit."""
    r1 = fx.step(case, 1)
    st, p, F = _wave_fields(pkg, case, r1["v1_in"], r1["f1_in"]["BCF"])
    tr = fx.trajectory(case)
    kept = set(fx.steps(case))
    for s in range(1, int(tr["STEP"][-1]) + 1):
        rc, _ = F.step()
        assert rc == 0, s
        assert F.st.icgitr == tr["ITR"][s], s
        assert F.cdiv00() == tr["DV_FULL"][s], s
        if s in kept:
            ref = fx.step(case, s)["f1_out"]
            for k in ("UU", "VV", "WW", "PP", "FF", "NF"):
                assert np.array_equal(F[k], ref[k]), (s, k)

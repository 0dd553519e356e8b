"""The wave-makers (MODEL WAVE-BC X- FUNC MATRIX with a LEVEL-ON <name>.mtb
table; FUNC STOKES, STK-CND and STREAM, the theory waves) and the radiation boundary
(MODEL OPEN-BC X+ FUNC TYPE1) on the GPU.

Fixtures (tests/golden/make_golden.py): "wave_flume", a 60x4x20 flume at rest
at depth 0.2 whose X- side follows a solitary-wave table, 400 steps recorded
at 1, 2, 100, 200, 300, 400; "wave_open", the same on 40x4x20 with a
radiation boundary at the far end that the wave leaves through, 1000 steps
recorded at 1, 2, 250, 500, 750, 1000; "stokes_flume" and "cnoidal_flume",
40x4x20 flumes driven by a fifth-order Stokes and (by the Ursell number) a
third-order cnoidal wave, 600 steps recorded at 1, 2, 100, 300, 600;
"stream_flume", the same flume driven by the fifth-order stream-function wave
(FUNC STREAM 5, H = 3 cm, T = 1.2 s), recorded at the same steps;
"wave_off", the 40-cell flume driven by the same wave as a LEVEL-OFF table
(no level column: profiles unscaled), 400 steps recorded at 1, 2, 100, 200,
400; "wave_only", the same flume with its x- side held at the wave's level by
a LEVEL-ONLY table (faces taken as free faces, the cells next to the side
filled to the level); "wave_two", the flume with both ends driven, X- by
the LEVEL-ON table and X+ by a second table (FUNC MATRIX2, <name>.mtb2)
holding a fixed level; "oblique_basin", a 30x16x16 basin with the Stokes wave at 25 degrees to the
side's normal (MODEL WAVE-BC ANGLE), 300 steps recorded at 1, 2, 100, 300.  The
predictor and the boundary pass
that sets the special faces have no reduction, so they are compared bit for
bit, as is F1CAL (those faces take F like FREE); whole V1CAL and free-run
steps go through the Poisson solve and are compared to the plain-form bars.
"""
import numpy as np
import pytest

import fixtures as fx
from helpers import assert_close, assert_same_int, assert_vel_close, rel_err

pytestmark = pytest.mark.gpu

CASES = ["wave_flume", "wave_open", "wave_off", "wave_only", "wave_two", "stokes_flume",
         "cnoidal_flume", "oblique_basin", "stream_flume"]
RSL = {"wave_flume": (100, 400, 100), "wave_open": (250, 1000, 250),
       "stokes_flume": (300, 600, 300), "cnoidal_flume": (300, 600, 300), "stream_flume": (300, 600, 300),
       "oblique_basin": (100, 300, 200), "wave_off": (200, 400, 200),
       "wave_only": (200, 400, 200), "wave_two": (200, 400, 200)}
STATE = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF", "INDC")
KEYS = ("rho0", "grz0", "anu0", "scmvp", "flower", "fupper", "plower", "cgpara", "cgepsa",
        "cgepsr", "cgdiv", "icgtyp", "icgmax", "wbub", "idrop")


def _ctx(pkg, st, dtnow, tnow):
    numi, numj, numk = st["GGV"].shape
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    c = pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1])
    c.set_params(**{k: p[k] for k in KEYS}, dtnow=dtnow, tnow=tnow)
    c.upload_static(st)
    d = pkg.deck.parse(str(st["DECK"]))
    table = pkg.mtb.parse(str(st["FILE_MTB"])) if "FILE_MTB" in st else None
    table2 = pkg.mtb.parse(str(st["FILE_MTB2"])) if "FILE_MTB2" in st else None
    c.set_waves(d.waves, table, d.wvlvl, table2)
    c.set_opens(d.opens)
    return c, d


def _v1_state(r):
    a = {k: r["v1_in"][k] for k in STATE if k != "BCF"}
    a["BCF"] = r["f1_in"]["BCF"]
    return a


def _wave_faces(st):
    """The wave-maker and radiation faces."""
    t = np.asarray(st["INDB"])[2]
    return (t == 5) | (t == 7)


@pytest.mark.parametrize("case", CASES)
def test_init_with_special_sides(pkg, case):
    """cad_init_state with the wave-maker at TNOW = 0 and the radiation side's
    initial pass: the reference's initial state bit for bit."""
    st = fx.static(case)
    lay = pkg.layout.from_text(str(st["DECK"]))
    c, d = _ctx(pkg, st, lay.params()["dtnow"], 0.0)
    with c:
        c.upload_state(lay.raw_state())
        c.init_state(d.wvlvl, d.uini)
        out = c.download_state()
    r1 = fx.step(case, 1)
    for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC"):
        assert np.array_equal(out[k], r1["v1_in"][k]), k
    assert np.array_equal(out["BCF"], r1["f1_in"]["BCF"])


@pytest.mark.parametrize("case", CASES)
def test_predictor_with_wave_faces_bitexact(pkg, case):
    """VF_VFLXDU..VF_VEULER then VF_BWUWN with the wave-maker at the step's
    TNOW: every velocity and the wave faces' boundary values bit for bit."""
    st = fx.static(case)
    wave = _wave_faces(st)
    for s in fx.steps(case):
        r = fx.step(case, s)
        print(f"{case} step {s} predictor", flush=True)
        c, d = _ctx(pkg, st, float(r["v1_in"]["DTNOW"][0]), float(r["v1_in"]["TNOW"][0]))
        with c:
            c.upload_state(_v1_state(r))
            c.predictor()
            out = c.download_state()
            level = c.wave_level(1)
        for k, kk in (("UU", "US"), ("VV", "VS"), ("WW", "WS")):
            assert np.array_equal(out[k], r["ps_in"][kk]), f"step {s} {k}: {rel_err(out[k], r['ps_in'][kk]):.3e}"
        if case not in ("wave_only", "wave_two"):
            # a fixed level's faces are free faces, reset after the solve
            for k in ("BCU", "BCV", "BCW"):
                assert np.array_equal(out[k][wave], r["v1_out"][k][wave]), f"step {s} {k}"
        assert np.isfinite(level) and abs(level) < d.waves[1]["depth"]


@pytest.mark.parametrize("case", CASES)
def test_v1cal_with_wave_faces(pkg, case):
    """Whole VF_V1CAL: the fields to the plain-form bars, the wave faces'
    boundary values (set before the solve) bit for bit."""
    st = fx.static(case)
    wave = _wave_faces(st)
    for s in fx.steps(case):
        r = fx.step(case, s)
        print(f"{case} step {s} v1cal", flush=True)
        c, _ = _ctx(pkg, st, float(r["v1_in"]["DTNOW"][0]), float(r["v1_in"]["TNOW"][0]))
        with c:
            c.upload_state(_v1_state(r))
            info = c.v1cal()
            out = c.download_state()
        ref = r["v1_out"]
        assert info["icgitr"] == int(ref["ICGITR"][0]), f"step {s}"
        assert_vel_close(f"step {s} velocity", out, ref, 1e-10)
        assert_close(f"step {s} PP", out["PP"], ref["PP"], 1e-10)
        if case in ("wave_only", "wave_two"):
            # set from the solved pressure: to the bar, against the faces'
            # velocity scale (v is round-off in this quasi-2-D flume)
            scale = max(np.abs(ref[k][wave]).max() for k in ("BCU", "BCV", "BCW"))
            for k in ("BCU", "BCV", "BCW"):
                err = np.abs(out[k][wave] - ref[k][wave]).max() / scale
                assert err <= 1e-10, f"step {s} {k}: {err:.3e}"
        else:
            for k in ("BCU", "BCV", "BCW"):
                assert np.array_equal(out[k][wave], ref[k][wave]), f"step {s} {k}"


@pytest.mark.parametrize("case", CASES)
def test_f1cal_with_wave_faces_bitexact(pkg, case):
    """VF_F1CAL with the wave faces' F taken like FREE (vf_bwff.f:41-76)."""
    st = fx.static(case)
    for s in fx.steps(case):
        r = fx.step(case, s)
        fin, ref = r["f1_in"], r["f1_out"]
        print(f"{case} step {s} f1cal", flush=True)
        c, _ = _ctx(pkg, st, float(fin["DTNOW"][0]), float(fin["TNOW"][0]))
        with c:
            c.upload_state({k: fin[k] for k in STATE})
            c.f1cal()
            out = c.download_state()
        for k in ("FF", "BCF", "UU", "VV", "WW"):
            assert np.array_equal(out[k], ref[k]), f"step {s} {k}: {rel_err(out[k], ref[k]):.3e}"
        assert_same_int(f"step {s} NF", out["NF"], ref["NF"])


@pytest.mark.parametrize("case", CASES)
def test_wave_flume_free_run(pkg, tmp_path, case):
    """The deck through driver.run (the table read from <name>.mtb, the
    initial state by VF_CINIT's device part) to its last step: ITR every step
    as the reference's list file has it, the fields at the recorded steps
    from the .rsl to 1e-8."""
    st = fx.static(case)
    tr = fx.trajectory(case)
    a, b, n = RSL[case]
    deck = str(st["DECK"]) + f"\nFILE RSL STEP {a} {b} {n}\n"
    for ext in ("mtb", "mtb2"):
        if f"FILE_{ext.upper()}" in st:
            (tmp_path / f"flume.{ext}").write_text(str(st[f"FILE_{ext.upper()}"]))
    res = pkg.driver.run(deck, "flume", tmp_path)
    last = int(tr["STEP"][-1])
    assert res["nnow"] == last
    lines = (tmp_path / "flume.list").read_text().splitlines()
    itr = [int(ln.split("ITR =")[1]) for ln in lines if "ITR =" in ln]
    assert itr == [int(x) for x in tr["ITR"][: len(itr)]] and len(itr) == last + 1
    recs = [r for r in pkg.rsl.read(tmp_path / "flume.rsl")][1:]
    assert [r["NNOW"] for r in recs] == list(range(a, b + 1, n))
    if case == "wave_flume":
        # the .tran (FILE TRN in the deck): the probe list exactly, the values
        # every 20 steps within 1e-8 of each probe's scale; the wave-maker's
        # level (W-LEVEL ANS X-) depends on TNOW only, so to print precision
        with np.load(fx.HERE / f"{case}_tran.npz") as z:
            ref = z["TRAN"].tobytes().decode().splitlines()
        mine = (tmp_path / "flume.tran").read_text().splitlines()
        nprobe = sum(1 for ln in ref if ln.split() and ln.split()[0].isdigit())
        assert mine[: nprobe + 1] == ref[: nprobe + 1]
        ta = np.array([[float(x) for x in ln.split()] for ln in mine[nprobe + 1:]])
        tb = np.array([[float(x) for x in ln.split()] for ln in ref[nprobe + 1:]])
        assert ta.shape == tb.shape == (21, nprobe + 1)
        assert np.array_equal(ta[:, :2], tb[:, :2])
        scale = np.maximum(np.abs(tb).max(axis=0), 1e-30)
        assert (np.abs(ta - tb) / scale).max() <= 1e-8
    for rec in recs:
        s = rec["NNOW"]
        ref = fx.step(case, s)["f1_out"]
        assert_vel_close(f"step {s} velocity", rec, ref, 1e-8)
        for k in ("PP", "FF"):
            assert_close(f"step {s} {k}", rec[k], ref[k], 1e-8)
        assert_same_int(f"step {s} NF", rec["NF"], ref["NF"])


def test_special_side_errors(pkg):
    """The wave-maker and radiation entry points refuse what the reference
    would stop on or what would leave faces unset: stepping a context whose
    INDB has wave or radiation faces before their sides are set, a side
    before its table, a side that is taken, bad side numbers and values, a
    table whose depths do not increase."""
    st = fx.static("wave_open")
    d = pkg.deck.parse(str(st["DECK"]))
    table = pkg.mtb.parse(str(st["FILE_MTB"]))
    E = pkg.capi.CadError
    lay = pkg.layout.from_text(str(st["DECK"]))
    numi, numj, numk = st["GGV"].shape
    with pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1]) as c:
        p = lay.params()
        c.set_params(**{k: v for k, v in p.items() if k in c.params})
        c.upload_static(st)
        c.upload_state(lay.raw_state())
        with pytest.raises(E) as e:
            c.init_state(d.wvlvl, d.uini)          # type-5 faces, no wave side yet
        assert e.value.code == -3
        lib = pkg.capi.lib()
        assert lib.cad_set_wave(c._h, 1, 2, numj - 1, 0.2, 2.0, 0.0, 0.2) == -3   # no table yet
        with pytest.raises(E):
            bad = pkg.mtb.Table(table.kind, table.t0, table.zz[::-1].copy(), table.tt, table.hh,
                                table.un, table.ut)
            c.set_waves(d.waves, bad, d.wvlvl)
        c.set_waves(d.waves, table, d.wvlvl)
        with pytest.raises(E) as e:
            c.init_state(d.wvlvl, d.uini)          # type-7 faces, no radiation side yet
        assert e.value.code == -3
        assert lib.cad_set_open(c._h, 1, 0.2, 2.0) == -1      # the wave-maker's side
        assert lib.cad_set_open(c._h, 5, 0.2, 2.0) == -1      # no such side
        assert lib.cad_set_open(c._h, 2, 0.0, 2.0) == -1      # no depth
        assert lib.cad_set_wave(c._h, 2, 2, numj, 0.2, 2.0, 0.0, 0.2) == -1   # columns past the side
        c.set_opens(d.opens)
        with pytest.raises(E):
            c.wave_level(2)                        # a radiation side has no wave level
        c.init_state(d.wvlvl, d.uini)
        # the level of VF_CINIT's pass: the table's level at its phase origin
        # (the same table as the wave flume's, whose .tran prints 3.61819724E-06)
        assert abs(c.wave_level(1) - 3.61819724e-06) <= 1e-14


def test_theory_and_second_table_errors(pkg):
    """cad_set_wave_theory refuses a bad side or function, missing depth,
    height or period, columns outside the side and a radiation side;
    cad_set_wave2 needs its table first, and cad_set_wave_table2 checks its
    arrays as the first table's entry does."""
    st = fx.static("stokes_flume")
    lay = pkg.layout.from_text(str(st["DECK"]))
    numi, numj, numk = st["GGV"].shape
    lib = pkg.capi.lib()
    with pkg.capi.Context(numi, numj, numk, st["INDB"].shape[1]) as c:
        p = lay.params()
        c.set_params(**{k: v for k, v in p.items() if k in c.params})
        c.upload_static(st)
        good = (2, numj - 1, 0.2, 0.02, 1.0, 1.0, 0.2, 0.0, 0.0, 0.0)
        assert lib.cad_set_wave_theory(c._h, 1, -2, *good) == 0
        assert lib.cad_set_wave_theory(c._h, 5, -2, *good) == -1
        assert lib.cad_set_wave_theory(c._h, 1, -3, *good) == -1
        assert lib.cad_set_wave_theory(c._h, 1, 1, *good) == -1
        for i in (2, 3, 4):   # depth, height, period
            bad = list(good)
            bad[i] = 0.0
            assert lib.cad_set_wave_theory(c._h, 1, -2, *bad) == -1
        assert lib.cad_set_wave_theory(c._h, 1, -2, 1, numj - 1, *good[2:]) == -1
        assert lib.cad_set_wave_theory(c._h, 1, -2, 2, numj, *good[2:]) == -1
        assert lib.cad_set_open(c._h, 2, 0.2, 1.0) == 0
        assert lib.cad_set_wave_theory(c._h, 2, -2, *good) == -1    # a radiation side
        assert lib.cad_set_wave2(c._h, 3, 2, numi - 1, 0.2, 1.0, 0.0, 0.2) == -3   # no table 2
        t = np.linspace(0.0, 1.0, 5)
        assert lib.cad_set_wave_table2(c._h, 1, 5, None, t.ctypes.data_as(pkg.capi.PD), None,
                                       None, None, 0.0) == -1
        z = np.array([0.0, -1.0])
        assert lib.cad_set_wave_table2(c._h, 2, 5, z.ctypes.data_as(pkg.capi.PD),
                                       t.ctypes.data_as(pkg.capi.PD), None,
                                       np.zeros(10).ctypes.data_as(pkg.capi.PD),
                                       np.zeros(10).ctypes.data_as(pkg.capi.PD), 0.0) == -1

"""The per-step outputs in the reference's formats (rsl.py) against the files
the reference itself wrote (tests/golden/make_golden.py runs it with FILE RSL
and keeps the .rsl bytes and the list-file text).

The writer is fed from the phase records of the same run (an independent
source) and must reproduce the reference's file byte for byte; the reader must
give back those records; the list lines must equal the reference's text.
"""
import numpy as np
import pytest

import fixtures as fx


def _rsl_bytes(tmp_path, case):
    with np.load(fx.HERE / f"{case}_rsl.npz") as z:
        raw = z["RSL"].tobytes()
    p = tmp_path / f"{case}.rsl"
    p.write_bytes(raw)
    return p, raw


def _after_step(case, s):
    """The state the reference writes after step s: the F1CAL output of step
    s, or for s = 0 the input of step 1's V1CAL (with its F1CAL's BCF)."""
    if s == 0:
        r = fx.step(case, 1)
        rec = dict(r["v1_in"])
        rec["BCF"] = r["f1_in"]["BCF"]
        for k in ("TBUB", "DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW"):
            rec[k] = r["f1_in"][k]
        return rec
    return fx.step(case, s)["f1_out"]


def _info(rec):
    return {"fsum": float(rec["FSUM"][0]), "fcut": float(rec["FCUT"][0]),
            "cgbnrm": float(rec["CGBNRM"][0]), "cgxnrm": float(rec["CGXNRM"][0]),
            "icgitr": int(rec["ICGITR"][0])}


@pytest.mark.parametrize("case,steps,idrop", [("tiny", [0, 1, 2, 3, 4, 5], 0),
                                              ("tdoor", [100, 200, 300], 1)])
def test_writer_reproduces_reference_file(pkg, tmp_path, case, steps, idrop):
    ref_path, raw = _rsl_bytes(tmp_path, case)
    st = fx.static(case)
    out = tmp_path / "ours.rsl"
    with pkg.rsl.Writer(out, st["GGV"].shape, st["INDB"].shape[1], st["GGV"], idrop=idrop) as w:
        tr = fx.trajectory(case)
        for s in steps:
            rec = _after_step(case, s)
            tnow = float(tr["TNOW_FULL"][s]) if "TNOW_FULL" in tr else float(rec["TNOW"][0])
            dtnow = float(tr["DT_FULL"][s]) if "DT_FULL" in tr else float(rec["DTNOW"][0])
            w.write_step(s, tnow, dtnow, _info(rec), rec, rec)
    mine = out.read_bytes()
    assert len(mine) == len(raw)
    assert mine == raw


@pytest.mark.parametrize("case", ["tiny", "tdoor"])
def test_reader_gives_the_recorded_state(pkg, tmp_path, case):
    path, _ = _rsl_bytes(tmp_path, case)
    it = pkg.rsl.read(path)
    hdr = next(it)
    st = fx.static(case)
    assert hdr["version"] == pkg.rsl.VERSION and hdr["shape"] == st["GGV"].shape
    assert np.array_equal(hdr["GGV"], st["GGV"])
    n = 0
    for rec in it:
        s = rec["NNOW"]
        ref = _after_step(case, s)
        for k in ("NF", "UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "TBUB"):
            assert np.array_equal(rec[k], ref[k]), (s, k)
        assert rec["FSUM"] == ref["FSUM"][0] and rec["ICGITR"] == ref["ICGITR"][0]
        n += 1
    assert n > 0


def test_restart_state_matches_the_record(pkg, tmp_path):
    """VF_IR1INP: droplet fields rebuilt from the droplet records, INDC from NF."""
    path, _ = _rsl_bytes(tmp_path, "tdoor")
    r = pkg.rsl.restart(path, 200)
    ref = fx.step("tdoor", 200)["f1_out"]
    for k in ("UU", "VV", "WW", "PP", "FF", "NF", "INDC"):
        assert np.array_equal(r["state"][k], ref[k]), k
    for k in ("TBUB", "DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW"):
        assert np.array_equal(r["tdoor"][k], ref[k]), k
    assert np.count_nonzero(r["tdoor"]["DROPTX"] >= 0.0) > 0
    with pytest.raises(KeyError):
        pkg.rsl.restart(path, 150)


@pytest.mark.parametrize("case", ["tiny", "tdoor"])
def test_list_lines_match_reference_text(pkg, case):
    tr = fx.trajectory(case)
    lines = [str(x) for x in tr["LIST"]]
    for s in fx.steps(case):
        rec = fx.step(case, s)["f1_out"]
        a, b = pkg.rsl.list_lines(s, float(rec["TNOW"][0]), float(rec["DTNOW"][0]),
                                  float(rec["FSUM"][0]), float(rec["FCUT"][0]), float(tr["DV_FULL"][s]),
                                  float(rec["CGBNRM"][0]), float(rec["CGXNRM"][0]), int(rec["ICGITR"][0]))
        assert a + "\n" + b == lines[s], (s, a, b, lines[s])


def test_fortran_e_format(pkg):
    f = pkg.rsl._e12_5
    assert f(0.0) == " 0.00000E+00"
    assert f(-1.234567e-3) == "-1.23457E-03"
    assert f(1.0e-100) == " 1.00000-100"
    assert f(123456.0) == " 1.23456E+05"


@pytest.mark.parametrize("case,steps", [("tiny", [0, 1, 2, 3, 4, 5]), ("tdoor", [100, 200, 300])])
def test_graphics_file_matches_reference(pkg, case, steps):
    """grp.Writer fed from the phase records reproduces the reference's .grp
    (the tdoor run uses FILE GRP AREA, so the output box and the graphics
    order of boundary faces are exercised)."""
    import tempfile
    from pathlib import Path
    with np.load(fx.HERE / f"{case}_grp.npz") as z:
        raw = z["GRP"].tobytes()
    st = fx.static(case)
    d = pkg.deck.parse(str(st["DECK"]))
    tr = fx.trajectory(case)
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "ours.grp"
        with pkg.grp.Writer(out, st["XX"], st["YY"], st["ZZ"], st["INDB"], st["GGV"], d.grp_area) as w:
            for s in steps:
                rec = _after_step(case, s)
                w.write_step(s, float(tr["TNOW_FULL"][s]) if "TNOW_FULL" in tr else float(rec["TNOW"][0]), rec)
        mine = out.read_bytes()
    assert len(mine) == len(raw)
    assert mine == raw


def test_time_series_matches_reference(pkg):
    """tran.py from the recorded states reproduces the reference's .tran:
    the probe list, the column line and the value lines of every recorded
    step that FILE TRN STEP 0 300 10 selects (water levels, point values,
    box min/max/mean/integral, pressure forces)."""
    case = "tdoor"
    with np.load(fx.HERE / f"{case}_tran.npz") as z:
        text = z["TRAN"].tobytes().decode()
    ref_lines = text.splitlines(keepends=True)
    st = fx.static(case)
    d = pkg.deck.parse(str(st["DECK"]))
    shape = st["GGV"].shape
    probes = [pkg.tran.parse(w, shape) for w in d.trn_probes]
    hdr = pkg.tran.header(probes)
    assert hdr == "".join(ref_lines[: len(probes) + 1])
    body = ref_lines[len(probes) + 1:]
    tr = fx.trajectory(case)
    checked = 0
    for s in fx.steps(case):
        if s % 10:
            continue
        rec = _after_step(case, s)
        vals = pkg.tran.values(probes, st, rec, d.wvlvl)
        got = pkg.tran.line(float(tr["TNOW_FULL"][s]), vals)
        assert got == body[s // 10], (s, got, body[s // 10])
        checked += 1
    assert checked >= 6


def test_wave_flume_time_series_matches_reference(pkg):
    """The wave flume's .tran (FILE TRN STEP 0 400 20 with W-LEVEL ANS X-,
    water levels, points and an F mean): the probe list and column line, and
    the value lines of step 0 and the recorded steps.  The wave-maker level
    (BCTYP(7,1)) of a line is the oracle's boundary pass at that step's TNOW,
    which pins the oracle's table level and ramp against the reference's file."""
    import pyoracle as po
    case = "wave_flume"
    with np.load(fx.HERE / f"{case}_tran.npz") as z:
        text = z["TRAN"].tobytes().decode()
    ref_lines = text.splitlines(keepends=True)
    st = fx.static(case)
    d = pkg.deck.parse(str(st["DECK"]))
    shape = st["GGV"].shape
    probes = [pkg.tran.parse(w, shape) for w in d.trn_probes]
    assert pkg.tran.header(probes) == "".join(ref_lines[: len(probes) + 1])
    body = ref_lines[len(probes) + 1:]
    tr = fx.trajectory(case)
    table = pkg.mtb.parse(str(st["FILE_MTB"]))

    def level(rec, tnow):
        p = fx.deck_params(st["DECK"])
        p.update(fx.solver_params(st))
        p["dtnow"] = float(rec["DTNOW"][0])
        F = po.Fields(fx.state_arrays(st, rec), p)
        F.set_waves(d.waves, table, d.wvlvl)
        F.st.tnow = tnow
        assert F.predictor() == 0
        return F.st.wv_level[0]

    checked = 0
    for s in [0] + [s for s in fx.steps(case) if s % 20 == 0]:
        if s == 0:
            r1 = fx.step(case, 1)
            rec = dict(r1["v1_in"])
            rec["BCF"] = r1["f1_in"]["BCF"]
            lv = level(r1["v1_in"], 0.0)
        else:
            r = fx.step(case, s)
            rec = r["f1_out"]
            lv = level(r["v1_in"], float(r["v1_in"]["TNOW"][0]))
        vals = pkg.tran.values(probes, st, rec, d.wvlvl, {1: lv})
        got = pkg.tran.line(float(tr["TNOW_FULL"][s]), vals)
        assert got == body[s // 20], (s, got, body[s // 20])
        checked += 1
    assert checked == 5

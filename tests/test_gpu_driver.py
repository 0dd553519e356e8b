"""Whole runs from the reference's input deck through driver.run on the GPU:
the result file and list-file lines it writes, and a restart from the
reference's own result file (FILE RES)."""
import numpy as np
import pytest

import fixtures as fx
from helpers import assert_close, assert_same_int, assert_vel_close

pytestmark = pytest.mark.gpu


def _ref_rsl(tmp_path, case, name):
    with np.load(fx.HERE / f"{case}_rsl.npz") as z:
        raw = z["RSL"].tobytes()
    p = tmp_path / name
    p.write_bytes(raw)
    return p


def test_run_writes_the_reference_outputs(pkg, tmp_path):
    """tiny (FILE RSL STEP 0 5 1): the same records in the .rsl, fields to the
    plain-form bars, flags exactly; list lines with the same step, ITR and
    numbers to print precision."""
    case = "tiny"
    deck = str(fx.static(case)["DECK"])
    res = pkg.driver.run(deck, "run", tmp_path)
    assert res["nnow"] == 5
    ref = list(pkg.rsl.read(_ref_rsl(tmp_path, case, "ref.rsl")))
    mine = list(pkg.rsl.read(tmp_path / "run.rsl"))
    assert len(mine) == len(ref) == 7
    assert np.array_equal(mine[0]["GGV"], ref[0]["GGV"])
    for a, b in zip(mine[1:], ref[1:]):
        s = b["NNOW"]
        assert a["NNOW"] == s and a["ICGITR"] == b["ICGITR"]
        assert a["TNOW"] == b["TNOW"] and a["DTNOW"] == b["DTNOW"]
        assert_vel_close(f"step {s} velocity", a, b, 1e-10)
        for k in ("PP", "FF"):
            assert_close(f"step {s} {k}", a[k], b[k], 1e-10)
        assert_same_int(f"step {s} NF", a["NF"], b["NF"])
    with np.load(fx.HERE / f"{case}_grp.npz") as z:
        (tmp_path / "ref.grp").write_bytes(z["GRP"].tobytes())
    # the graphics file: its header exactly, each step's time and flags
    # exactly, its fields to the same bars as the result file
    gh, gsteps = pkg.grp.read(tmp_path / "run.grp")
    rh, rsteps = pkg.grp.read(tmp_path / "ref.grp")
    assert gh.keys() == rh.keys()
    for k in gh:
        assert np.array_equal(gh[k], rh[k]), f".grp header {k}"
    assert len(gsteps) == len(rsteps) == 6
    for a, b in zip(gsteps, rsteps):
        s = b["NNOW"]
        assert a.keys() == b.keys()
        assert a["NNOW"] == s and a["TNOW"] == b["TNOW"]
        assert_same_int(f".grp step {s} NF", a["NF"], b["NF"])
        assert_vel_close(f".grp step {s} velocity", a, b, 1e-10)
        if "BCU" in b:
            assert_vel_close(f".grp step {s} boundary velocity", a, b, 1e-10, ("BCU", "BCV", "BCW"))
        for k in ("PP", "FF", "BCP", "BCF"):
            if k in b:
                assert_close(f".grp step {s} {k}", a[k], b[k], 1e-10)
    lines = (tmp_path / "run.list").read_text().splitlines()
    ref_lines = [ln for x in fx.trajectory(case)["LIST"] for ln in str(x).splitlines()]
    assert len(lines) == len(ref_lines)
    for x, y in zip(lines, ref_lines):
        assert x.split(":")[0] == y.split(":")[0]
        fx_ = [w for w in x.replace("=", " ").split() if w[0].isdigit() or w[0] == "-"]
        fy_ = [w for w in y.replace("=", " ").split() if w[0].isdigit() or w[0] == "-"]
        assert len(fx_) == len(fy_)
        for u, v in zip(fx_, fy_):
            assert abs(float(u) - float(v)) <= 2e-5 * max(abs(float(v)), 1e-30), (x, y)


def test_restart_from_reference_result_file(pkg, tmp_path):
    """FILE RES 200 from the droplet case's .rsl: ten steps later the state
    matches the reference's step 210."""
    case = "tdoor"
    _ref_rsl(tmp_path, case, "rst.res")
    deck = str(fx.static(case)["DECK"])
    deck = "\n".join(ln for ln in deck.splitlines() if not ln.startswith("FILE RSL"))
    deck += "\nFILE RES 200\nFILE RSL STEP 210 210 1\n"
    res = pkg.driver.run(deck, "rst", tmp_path, nmax=10)
    assert res["nnow"] == 210
    ref = fx.step(case, 210)["f1_out"]
    assert res["tnow"] == pytest.approx(float(ref["TNOW"][0]), rel=1e-14)
    assert res["info"]["icgitr"] == int(ref["ICGITR"][0])
    recs = list(pkg.rsl.read(tmp_path / "rst.rsl"))
    assert len(recs) == 2 and recs[1]["NNOW"] == 210
    a = recs[1]
    assert_vel_close("step 210 velocity", a, ref, 1e-8)
    for k in ("PP", "FF"):
        assert_close(f"step 210 {k}", a[k], ref[k], 1e-8)
    assert_same_int("step 210 NF", a["NF"], ref["NF"])
    assert np.array_equal(a["TBUB"], ref["TBUB"])


def test_run_time_series_against_reference(pkg, tmp_path):
    """The droplet case (300 steps, bubbles and droplets on) run from its deck:
    the .tran it writes has the reference's probe list and, every 10 steps,
    probe values within 1e-8 of each probe's scale."""
    case = "tdoor"
    deck = str(fx.static(case)["DECK"])
    res = pkg.driver.run(deck, "td", tmp_path)
    assert res["nnow"] == 300
    with np.load(fx.HERE / f"{case}_tran.npz") as z:
        ref = z["TRAN"].tobytes().decode().splitlines()
    mine = (tmp_path / "td.tran").read_text().splitlines()
    nprobe = sum(1 for ln in ref if ln.split() and ln.split()[0].isdigit())
    assert mine[: nprobe + 1] == ref[: nprobe + 1]
    a = np.array([[float(x) for x in ln.split()] for ln in mine[nprobe + 1:]])
    b = np.array([[float(x) for x in ln.split()] for ln in ref[nprobe + 1:]])
    assert a.shape == b.shape == (31, nprobe + 1)
    scale = np.maximum(np.abs(b).max(axis=0), 1e-30)
    # velocity-valued probes (point, min, max, mean of U/V/W) against the
    # velocity scale: a component that is round-off in this y-symmetric
    # flow is judged against the flow
    d = pkg.deck.parse(deck)
    vel = [n + 1 for n, w in enumerate(d.trn_probes)
           if w[0] in ("POINT", "MIN", "MAX", "AV") and w[1] in ("U", "V", "W")]
    scale[vel] = np.abs(b[:, vel]).max()
    col = (np.abs(a - b) / scale).max(axis=0)
    worst = int(np.argmax(col))
    row = int(np.argmax(np.abs(a[:, worst] - b[:, worst])))
    assert col.max() <= 1e-8, (list(np.round(col, 12)), worst, row, a[row, worst], b[row, worst],
                               mine[nprobe + 1 + row][:40])

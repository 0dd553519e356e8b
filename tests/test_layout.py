"""Deck reader and host-side field layout against the reference's own set-up.

The reference's set-up arrays were recorded from its VF_V1CAL arguments on the
first step (tests/golden/*_static.npz); the host layout must reproduce them
bit for bit, including the boundary-face numbering order of VF_CINDX.
"""
import numpy as np
import pytest

import fixtures as fx

CASES = ["tiny", "beach_small", "corner3d", "dam_a", "tdoor", "auto_dam", "subloop", "wave_flume",
         "porous_df", "wave_damp", "damp3d", "stokes_flume", "cnoidal_flume",
         "oblique_basin", "wave_off", "wave_only", "wave_two"]
STATIC = ("XX", "YY", "ZZ", "ANU", "CD0", "GGV", "GGX", "GGY", "GGZ", "GLV", "GLX", "GLY", "GLZ",
          "INDX", "INDY", "INDZ", "INDB", "BCVI")


@pytest.mark.parametrize("case", CASES)
def test_layout_matches_reference_setup(pkg, case):
    st = fx.static(case)
    lay = pkg.layout.from_text(str(st["DECK"]))
    a = lay.static_arrays()
    assert lay.numb == st["INDB"].shape[1]
    for k in STATIC:
        assert a[k].shape == st[k].shape, k
        assert np.array_equal(a[k], st[k]), k


def test_metrics_closed_form(pkg):
    """VF_CGRID on a stretched axis: dx, cx=(dx_i+dx_{i-1})/2, reciprocals, ghosts."""
    x = [0.0, 0.1, 0.3, 0.6, 1.0]
    m = pkg.layout.grid_metrics(x)
    dx = np.diff(x)
    assert m.shape == (6, 6)
    assert np.allclose(m[0, 1:], x)
    assert np.allclose(m[1, 1:5], dx)
    assert m[1, 0] == dx[0] and m[1, 5] == dx[-1]
    assert np.isclose(m[0, 0], -dx[0])
    assert np.allclose(m[2, 2:5], (dx[1:] + dx[:-1]) / 2)
    assert np.allclose(m[4, 1:], 1 / m[2, 1:])
    assert np.allclose(m[5, 1:], 1 / (2 * m[2, 1:]))
    assert m[2, 0] == m[4, 0] == m[5, 0] == 0.0


@pytest.mark.parametrize("text,exc", [
    ("GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nOPTION T-DOOR DROP FALL\n"
     "TIME CONST 1e-3\n", "DeckError"),
    ("GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nOPTION T-DOOR BUB\n"
     "TIME CONST 1e-3\n", "DeckError"),
    ("GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nOPTION T-DOOR BUB OFF\n"
     "OPTION T-DOOR DROP OFF\nTIME CONST 1e-3\nMODEL WAVE-BC X- FUNC STREAM 22\n", "UnsupportedDeck"),
    ("GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nOPTION T-DOOR BUB OFF\n"
     "OPTION T-DOOR DROP OFF\nTIME CONST 1e-3\nPARALLEL X 3\n", "UnsupportedDeck"),
    ("GRID X 0 1 2 END\nGRID Y 0 1 2 END\nFOO 1\n", "DeckError"),
])
def test_deck_rejects_outside_the_path(pkg, text, exc):
    with pytest.raises(getattr(pkg.deck, exc)):
        pkg.deck.parse(text)


def test_deck_timer_door_words(pkg):
    """OPTION T-DOOR as Source code/vf_iiopt.f:57-77 reads it, with the
    defaults of vf_a2dflt.f:151,205 (both models on)."""
    base = "GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nTIME CONST 1e-3\n"
    d = pkg.deck.parse(base)
    assert (d.wbub, d.idrop) == (0.2, 1)
    d = pkg.deck.parse(base + "OPTION T-DOOR BUB 0.5\nOPTION T-DOOR DROP OFF\n")
    assert (d.wbub, d.idrop) == (0.5, 0)
    d = pkg.deck.parse(base + "OPTION T-DOOR BUB OFF\nOPTION T-DOOR DROP FREE-RUNDOWN\n")
    assert (d.wbub, d.idrop) == (0.0, 1)
    lay = pkg.layout.from_text(base + "B.C. D VP SLIP\nB.C. D F FREE\n")
    assert (lay.params()["wbub"], lay.params()["idrop"]) == (0.2, 1)


def test_deck_drag_df_words(pkg):
    """OPTION DRAG-DF yu n and its n table lines as Source code/vf_iiopt.f:86-108
    reads them: the table lines are consumed by the option (a deck line after
    it is read as usual), and each check of the reference is an error."""
    base = "GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nTIME CONST 1e-3\n"
    assert pkg.deck.parse(base).drag_df is None
    d = pkg.deck.parse(base + "OPTION DRAG-DF 1.0E-6 2\n 0.01 1000 1.1\n"
                       "# a comment between rows\n 0.05 1500.0 2.0\nOPTION SUB-LOOP 2\n")
    assert d.drag_df == (pkg.deck.zstor("1.0E-6"), [(pkg.deck.zstor("0.01"), 1000.0, 1.1),
                                                    (pkg.deck.zstor("0.05"), 1500.0, 2.0)])
    assert d.loops == 2
    bad = ["OPTION DRAG-DF 1.0E-6\n",                     # no row count
           "OPTION DRAG-DF 0.0 1\n0.01 1000 1.1\n",      # YU <= 0
           "OPTION DRAG-DF 1.0E-6 0\n",                   # no rows
           "OPTION DRAG-DF 1.0E-6 11\n",                  # more than MAXDRG = 10
           "OPTION DRAG-DF 1.0E-6 1\n0.01 1000\n",       # two words in a row
           "OPTION DRAG-DF 1.0E-6 2\n0.01 1000 1.1\n",   # deck ends inside the table
           "OPTION DRAG-DF 1.0E-6 1\n0.0 1000 1.1\n",    # DR <= 0
           "OPTION DRAG-DF 1.0E-6 1\n0.01 -1 1.1\n",     # AP < 0
           "OPTION DRAG-DF 1.0E-6 1\n0.01 1000 -0.1\n"]  # BT < 0
    for t in bad:
        with pytest.raises(pkg.deck.DeckError):
            pkg.deck.parse(base + t)


def test_deck_theory_wave_words(pkg):
    """MODEL WAVE-BC FUNC STOKES / CNOIDAL / STK-CND / STREAM n and ANGLE
    (Source code/vf_iimdl.f:51-70, 107-111) with the set-up checks of
    vf_csetup.f:86-100: a theory wave needs DEPTH, HEIGHT and PERIOD, the table
    takes no ANGLE; STREAM takes an order 1..22 (22 has no defined result in
    the reference and is refused as unsupported)."""
    base = ("GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nTIME CONST 1e-3\n"
            "B.C. D VP SLIP\nB.C. D F FREE\n")
    full = ("MODEL WAVE-BC X- DEPTH 0.2\nMODEL WAVE-BC X- HEIGHT 0.02\n"
            "MODEL WAVE-BC X- PERIOD 1.0\n")
    for func in ("STOKES", "CNOIDAL", "STK-CND"):
        d = pkg.deck.parse(base + f"MODEL WAVE-BC X- FUNC {func}\n" + full)
        assert d.waves[1]["func"] == func and d.waves[1]["height"] == pkg.deck.zstor("0.02")
    with pytest.raises(pkg.deck.DeckError):      # no HEIGHT
        pkg.deck.parse(base + "MODEL WAVE-BC X- FUNC STOKES\nMODEL WAVE-BC X- DEPTH 0.2\n"
                       "MODEL WAVE-BC X- PERIOD 1.0\n")
    with pytest.raises(pkg.deck.DeckError):
        pkg.deck.parse(base + "MODEL WAVE-BC X- FUNC AIRY\n" + full)
    d = pkg.deck.parse(base + "MODEL WAVE-BC X- FUNC STREAM 5\n" + full)
    assert d.waves[1]["func"] == "STREAM" and d.waves[1]["order"] == 5
    assert pkg.capi.wave_func(d.waves[1]) == 5
    for func in ("STREAM", "STREAM 0", "STREAM 23"):   # SYNTAX ERROR / INVALID VALUE
        with pytest.raises(pkg.deck.DeckError):
            pkg.deck.parse(base + f"MODEL WAVE-BC X- FUNC {func}\n" + full)
    with pytest.raises(pkg.deck.UnsupportedDeck):
        pkg.deck.parse(base + "MODEL WAVE-BC X- FUNC STREAM 22\n" + full)
    d = pkg.deck.parse(base + "MODEL WAVE-BC X- FUNC STOKES\n" + full +
                       "MODEL WAVE-BC X- ANGLE 10.0 0.5 -0.25\n")
    assert (d.waves[1]["angle"], d.waves[1]["ax"], d.waves[1]["ay"]) == (10.0, 0.5, -0.25)
    with pytest.raises(pkg.deck.DeckError):        # the table refuses an angle
        pkg.deck.parse(base + "MODEL WAVE-BC X- FUNC MATRIX\nMODEL WAVE-BC X- DEPTH 0.2\n"
                       "MODEL WAVE-BC X- PERIOD 1.0\nMODEL WAVE-BC X- ANGLE 10.0 0.0 0.0\n")


def test_deck_damp_words(pkg):
    """MODEL DAMP as Source code/vf_iimdl.f:216-258 reads it, with the
    defaults of vf_a2dflt.f:150,189-204 and the set-up checks of
    vf_csetup.f:221-234 (a named side needs DEGREE, WIDTH and DEPTH)."""
    base = "GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\nTIME CONST 1e-3\n"
    assert pkg.deck.parse(base).damps == {}
    d = pkg.deck.parse(base + "MODEL DAMP Y+ DEGREE 2\nMODEL DAMP Y+ WIDTH 0.5\n"
                       "MODEL DAMP Y+ DEPTH 1.0\nMODEL DAMP Y+ PARAM-Z 0.3\n")
    assert d.damps == {4: dict(degree=2, pxy=0.6, pz=pkg.deck.zstor("0.3"),
                               width=pkg.deck.zstor("0.5"), depth=1.0)}
    bad = ["MODEL DAMP\n", "MODEL DAMP Z- DEGREE 1\n", "MODEL DAMP X- DEGREE\n",
           "MODEL DAMP X- DEGREE -1\n", "MODEL DAMP X- WIDTH 0.0\n", "MODEL DAMP X- COLOR 1\n",
           # named without DEGREE, WIDTH or DEPTH
           "MODEL DAMP X- WIDTH 0.5\nMODEL DAMP X- DEPTH 1.0\n",
           "MODEL DAMP X- DEGREE 1\nMODEL DAMP X- DEPTH 1.0\n",
           "MODEL DAMP X- DEGREE 1\nMODEL DAMP X- WIDTH 0.5\n"]
    for t in bad:
        with pytest.raises(pkg.deck.DeckError):
            pkg.deck.parse(base + t)


def test_workload_decks_parse(pkg):
    """The benchmark workloads are valid decks with the documented sizes."""
    for name, cells in (("runup_b", (512, 32, 96)), ("dam_a", (64, 16, 48))):
        case = pkg.configs.named(name)
        d = pkg.deck.parse(case.deck)
        assert d.ncells == cells


def test_deck_time_words(pkg):
    """TIME CONST / AUTO / LIMIT / END and OPTION SUB-LOOP as vf_iitime.f:28-60
    and vf_iiopt.f:41-45 read them (defaults vf_a2dflt.f:223-234)."""
    base = "GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\n"
    d = pkg.deck.parse(base)
    assert (d.idttyp, d.dtinit, d.dtsafe, d.dtmin, d.dtmax, d.loops) == (1, 1e-6, 0.1, 1e-20, 1e20, 1)
    d = pkg.deck.parse(base + "TIME AUTO 1e-4 0.3\nTIME LIMIT 1e-6 1e-2\nTIME END 7 2.5\nOPTION SUB-LOOP 3\n")
    assert (d.idttyp, d.dtinit, d.dtsafe, d.dtmin, d.dtmax, d.nend, d.tend, d.loops) == \
        (1, 1e-4, 0.3, 1e-6, 1e-2, 7, 2.5, 3)
    d = pkg.deck.parse(base + "TIME AUTO 1e-4 0.3\nTIME CONST 2e-3\n")
    assert (d.idttyp, d.dt_const) == (0, 2e-3)
    for bad in ("TIME LIMIT 1e-2 1e-3\n", "TIME AUTO 1e-4\n", "TIME FOO 1\n", "OPTION SUB-LOOP 0\n"):
        with pytest.raises(pkg.deck.DeckError):
            pkg.deck.parse(base + bad)


class _FakeCtx:
    """Stands in for capi.Context: a step advances TNOW by DTNOW."""

    def __init__(self, dts):
        self.p = {"tnow": 0.0, "dtnow": 0.0}
        self.dts = list(dts)
        self.calls = []

    def get_params(self):
        return dict(self.p)

    def set_params(self, **kw):
        self.p.update(kw)

    def cdtcal(self, nnow, *limits):
        self.calls.append(nnow)
        return self.dts[nnow]

    def step(self):
        self.p["tnow"] = self.p["tnow"] + self.p["dtnow"]
        return {"icgitr": 1}


def test_time_loop_end_rule(pkg):
    """vf_a1main.f:937-947: the step size is computed first, then the loop ends
    when NNOW >= NEND or TNOW + DT/2 >= TEND."""
    base = "GRID X 0 1 2 END\nGRID Y 0 1 2 END\nGRID Z 0 1 2 END\n"
    d = pkg.deck.parse(base + "TIME AUTO 1e-4 0.3\nTIME END 100 0.35\n")
    ctx = _FakeCtx([0.1] * 10)
    recs = list(pkg.timeloop.run(ctx, d))
    # t = 0.1, 0.2, 0.3; at t = 0.3 the next step would pass 0.35 - 0.05
    assert [r.nnow for r in recs] == [1, 2, 3] and ctx.calls == [0, 1, 2, 3]
    d = pkg.deck.parse(base + "TIME CONST 0.1\nTIME END 2 10.0\n")
    recs = list(pkg.timeloop.run(_FakeCtx([]), d))
    assert [r.nnow for r in recs] == [1, 2] and recs[-1].dtnow == 0.1


def test_number_reading_follows_the_reference(pkg):
    """deck.zstor reads numbers the reference's way (vf_zstor.f): digits
    accumulated in binary, divided by the power of ten of the decimals, then
    scaled by 10.0**exponent -- not the correctly rounded conversion.  The
    expected values are what those operations give, worked out here by hand."""
    z = pkg.deck.zstor
    # 3*10**-4: fl(3 * fl(1e-4)) is one ulp above the nearest double to 0.0003
    assert z("3.000000e-04") == 3.0 * 1.0e-4 and z("3.000000e-04") != 0.0003
    # 0.1 read as 1/10 is the nearest double; 0.15 as 15/100 likewise
    assert z("0.1") == 0.1 and z("-0.15") == -0.15
    # digits / 10**decimals, then the exponent (D and E both mark it)
    assert z("12.5D2") == 1250.0 and z("+7e2") == 700.0
    assert z("7E-1") == 7.0 * 0.1 and z("7E-1") != 0.7
    assert z("1.25e-3") == 125.0 / 100.0 * 1.0e-3
    # a second point or any other character ends the number; "1.0e" has no exponent
    assert z("1.5.3") == 1.5 and z("2x4") == 2.0 and z("1.0e") == 1.0
    # leading blanks, an empty field
    assert z("   4") == 4.0 and z("") == 0.0
    assert pkg.deck.zstoi("-12x") == -12 and pkg.deck.zstoi("+3") == 3


def test_deck_special_side_words(pkg):
    """MODEL WAVE-BC ... FUNC MATRIX and MODEL OPEN-BC ... FUNC TYPE1 as
    Source code/vf_iimdl.f:33-210 and vf_csetup.f:37-205 read them: the
    numbers a side needs, one model per side, and the boundary types the
    layout gives the side's faces (5: wave-maker, 7: radiation)."""
    base = ("GRID X 0 1 2 3 4 END\nGRID Y 0 1 2 3 END\nGRID Z 0 1 2 3 END\nTIME CONST 1e-3\n"
            "B.C. D VP SLIP\nB.C. D F FREE\nMATE W-LEVEL 1.5\n")
    wave = "MODEL WAVE-BC X- FUNC MATRIX\nMODEL WAVE-BC X- DEPTH 1.5\nMODEL WAVE-BC X- PERIOD 2.0\n"
    opn = "MODEL OPEN-BC X+ FUNC TYPE1\nMODEL OPEN-BC X+ DEPTH 1.5\nMODEL OPEN-BC X+ PERIOD 2.0\n"
    d = pkg.deck.parse(base + wave + opn)
    assert d.opens == {2: {"func": "TYPE1", "depth": 1.5, "period": 2.0}}
    assert d.waves[1]["depth"] == 1.5 and d.waves[1]["period"] == 2.0
    lay = pkg.layout.from_text(base + wave + opn)
    indb, indx = lay.static_arrays()["INDB"], lay.static_arrays()["INDX"]
    lx = indx[1, 1:-1, 1:-1].ravel() - 1          # faces of the x- side
    rx = indx[-1, 1:-1, 1:-1].ravel() - 1         # faces of the x+ side
    assert np.all(indb[2, lx] == 5) and np.all(indb[3, lx] == 5)
    assert np.all(indb[2, rx] == 7) and np.all(indb[3, rx] == 7)
    assert np.count_nonzero(indb[2] == 5) == 3 * 3 and np.count_nonzero(indb[2] == 7) == 3 * 3
    # a side takes one model; a radiation side needs its depth and period
    with pytest.raises(pkg.deck.DeckError):
        pkg.deck.parse(base + wave + "MODEL OPEN-BC X- FUNC TYPE1\n")
    with pytest.raises(pkg.deck.DeckError):
        pkg.deck.parse(base + "MODEL OPEN-BC Y- FUNC TYPE1\nMODEL OPEN-BC Y- DEPTH 1.0\n")
    with pytest.raises(pkg.deck.DeckError):
        pkg.deck.parse(base + "MODEL OPEN-BC Y- FUNC TYPE2\n")
    with pytest.raises(pkg.deck.DeckError):
        pkg.deck.parse(base + "MODEL WAVE-BC X- FUNC MATRIX\nMODEL WAVE-BC X- DEPTH 1.5\n")


def test_mtb_table_kinds(pkg):
    """<name>.mtb as VF_IM1INP reads it (Source code/vf_im1inp.f:61-175): a
    LEVEL-OFF table has no level after each phase (its level is 0); a
    LEVEL-ONLY table has no z levels and no profiles; the kind word is
    checked."""
    on = pkg.configs.solitary_mtb(0.2, 0.04, 2.0, 0.2, nz=5, nt=4)
    off = pkg.configs.solitary_mtb(0.2, 0.04, 2.0, 0.2, nz=5, nt=4, kind="LEVEL-OFF")
    a, b = pkg.mtb.parse(on), pkg.mtb.parse(off)
    assert (a.kind, b.kind) == ("LEVEL-ON", "LEVEL-OFF")
    assert np.array_equal(a.un, b.un) and np.array_equal(a.tt, b.tt) and np.array_equal(a.zz, b.zz)
    assert np.any(a.hh != 0.0) and np.all(b.hh == 0.0)
    only = pkg.mtb.parse(pkg.configs.solitary_mtb(0.2, 0.04, 2.0, 0.2, nz=5, nt=4, kind="LEVEL-ONLY"))
    assert only.kind == "LEVEL-ONLY" and len(only.zz) == 0 and np.array_equal(only.hh, a.hh)
    with pytest.raises(pkg.mtb.MtbError):
        pkg.mtb.parse(on.replace("LEVEL-ON", "LEVEL-X", 1))

"""The wave-maker's theory functions (MODEL WAVE-BC FUNC STOKES / STK-CND /
STREAM; Source code/vf_cwmak0.f, vf_cwmak1.f, vf_cwmak2.f, vf_wstk0.f,
vf_wcnd0.f, vf_wsfmb2.f, vf_cwzero.f), evaluated on the host by
cad_wave_theory_eval, against the
boundary values the reference wrote on the wave faces (vf_bwuwn.f:331-380).

The reference calls VF_CWMAK0 once at set-up (VF_CSETUP) and once in every
boundary pass (VF_CINIT's, then each step's predictor); VF_STK0 carries its
last secant step from call to call, so the replay makes the same calls in the
same order.  The columns' depth scale VWS is restated here from the recorded F
and flags (vf_bwuwn.f:300-330).  No GPU: the evaluation is host code.

Fixtures (tests/golden/make_golden.py): "stokes_flume" (fifth-order Stokes,
H = 2 cm, T = 1 s in 0.2 m) and "cnoidal_flume" (STK-CND with an Ursell
number of 29, so the third-order cnoidal wave), 40x4x20, ramped in over one
period, recorded at steps 1, 2, 100, 300, 600; "stream_flume" (the
fifth-order stream-function wave, H = 3 cm, T = 1.2 s, FUNC STREAM 5), the
same flume and steps.  The stream function's set-up is a least-squares
iteration in double precision from single-precision inputs; a one-literal
change there (1.01 for the reference's single-precision 1.01) already moves
the faces' bits.
"""
import numpy as np
import pytest

import fixtures as fx

CASES = ["stokes_flume", "cnoidal_flume", "oblique_basin", "stream_flume"]


def _kind(pkg, wv, grz0):
    """FUNC -> the wave evaluated (STK-CND by the Ursell number, vf_csetup.f:101-111;
    STREAM n -> n)."""
    f = pkg.capi.wave_func(wv)
    if f == 0:
        wus = grz0 * wv["height"] * wv["period"] * wv["period"] / wv["depth"] / wv["depth"]
        f = -2 if wus <= 25.0 else -1
    return f


def _phase(t0, tnow, t):
    """The single-precision phase of vf_bwuwn.f:270-276."""
    r = np.float32(t0 - tnow / t)
    r = np.float32(r - np.float32(int(r)))
    if r < np.float32(0.0):
        r = np.float32(r + np.float32(1.0))
    return float(r)


def _phase_ray(t0, tnow, t, rr, dl):
    """The phase of a column at distance rr along an oblique wave's ray."""
    r = np.float32(t0 - (tnow / t - rr / dl))
    r = np.float32(r - np.float32(int(r)))
    if r < np.float32(0.0):
        r = np.float32(r + np.float32(1.0))
    return float(r)


def _ramp(tnow, t, ampl):
    if ampl >= 1.0e-20:
        a = tnow / t / ampl
        if a < 1.0:
            return 0.5 * np.sin(3.141592653589794 * (a - 0.5)) + 0.5
    return 1.0


def _scale(st, rec, j, wvz, dep, wvlvl):
    """VWS of column j next to the x- side (cells I = 2)."""
    nf, ff, zz = rec["NF"], rec["FF"], st["ZZ"]
    val = 0.0
    for k in range(2, nf.shape[2]):
        nw = int(nf[1, j - 1, k - 1])
        dz = zz[1, k - 1]
        if nw == -1:
            val = val + dz * 1.0
        elif nw == 0:
            val = val + dz * ff[1, j - 1, k - 1]
        elif nw != 8:
            val = val + dz * ff[1, j - 1, k - 1]
            break
    val = val - (wvlvl - zz[0, 1])
    return (wvz + dep) / (val + dep)


@pytest.mark.parametrize("case", CASES)
def test_theory_wave_faces_match_reference(pkg, case):
    """BCU and BCW of every wave face after the predictor's boundary pass,
    bit for bit at every recorded step."""
    st = fx.static(case)
    d = pkg.deck.parse(str(st["DECK"]))
    wv = d.waves[1]
    grz0 = fx.deck_params(st["DECK"])["grz0"]
    kind = _kind(pkg, wv, grz0)
    assert kind == {"cnoidal_flume": -1, "stream_flume": 5}.get(case, -2)
    dep, t, wvlvl = wv["depth"], wv["period"], d.wvlvl
    zz, indx, xx, yy = st["ZZ"], st["INDX"], st["XX"], st["YY"]
    # oblique incidence (vf_bwuwn.f:297-325): the wave's direction, and each
    # column's phase shifted along the ray by the set-up wave length
    isa = abs(wv["angle"]) > 1.0e-20
    sa, sb, sc = 1.0, 0.0, 0.0
    if isa:
        ang = wv["angle"] * 3.141592653589794 / 180.0
        sb, sa = np.sin(ang), np.cos(ang)
        sc = -(sa * wv["ax"] + sb * wv["ay"])
    numk = zz.shape[1]
    for s in fx.steps(case):
        r = fx.step(case, s)
        if "v1_out" not in r:
            continue
        w = pkg.capi.TheoryWave(kind, dep, wv["height"], t)
        _, _, _, t0 = w.call(origin=True)            # VF_CSETUP
        dl = w.length
        if kind < 0:
            for _ in range(s):                       # VF_CINIT, steps 1 .. s-1 (VF_STK0's state)
                w.call()
        tnow = float(r["v1_in"]["TNOW"][0])
        wvt = _phase(t0, tnow, t)
        aw = _ramp(tnow, t, wv["ampl"])
        keep = w.state.value
        wvz, _, _, _ = w.call(wvt)
        w.state.value = keep                          # one VF_CWMAK0 for the step
        bcu, bcw = r["v1_out"]["BCU"], r["v1_out"]["BCW"]
        checked = 0
        bcv = r["v1_out"]["BCV"]
        for j in range(2, indx.shape[1]):
            wq, zq = wvt, wvz
            if isa:
                x0, y0 = xx[0, 1], (yy[0, j - 1] + yy[0, j]) * 0.5
                rr = (sa * x0 + sb * y0 + sc) / np.sqrt(sa * sa + sb * sb)
                wq = _phase_ray(t0, tnow, t, rr, dl)
                keep = w.state.value
                zq = w.call(wq)[0]
                w.state.value = keep
            vws = _scale(st, r["v1_in"], j, zq, dep, wvlvl)
            ks = [k for k in range(2, numk) if indx[1, j - 1, k - 1] >= 1]
            zc = [vws * ((zz[0, k - 1] + zz[0, k]) * 0.5 - wvlvl + dep) - dep for k in ks]
            keep = w.state.value
            _, un, ut, _ = w.call(wq, zc)
            w.state.value = keep
            for k, a, b in zip(ks, un, ut):
                ll = int(indx[1, j - 1, k - 1]) - 1
                assert a * vws * aw * sa == bcu[ll], (s, j, k)
                assert a * vws * aw * sb == bcv[ll], (s, j, k)
                assert b * vws * aw == bcw[ll], (s, j, k)
                checked += 1
        assert checked > 0


def test_theory_wave_is_a_wave(pkg):
    """What the functions must be, independent of the reference: the level
    crosses zero at the phase origin, its crest-to-trough range over a period
    is the height H (within 2%), and the horizontal velocity under the crest
    is positive and grows towards the surface."""
    for func in ("STOKES", "CNOIDAL"):
        d, h, t = (0.2, 0.02, 1.0) if func == "STOKES" else (0.2, 0.03, 2.0)
        w = pkg.capi.TheoryWave(func, d, h, t)
        _, _, _, t0 = w.call(origin=True)
        lv0 = w.call(t0)[0]
        assert abs(lv0) < 1e-7
        ph = np.linspace(0.0, 1.0, 201)[:-1]
        lv = np.array([w.call(p)[0] for p in ph])
        assert abs((lv.max() - lv.min()) - h) < 0.02 * h
        crest = ph[int(np.argmax(lv))]
        _, un, _, _ = w.call(crest, [-0.15, -0.1, -0.05, 0.0])
        assert np.all(un > 0) and np.all(np.diff(un) > 0)


def test_theory_eval_refuses_bad_arguments(pkg):
    """cad_wave_theory_eval refuses a function other than Stokes (-2) or
    cnoidal (-1), a missing state, and a depth, height or period below the
    set-up checks' 1e-20 (vf_csetup.f:86-94)."""
    import ctypes as C
    lib = pkg.capi.lib()
    st = C.c_float(0.0)
    lv = C.c_double()
    for func, d, h, t in ((0, 0.2, 0.02, 1.0), (-3, 0.2, 0.02, 1.0), (-2, 0.0, 0.02, 1.0),
                          (-2, 0.2, 0.0, 1.0), (-1, 0.2, 0.02, 0.0)):
        assert lib.cad_wave_theory_eval(func, d, h, t, C.byref(st), 0.5, 0, None, None, None,
                                        C.byref(lv), None, None) == -1
    assert lib.cad_wave_theory_eval(-2, 0.2, 0.02, 1.0, None, 0.5, 0, None, None, None,
                                    C.byref(lv), None, None) == -1
    assert lib.cad_wave_theory_eval(-2, 0.2, 0.02, 1.0, C.byref(st), 0.5, 0, None, None, None,
                                    C.byref(lv), None, None) == 0

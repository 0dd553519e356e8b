"""The benchmark tanks at their full sizes, against the reference's own runs.

The reference (its Fortran, one rank) ran the 512x32x96 solitary run-up (40
steps) and the 512x256x128 buildings tank (3 steps) restarted from the same
initial state (tests/golden/make_golden_full.py).  Here the GPU runs the same
tanks from the deck and the wave, through the C ABI, in the plain FP64 form,
and must keep, at every step: ITR equal to the reference's, FSUM within 1e-12,
||div|| within 1e-6; at the last step U, V, W, P, F on the recorded J planes
within 1e-8 and every cell flag equal.  These are the sizes where the sweep
tiles, the mailbox chain and the partial sums over millions of cells run as in
the benchmark.

The mixed form then runs the buildings tank for 200 steps beside the plain
form (FP64, 1e-10): its true relative residual is at most 1e-8 at every step,
and after 200 steps its fields are within 1e-6 of the plain form's with every
flag equal.
"""
import ast

import numpy as np
import pytest

import fixtures as fx
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _load(name):
    with np.load(fx.HERE / f"{name}_full_trajectory.npz") as z:
        traj = {k: z[k] for k in z.files}
    with np.load(fx.HERE / f"{name}_full_final.npz") as z:
        fin = {k: z[k] for k in z.files}
    return traj, fin


def _case(pkg, traj, name, steps=None):
    kw = ast.literal_eval(str(traj["KWARGS"]))
    n = int(traj["NSTEPS"]) if steps is None else steps
    return getattr(pkg.configs, str(traj["FACTORY"]))(steps=n, name=name, **kw)


def _context(pkg, case, precision=0):
    lay = pkg.layout.from_text(case.deck)
    state = lay.raw_state()
    case.ic(lay.XX, lay.ZZ, lay.NF0, state)
    ni, nj, nk = lay.shape
    c = pkg.capi.Context(ni, nj, nk, lay.numb)
    p = lay.params()
    c.set_params(**{k: v for k, v in p.items() if k in c.params})
    c.set_params(precision=precision)
    c.upload_static(lay.static_arrays())
    c.upload_state(state)
    c.init_state(lay.deck.wvlvl, lay.deck.uini, 0)
    return c


def _vel_slices_err(out, fin, jj):
    vs = max(np.abs(fin[k]).max() for k in ("UU", "VV", "WW"))
    return max(np.abs(out[k][:, jj, :] - fin[k]).max() for k in ("UU", "VV", "WW")) / vs


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", ["runup_b", "buildings_c"])
def test_full_size_plain_form_matches_reference(pkg, name):
    if not (fx.HERE / f"{name}_full_trajectory.npz").exists():
        pytest.skip("no full-size record")
    traj, fin = _load(name)
    case = _case(pkg, traj, name)
    nsteps = int(traj["NSTEPS"])
    jj = [int(j) - 1 for j in fin["JPLANES"]]
    print(f"{name}: {case.ni}x{case.nj}x{case.nk}, {nsteps} steps", flush=True)
    with _context(pkg, case) as c:
        for s in range(1, nsteps + 1):
            info = c.step()
            dv = c.cdiv00()
            print(f"{name} step {s} ITR={info['icgitr']} ref {int(traj['ITR'][s])}", flush=True)
            assert info["icgitr"] == int(traj["ITR"][s]), (name, s)
            assert abs(info["fsum"] - traj["FSUM"][s]) <= 1e-12 * traj["FSUM"][s], (name, s)
            assert abs(dv - traj["DV"][s]) <= 1e-6 * traj["DV"][s], (name, s, dv, traj["DV"][s])
        out = c.download_state()
    assert _vel_slices_err(out, fin, jj) <= 1e-8
    for k in ("PP", "FF"):
        e = rel_err(out[k][:, jj, :], fin[k])
        assert e <= 1e-8, (k, e)
    bad = np.count_nonzero(out["NF"].astype(np.int8) != fin["NF"])
    assert bad == 0, f"{bad} flags differ"
    assert info["nums"] == int(fin["NUMS"][0])


@pytest.mark.timeout(900)
def test_full_size_mixed_form_200_steps(pkg):
    name = "buildings_c"
    if not (fx.HERE / f"{name}_full_trajectory.npz").exists():
        pytest.skip("no full-size record")
    traj, _ = _load(name)
    steps = 200
    case = _case(pkg, traj, name, steps)
    print("buildings_c plain form, 200 steps", flush=True)
    with _context(pkg, case) as c:
        for s in range(steps):
            c.step()
        plain = c.download_state()
    print("buildings_c mixed form, 200 steps", flush=True)
    worst = 0.0
    with _context(pkg, case, precision=pkg.capi.PRECISION_MIXED) as c:
        c.set_params(cgepsr=1.0e-8)
        for s in range(steps):
            info = c.step()
            worst = max(worst, info["cgxnrm"])
            assert info["cgxnrm"] <= 1e-8, (s, info["cgxnrm"])
        mixed = c.download_state()
    print(f"worst true residual {worst:.3e}", flush=True)
    vs = max(np.abs(plain[k]).max() for k in ("UU", "VV", "WW"))
    ev = max(np.abs(mixed[k] - plain[k]).max() for k in ("UU", "VV", "WW")) / vs
    assert ev <= 1e-6, ev
    for k in ("PP", "FF"):
        assert rel_err(mixed[k], plain[k]) <= 1e-6, k
    assert np.array_equal(mixed["NF"], plain["NF"])

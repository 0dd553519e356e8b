"""The link-level drop-in of the pressure solve on the GPU.

* vf_vpsol_ / vf_m1bcgs_ called through ctypes and from a Fortran program
  (tests/fortran/vpsol_driver.f90) on the Poisson systems the reference
  recorded: ITR equal, PT within 1e-10, norms within 1e-12.
* The reference itself, relinked with this library in place of its
  vf_vpsol.o + vf_m1bcgs.o or of vf_m1bcgs.o alone (oracle/build_ref.sh; no
  reference file changed): its own time loop, input reader and output writer
  run around the GPU solve.  Its list-file ITR equals the unmodified
  reference's at every step and its result file agrees with the reference's.
"""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import fixtures as fx
from helpers import assert_close, assert_same_int, assert_vel_close
from test_dropin_abi import DRIVER, run_driver, write_system

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
REF_DIR = ROOT / "oracle" / "_ref"
CASES = ["tiny", "beach_small", "corner3d"]
TOL = 1.0e-10


def _solver_params(case):
    st = fx.static(case)
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    return {k: p[k] for k in ("cgpara", "cgepsa", "cgepsr", "cgdiv", "icgtyp", "icgmax")}


def _check(case, s, pt, bn, xn, it):
    ref = fx.step(case, s)["ps_out"]
    assert it == int(ref["ICGITR"][0]), (case, s, it)
    assert_close(f"{case} step {s} PT", pt, ref["PT"], TOL)
    assert abs(bn - ref["CGBNRM"][0]) <= 1e-12 * ref["CGBNRM"][0]
    assert abs(xn - ref["CGXNRM"][0]) <= 1e-6 * ref["CGXNRM"][0] + 1e-30


@pytest.mark.parametrize("case", CASES)
def test_vpsol_entry_matches_reference(pkg, case):
    p = _solver_params(case)
    for s in fx.steps(case):
        ps = fx.step(case, s)["ps_in"]
        print(f"vf_vpsol_ {case} step {s}", flush=True)
        pt, bn, xn, it = pkg.capi.vpsol(ps, ps["INDC"], **p)
        _check(case, s, pt, bn, xn, it)
        # inactive cells keep VF_VPSOL's PT = 0
        assert np.all(pt[ps["INDC"] == -1] == 0.0)


def test_m1bcgs_on_arrays_larger_than_the_grid(pkg):
    """MI,MJ,MK > NI,NJ,NK: the solve runs on the (NI,NJ,NK) corner, equal bit
    for bit to the same system in arrays of its own size; the rest of the
    arrays is not read and X0 there is left as it was."""
    case = "corner3d"
    p = _solver_params(case)
    ps = fx.step(case, fx.steps(case)[0])["ps_in"]
    shape = ps["AD"].shape
    big = (shape[0] + 3, shape[1] + 2, shape[2] + 1)
    sl = tuple(slice(0, n) for n in shape)
    padded = {}
    for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB"):
        a = np.full(big, 1.0e300, order="F")
        a[sl] = ps[k]
        padded[k] = a
    ic = np.full(big, 7, dtype=np.int32, order="F")
    ic[sl] = ps["INDC"]
    x0 = np.full(big, -5.0, order="F")
    x0[sl] = 0.0
    print("vf_m1bcgs_ padded corner3d", flush=True)
    xp, bnp, xnp, itp = pkg.capi.m1bcgs(padded, ic, x0=x0, loop=shape, **p)
    x, bn, xn, it = pkg.capi.m1bcgs(ps, ps["INDC"], **p)
    assert itp == it and bnp == bn and xnp == xn
    assert np.array_equal(xp[sl], x)
    out = np.ones(big, bool)
    out[sl] = False
    assert np.all(xp[out] == -5.0)


@pytest.mark.parametrize("mode", ["vpsol", "m1bcgs"])
def test_fortran_caller_on_recorded_systems(pkg, tmp_path, mode):
    """The Fortran caller solves the reference's recorded systems through
    VF_VPSOL (COMMON blocks) and VF_M1BCGS (full argument list)."""
    if not DRIVER.exists():
        pytest.skip("Fortran driver not built (flang absent at build time)")
    for case in CASES:
        p = _solver_params(case)
        s = fx.steps(case)[-1]
        ps = fx.step(case, s)["ps_in"]
        d = tmp_path / f"{case}_{mode}"
        d.mkdir()
        write_system(d, ps, ps["INDC"], ps["AD"].shape, p)
        print(f"Fortran {mode} {case} step {s}", flush=True)
        r = run_driver(d, mode)
        assert r.returncode == 0, r.stdout + r.stderr
        w = r.stdout.split()
        pt = np.fromfile(d / "PT.bin", dtype=np.float64).reshape(ps["AD"].shape, order="F")
        _check(case, s, pt, float(w[2]), float(w[3]), int(w[1]))


# a step's list line and the ITR on the line after it (the ">> FILE-... :
# STEP=" lines of the output files are not step lines)
_STEP_RE = re.compile(r"^ STEP=\s*(\d+)\s*:[^\n]*\n[^\n]*ITR\s*=\s*(\S+)", re.M)


def _run_reference(exe, name, deck, tmp_path, timeout=600):
    d = tmp_path / name
    d.mkdir()
    (d / f"{name}.in").write_text(deck)
    (d / "data.env").write_text(f"1\n{name} 1 0\n")
    print(f"reference relinked: {exe.name} on {name}", flush=True)
    r = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0 and "NORMAL END" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    itr = {int(a): int(b) for a, b in _STEP_RE.findall((d / f"{name}.list").read_text())}
    return d, itr


def test_reference_time_loop_with_gpu_vpsol(pkg, tmp_path):
    """The reference with vf_vpsol_ from this library runs the 64x16x48
    dam break for 100 steps: ITR equal to the unmodified reference's at every
    step, and the step-100 state within 1e-8, flags equal."""
    exe = REF_DIR / "cadmas_ref_vpsol_hip.x"
    if not exe.exists():
        pytest.skip("relinked reference not built (reference sources absent at build time)")
    deck = str(fx.static("dam_a")["DECK"]).replace("TIME END 500", "TIME END 100")
    deck += "FILE RSL STEP 100 100 1\n"
    d, itr = _run_reference(exe, "dam_a", deck, tmp_path)
    ref_itr = fx.trajectory("dam_a")["ITR"]
    assert len(itr) == 101
    assert [itr[s] for s in range(1, 101)] == [int(v) for v in ref_itr[1:101]]
    recs = list(pkg.rsl.read(d / "dam_a.rsl"))
    a = recs[-1]
    assert a["NNOW"] == 100
    ref = fx.step("dam_a", 100)["f1_out"]
    assert_vel_close("step 100 velocity", a, ref, 1e-8)
    for k in ("PP", "FF"):
        assert_close(f"step 100 {k}", a[k], ref[k], 1e-8)
    assert_same_int("step 100 NF", a["NF"], ref["NF"])


def test_reference_time_loop_with_gpu_m1bcgs(pkg, tmp_path):
    """The reference with vf_m1bcgs_ from this library (its own VF_VPSOL
    calling it) runs the tiny dam break: the result file's records agree with
    the unmodified reference's (fields to 1e-10, flags exactly, same ITR)."""
    exe = REF_DIR / "cadmas_ref_m1bcgs_hip.x"
    if not exe.exists():
        pytest.skip("relinked reference not built (reference sources absent at build time)")
    deck = str(fx.static("tiny")["DECK"])
    d, itr = _run_reference(exe, "tiny", deck, tmp_path)
    ref_itr = fx.trajectory("tiny")["ITR"]
    assert [itr[s] for s in range(1, 6)] == [int(v) for v in ref_itr[1:6]]
    with np.load(fx.HERE / "tiny_rsl.npz") as z:
        (tmp_path / "ref.rsl").write_bytes(z["RSL"].tobytes())
    ref = list(pkg.rsl.read(tmp_path / "ref.rsl"))
    mine = list(pkg.rsl.read(d / "tiny.rsl"))
    assert len(mine) == len(ref) == 7
    for a, b in zip(mine[1:], ref[1:]):
        s = b["NNOW"]
        assert a["NNOW"] == s and a["ICGITR"] == b["ICGITR"]
        assert_vel_close(f"step {s} velocity", a, b, 1e-10)
        for k in ("PP", "FF"):
            assert_close(f"step {s} {k}", a[k], b[k], 1e-10)
        assert_same_int(f"step {s} NF", a["NF"], b["NF"])

"""The link-level drop-in of the pressure solve (vf_m1bcgs_, vf_vpsol_) on the
CPU: the COMMON blocks' layout agrees with the reference's Fortran, a Fortran
caller shares its blocks and its error exit with the library, and failures of
the library itself reach the caller as codes the reference's VF_VPSOL stops on
(vf_vpsol.f:70-75).  Every case here fails its argument checks before the
library touches a GPU."""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "fortran" / "bin" / "vpsol_driver"
DRIVER_OBJ = ROOT / "tests" / "fortran" / "bin" / "vpsol_driver.o"
NAMES = ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK", "BB")


def write_system(d, mat, indc, loop, params):
    """The driver's input: sys.txt and raw Fortran-order arrays."""
    d = Path(d)
    shape = mat["AD"].shape
    p = params
    (d / "sys.txt").write_text(" ".join(str(v) for v in (*loop, *shape)) +
                               f" {p['cgpara']!r} {p['cgepsa']!r} {p['cgepsr']!r} {p['cgdiv']!r}"
                               f" {int(p['icgtyp'])} {int(p['icgmax'])}\n")
    for k in NAMES:
        np.asarray(mat[k], dtype=np.float64).ravel(order="F").tofile(d / f"{k}.bin")
    np.asarray(indc, dtype=np.int32).ravel(order="F").tofile(d / "INDC.bin")


def run_driver(d, mode):
    return subprocess.run([str(DRIVER), str(d), mode], capture_output=True, text=True, timeout=300)


PARAMS = dict(cgpara=0.95, cgepsa=1.0e-15, cgepsr=1.0e-12, cgdiv=1.0e30, icgtyp=1, icgmax=7)


def _zero_system(shape):
    return {k: np.zeros(shape, order="F") for k in NAMES}, np.zeros(shape, np.int32, order="F")


def test_common_block_sizes_match_the_fortran_declaration(pkg):
    """sizeof of the C structs equals the size flang gives the COMMON blocks
    declared as the reference declares them (the driver's declarations)."""
    if not DRIVER_OBJ.exists():
        pytest.skip("Fortran driver not built (flang absent at build time)")
    out = subprocess.run(["nm", "-S", str(DRIVER_OBJ)], capture_output=True, text=True).stdout
    sizes = {}
    for ln in out.splitlines():
        w = ln.split()
        if len(w) == 4 and w[3] in ("vf_acompr_", "vf_acompi_", "vf_anumbi_"):
            sizes[w[3]] = int(w[1], 16)
    C = pkg.capi
    assert sizes == {"vf_acompr_": ctypes.sizeof(C.AcomprBlock), "vf_acompi_": ctypes.sizeof(C.AcompiBlock),
                     "vf_anumbi_": ctypes.sizeof(C.AnumbiBlock)}, sizes
    assert ctypes.sizeof(C.AcomprBlock) == 25 * 8
    assert ctypes.sizeof(C.AcompiBlock) == 18 * 4
    assert ctypes.sizeof(C.AnumbiBlock) == 10 * 4


def test_m1bcgs_library_failure_is_past_icgmax(pkg):
    """A loop range that is not a grid (NI = 3) is a failure of the library:
    ITER = ICGMAX + 1 (VF_VPSOL's NOT CONVERGED), the cause in cad_last_error."""
    mat, indc = _zero_system((3, 5, 5))
    x, bn, xn, it = pkg.capi.m1bcgs(mat, indc, icgmax=7)
    assert it == 8
    assert "ghost layer" in pkg.capi.lib().cad_last_error().decode()


def test_m1bcgs_rejects_a_loop_range_beyond_the_arrays(pkg):
    """NI > MI (the loops would run past the arrays) is refused, not run."""
    mat, indc = _zero_system((6, 6, 6))
    x, bn, xn, it = pkg.capi.m1bcgs(mat, indc, icgmax=11, loop=(8, 6, 6))
    assert it == 12
    assert np.all(x == 0.0)


def test_vpsol_failure_reaches_common_block(pkg):
    """Through the library's own COMMON blocks (no Fortran program): VF_VPSOL
    writes ICGITR = ICGMAX + 1 back into /VF_ACOMPI/ and PT is zeroed."""
    mat, indc = _zero_system((3, 5, 5))
    pt, bn, xn, it = pkg.capi.vpsol(mat, indc, icgmax=7)
    assert it == 8
    assert np.all(pt == 0.0)


def test_fortran_caller_shares_common_blocks_and_error_exit(pkg, tmp_path):
    """The Fortran caller's COMMON blocks are the ones the library reads (its
    NUMI = 3 is what the library refuses) and the caller's VF_A2ERR is what
    the library calls: the run stops with 'VF_VPSOL NOT CONVERGED.', as the
    reference does after a failed solve."""
    if not DRIVER.exists():
        pytest.skip("Fortran driver not built (flang absent at build time)")
    mat, indc = _zero_system((3, 5, 5))
    write_system(tmp_path, mat, indc, (3, 5, 5), PARAMS)
    r = run_driver(tmp_path, "vpsol")
    assert r.returncode == 3, r.stdout + r.stderr
    assert "VF_A2ERR: VF_VPSOL NOT CONVERGED." in r.stdout
    assert "NI,NJ,NK = 3,5,5" in r.stderr
    r = run_driver(tmp_path, "m1bcgs")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split()[1] == "8"

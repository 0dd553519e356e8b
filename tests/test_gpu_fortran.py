"""The reference-side binding: a Fortran program (tests/fortran/m1bcgs_driver.f90)
that uses the ISO_C_BINDING module include/cadmas_hip.f90 and calls the
VF_M1BCGS drop-in CADMAS_HIP_M1BCGS the way VF_VPSOL calls VF_M1BCGS
(vf_vpsol.f:64-67).  The driver builds a 7-point system with an obstacle block,
solves it on the GPU and checks the true relative residual (< 1e-8) in FP64."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent / "fortran" / "bin" / "m1bcgs_driver"


def test_fortran_caller_of_m1bcgs_dropin():
    if not DRIVER.exists():
        pytest.skip("Fortran driver not built (flang absent at build time)")
    r = subprocess.run([str(DRIVER)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout

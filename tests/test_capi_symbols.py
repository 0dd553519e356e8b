"""The C-ABI library loads without a GPU and exports every function that
include/cadmas_hip.h declares (no compute call is made here)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "cadmas_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cad[a-z0-9_]*|vf_[a-z0-9]+_)\s*\(", text)))


def test_header_declares_the_boundary():
    names = _declared()
    for n in ("cad_create", "cad_step", "cad_v1cal", "cad_f1cal", "cad_vpsol", "vf_m1bcgs_",
              "vf_vpsol_", "cad_init_state"):
        assert n in names


def test_library_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(str(pkg.capi.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_covers_the_header(pkg):
    assert set(_declared()) <= set(pkg.capi.EXPORTS)


def test_fortran_module_binds_exported_symbols(pkg):
    """Every bind(c, name='...') of the Fortran interface module names a symbol
    the library exports, and the module compiles where flang is present."""
    import shutil
    import subprocess
    import tempfile
    src = ROOT / "include" / "cadmas_hip.f90"
    names = re.findall(r"bind\(c,\s*name='([a-z0-9_]+)'\)", src.read_text())
    assert "cad_step" in names and "cad_create" in names
    lib = ctypes.CDLL(str(pkg.capi.LIB_PATH))
    assert [n for n in names if not hasattr(lib, n)] == []
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if Path(flang).exists():
        with tempfile.TemporaryDirectory() as d:
            r = subprocess.run([flang, "-fsyntax-only", "-module-dir", d, str(src)],
                               capture_output=True, text=True)
            assert r.returncode == 0, r.stderr


def test_params_struct_layout_agrees(pkg):
    """cad_params has one layout in C (the header), in the ctypes mirror and
    in the Fortran bind(c) type: same size and the same offsets of the fields
    appended last (IDROP, WBUB, TNOW)."""
    import shutil
    import subprocess
    import tempfile
    P = pkg.capi.Params
    fields = ("precision", "idrop", "wbub", "tnow")
    want = [ctypes.sizeof(P)] + [getattr(P, f).offset for f in fields]
    with tempfile.TemporaryDirectory() as d:
        csrc = Path(d) / "p.c"
        csrc.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cadmas_hip.h"\n'
                        "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(cad_params),"
                        + ",".join(f"offsetof(cad_params,{f})" for f in fields) + ");return 0;}\n")
        exe = Path(d) / "p"
        r = subprocess.run(["gcc", "-I", str(ROOT / "include"), str(csrc), "-o", str(exe)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
        assert got == want, (got, want)
        flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
        if Path(flang).exists():
            fsrc = Path(d) / "p.f90"
            fsrc.write_text("program p\n  use cadmas_hip\n  type(cad_params) :: q\n"
                            "  print '(i0)', c_sizeof(q)\nend program p\n")
            fexe = Path(d) / "pf"
            r = subprocess.run([flang, "-module-dir", d, str(ROOT / "include" / "cadmas_hip.f90"),
                                str(fsrc), "-o", str(fexe)], capture_output=True, text=True, cwd=d)
            assert r.returncode == 0, r.stderr
            out = subprocess.run([str(fexe)], capture_output=True, text=True).stdout.split()
            assert int(out[0]) == want[0]

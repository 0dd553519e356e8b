"""Build of the gfx950 library libcadmas_hip.so from csrc/ with hipcc.

Every source is compiled for gfx950 only, with
  -ffp-contract=off  a*b+c stays two roundings (the plain FP64 form reproduces
                     the reference's arithmetic bit for bit per cell), and
  -Xclang -target-feature -Xclang -packed-fp32-ops  no packed FP32 instructions.
The library is written next to this file (in-tree) so that it travels with the
source tree.
"""
import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libcadmas_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
         "-Wno-unused-value", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    return sorted(CSRC.glob("*.hip"))


def build(verbose=False, jobs=8):
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = objdir / (src.stem + ".o")
        objs.append(obj)
        cmd = [HIPCC, f"--offload-arch={ARCH}", *FLAGS, "-c", str(src), "-o", str(obj)]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
        if len(procs) >= jobs:
            _wait(procs.pop(0), verbose)
    for p in procs:
        _wait(p, verbose)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link of libcadmas_hip.so failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def _wait(item, verbose):
    src, p = item
    out, _ = p.communicate()
    lines = [ln for ln in (out or "").splitlines() if "packed-fp32-ops" not in ln]
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n" + "\n".join(lines))
    if verbose and lines:
        print("\n".join(lines))


if __name__ == "__main__":
    print(build(verbose=True))

"""Golden records of the reference on the full-size benchmark tanks.

The solitary-wave tanks cannot be written as a deck alone (the initial wave is
not deck-expressible, SURVEY's note on vf_ir1inp.f:184-205), so the reference
restarts from a result file: the tank's deck (configs.py) is laid out, the
wave is applied, the C restatement's VF_CINIT takes the state to the
reference's initial state and rsl.Writer writes it as <name>.res; the
reference (oracle/_ref/cadmas_ref.x, its own Fortran) then reads it with
FILE RES 0 and runs N steps.  Restarted this way the reference and the GPU
start from the same state.

Kept, small enough to commit (whole fields at these sizes are not):

  <name>_full_trajectory.npz  per step 1..N, full precision: ITR, ||VD||,
                              FSUM, FCUT, !B!, !R!, DT, TNOW; the deck and
                              the case's arguments
  <name>_full_final.npz       step N: U, V, W, P, F on a few J planes, NF of
                              the whole grid (int8), FSUM/FCUT/ITR; and the
                              sha256 of every set-up array the reference built
                              (VF_CSETUP and the rest), which the layout's
                              arrays must equal bit for bit

Usage (build container only: needs /root/reference and oracle/_ref):
    python tests/golden/make_golden_full.py runup_b buildings_c
"""
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(ROOT / "oracle"))
from dumpio import read_dump  # noqa: E402
from make_golden import REF_EXE, _load_pkg  # noqa: E402

# name -> (configs factory keyword arguments, steps, J planes kept)
CASES = {
    # config B: 512x32x96 solitary run-up, rubble block, 1:20 beach from 6 m
    "runup_b": ("solitary_runup", dict(eps_r=1.0e-10), 40, (2, 17, 33)),
    # config C: 512x256x128, 40 buildings, H/h = 0.15 on h = 0.8 m
    "buildings_c": ("buildings_tank", dict(eps_r=1.0e-10), 3, (2, 100, 180, 257)),
}

STATIC = ("XX", "YY", "ZZ", "ANU", "CD0", "GGV", "GGX", "GGY", "GGZ", "GLV", "GLX", "GLY", "GLZ",
          "INDX", "INDY", "INDZ", "INDB", "BCVI")


def sha(a):
    a = np.asarray(a)
    return hashlib.sha256(np.ascontiguousarray(a.ravel(order="F")).tobytes()).hexdigest()


def case_of(pkg, name, steps=None):
    fac, kw, n, _ = CASES[name]
    return getattr(pkg.configs, fac)(steps=steps or n, name=name, **kw)


def initial_state(pkg, case):
    """The layout of the deck with the wave applied and VF_CINIT run by the C
    restatement: (layout, state arrays)."""
    import pyoracle as po
    lay = pkg.layout.from_text(case.deck)
    state = lay.raw_state()
    case.ic(lay.XX, lay.ZZ, lay.NF0, state)
    arr = dict(lay.static_arrays())
    arr.update(state)
    F = po.Fields(arr, lay.params())
    rc = F.init_state(lay.deck.wvlvl, lay.deck.uini, 0)
    if rc != 0:
        raise RuntimeError(f"oracle VF_CINIT failed ({rc})")
    keys = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF")
    return lay, {k: np.array(F[k]) for k in keys}


def run(name):
    pkg = _load_pkg()
    fac, kw, nsteps, jplanes = CASES[name]
    case = case_of(pkg, name)
    t0 = time.time()
    lay, st = initial_state(pkg, case)
    print(f"{name}: layout + VF_CINIT {time.time() - t0:.1f}s", flush=True)
    body = "\n".join(ln for ln in case.deck.splitlines() if not ln.startswith(("TIME END", "FILE ")))
    info = {"fsum": 0.0, "fcut": 0.0, "cgbnrm": 0.0, "cgxnrm": 0.0, "icgitr": 0}
    with tempfile.TemporaryDirectory(prefix=f"golden_{name}_", dir="/root") as td:
        td = Path(td)
        (td / "dump").mkdir()
        with pkg.rsl.Writer(td / f"{name}.res", lay.shape, lay.numb, lay.static_arrays()["GGV"],
                            idrop=lay.deck.idrop) as w:
            w.write_step(0, 0.0, lay.params()["dtnow"], info, st, {"TBUB": np.zeros(lay.shape[2])})
        (td / f"{name}.in").write_text(body + f"\nTIME END {nsteps} 1.0D9\nFILE RES 0\n")
        (td / "data.env").write_text(f"1\n{name} 1 0\n")
        env = dict(os.environ, CADMAS_DUMP_DIR=str(td / "dump"), CADMAS_DUMP_STEPS=str(nsteps),
                   CADMAS_DUMP_TAGS="static,f1_out", OMP_NUM_THREADS="1")
        t0 = time.time()
        r = subprocess.run([str(REF_EXE)], cwd=td, env=env, capture_output=True, text=True)
        if r.returncode != 0 or "NORMAL END" not in r.stdout:
            raise RuntimeError(f"reference run {name} failed rc={r.returncode}\n{r.stdout[-3000:]}")
        print(f"{name}: reference {nsteps} steps {time.time() - t0:.1f}s", flush=True)
        traj = {k: [] for k in ("STEP", "ITR", "DV", "FSUM", "FCUT", "CGBNRM", "CGXNRM", "DT", "TNOW")}
        for f in sorted((td / "dump").glob("s*_dv.bin")):
            d = read_dump(f)
            traj["STEP"].append(int(d["NNOW"][0]))
            traj["ITR"].append(int(d["ICGITR"][0]))
            for k, dk in (("DV", "DV"), ("FSUM", "FSUM"), ("FCUT", "FCUT"), ("CGBNRM", "CGBNRM"),
                          ("CGXNRM", "CGXNRM"), ("DT", "DTNOW"), ("TNOW", "TNOW")):
                traj[k].append(float(d[dk][0]))
        lines = (td / f"{name}.list").read_text().splitlines()
        list_lines = [a + "\n" + b for a, b in zip(lines, lines[1:]) if a.startswith(" STEP=")]
        np.savez_compressed(HERE / f"{name}_full_trajectory.npz",
                            **{k: np.array(v) for k, v in traj.items()}, LIST=np.array(list_lines),
                            DECK=np.array(case.deck), FACTORY=np.array(fac),
                            KWARGS=np.array(repr(kw)), NSTEPS=np.array(nsteps))
        static = read_dump(td / "dump" / "static.bin")
        fin = read_dump(td / "dump" / f"s{nsteps:05d}_f1_out.bin")
        out = {"JPLANES": np.array(jplanes, dtype=np.int32)}
        for k in ("UU", "VV", "WW", "PP", "FF"):
            out[k] = np.ascontiguousarray(fin[k][:, [j - 1 for j in jplanes], :])
        out["NF"] = fin["NF"].astype(np.int8)
        for k in ("FSUM", "FCUT", "CGBNRM", "CGXNRM", "ICGITR", "NUMS"):
            if k in fin:
                out[k] = fin[k]
        for k in STATIC:
            if k in static:
                out["SHA_" + k] = np.array(sha(static[k]))
        np.savez_compressed(HERE / f"{name}_full_final.npz", **out)
        print(f"{name}: ITR {traj['ITR']}", flush=True)


if __name__ == "__main__":
    if not REF_EXE.exists():
        subprocess.run([str(ROOT / "oracle" / "build_ref.sh")], check=True)
    for n in sys.argv[1:] or list(CASES):
        run(n)
    shutil.rmtree(HERE / "__pycache__", ignore_errors=True)

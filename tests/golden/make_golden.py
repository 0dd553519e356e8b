"""Generate golden fixtures from the reference CADMAS-SURF/3D solver.

Runs oracle/_ref/cadmas_ref.x (the reference compiled from its own Fortran
sources by oracle/build_ref.sh, with the phase recorder oracle/ref_dump.c) on
the synthetic decks of configs.py and stores what each phase read and wrote as
compressed .npz files next to this script:

  <case>_static.npz          set-up arrays (metrics, porosities, face indices,
                             boundary list) and solver settings
  <case>_sNNNNN.npz          for step NNNNN: v1_in/ps_in/ps_out/v1_out/f1_in/
                             f1_out records, keys prefixed "<tag>:"
  <case>_trajectory.npz      per-step list-file quantities for every step:
                             STEP, FSUM, FCUT, DV, BNORM, RNORM, ITR

Usage: python tests/golden/make_golden.py [case ...]
Only runs in the build container (needs /root/reference and flang); the
fixtures it writes are what the tests read everywhere else.
"""
import os
import re
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(HERE))
from dumpio import read_dump  # noqa: E402

REF_EXE = ROOT / "oracle" / "_ref" / "cadmas_ref.x"


def _load_pkg():
    import importlib.util
    name = "cadmas_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = ROOT / "multiscale-and-multiphysics-integrated-simulator-for-tsunami_amd"
    spec = importlib.util.spec_from_file_location(name, pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


# case name -> (case factory, steps whose phase records are kept)
def _cases():
    cfg = _load_pkg().configs
    return {
        "tiny": (lambda: cfg.dam_break(8, 3, 6, 0.01, 0.03, 0.04, steps=5, name="tiny"),
                 [1, 2, 3, 4, 5]),
        "beach_small": (lambda: cfg.beach_small(steps=40, name="beach_small"), [1, 2, 3, 20, 40]),
        "corner3d": (lambda: cfg.corner3d(steps=60, name="corner3d"), [1, 2, 10, 30, 60]),
        "dam_a": (lambda: cfg.dam_break(name="dam_a"), [1, 100, 500]),
        # TimerDoor bubble / droplet models (vf_fbubup.f, vf_fdropf.f)
        "tdoor": (lambda: cfg.tdoor_drops(name="tdoor"),
                  [1, 2, 100, 101, 145, 170, 171, 180, 195, 200, 210, 220, 230, 245, 300]),
        "tdoor_dam": (lambda: cfg.tdoor_dam(name="tdoor_dam"),
                      [1, 2, 26, 200, 201, 400, 533, 571, 600, 601, 700, 900]),
        # automatic time step (vf_cdtcal.f) and OPTION SUB-LOOP
        "auto_dam": (lambda: cfg.auto_dam(name="auto_dam"), [1, 2, 100, 250, 400]),
        "subloop": (lambda: cfg.subloop(name="subloop"), [1, 2, 20, 40]),
        # table wave-maker (MODEL WAVE-BC FUNC MATRIX, vf_bwuwn.f:257-480)
        "wave_flume": (lambda: cfg.wave_flume(name="wave_flume"), [1, 2, 100, 200, 300, 400]),
        # ... with a radiation boundary at the far end (MODEL OPEN-BC X+ FUNC TYPE1)
        "wave_open": (lambda: cfg.wave_flume(ni=40, steps=1000, open_end=True, name="wave_open"),
                      [1, 2, 250, 500, 750, 1000]),
        # ... driven by a LEVEL-OFF table (profiles not scaled by the level)
        "wave_off": (lambda: cfg.wave_flume(ni=40, steps=400, table="LEVEL-OFF", name="wave_off"),
                     [1, 2, 100, 200, 400]),
        # ... and by a LEVEL-ONLY table (a fixed water level, vf_bwffsf.f)
        "wave_only": (lambda: cfg.wave_flume(ni=40, steps=400, table="LEVEL-ONLY", name="wave_only"),
                      [1, 2, 100, 200, 400]),
        # two tables: MATRIX (LEVEL-ON) at X-, MATRIX2 (LEVEL-ONLY) at X+
        "wave_two": (lambda: cfg.wave_two(name="wave_two"), [1, 2, 100, 200, 400]),
        # Dupuit-Forchheimer drag (OPTION DRAG-DF, vf_vgene.f:111-146)
        "porous_df": (lambda: cfg.porous_df(name="porous_df"), [1, 2, 3, 30, 60]),
        # implicit damping zones (MODEL DAMP, vf_veuler.f:131-166)
        "wave_damp": (lambda: cfg.wave_damp(name="wave_damp"), [1, 2, 300, 450, 600]),
        "damp3d": (lambda: cfg.damp3d(name="damp3d"), [1, 2, 10, 30, 60]),
        # theory wave-makers (VF_CWMAK0/1/2): fifth-order Stokes, and STK-CND
        # whose Ursell number (29) picks the third-order cnoidal wave
        "stokes_flume": (lambda: cfg.theory_flume("STOKES", name="stokes_flume"),
                         [1, 2, 100, 300, 600]),
        "cnoidal_flume": (lambda: cfg.theory_flume("STK-CND", H=0.03, period=2.0,
                                                   name="cnoidal_flume"), [1, 2, 100, 300, 600]),
        # the fifth-order stream-function wave (FUNC STREAM 5, vf_wsfmb2.f)
        "stream_flume": (lambda: cfg.theory_flume("STREAM 5", H=0.03, period=1.2, name="stream_flume"),
                         [1, 2, 100, 300, 600]),
        # the Stokes wave at 25 degrees to the x- side's normal (MODEL WAVE-BC ANGLE)
        "oblique_basin": (lambda: cfg.theory_flume("STOKES", ni=30, nj=16, nk=16, h=0.16, period=0.8,
                                                   steps=300, angle=(25.0, 0.0, 0.1),
                                                   name="oblique_basin"), [1, 2, 100, 300]),
    }


# keep only these records for large cases (tag -> keys or None for all)
_F1_ONLY = {"f1_in": None, "f1_out": None}
_KEEP = {
    "dam_a": {1: {"v1_in": None, "f1_in": ("BCF",)}, 100: {"f1_out": None}, 500: {"f1_out": None}},
    "tdoor": {s: _F1_ONLY for s in (2, 100, 101, 145, 170, 171, 180, 195, 200, 210, 220, 230, 245, 300)},
    "tdoor_dam": {s: _F1_ONLY for s in (2, 26, 200, 201, 400, 533, 571, 600, 601, 700, 900)},
    "auto_dam": {s: _F1_ONLY for s in (100, 250, 400)},
    "wave_flume": {s: {"v1_in": None, "ps_in": ("US", "VS", "WS"), "v1_out": None, "f1_in": None,
                       "f1_out": None} for s in (100, 200, 300, 400)},
    "wave_open": {s: {"v1_in": None, "ps_in": ("US", "VS", "WS"), "v1_out": None, "f1_in": None,
                      "f1_out": None} for s in (250, 500, 750, 1000)},
    "wave_off": {s: {"v1_in": None, "ps_in": ("US", "VS", "WS"), "v1_out": None, "f1_in": None,
                     "f1_out": None} for s in (100, 200, 400)},
    "wave_only": {s: {"v1_in": None, "ps_in": None, "v1_out": None, "f1_in": None,
                      "f1_out": None} for s in (100, 200, 400)},
    "wave_two": {s: {"v1_in": None, "ps_in": None, "v1_out": None, "f1_in": None,
                     "f1_out": None} for s in (100, 200, 400)},
    "wave_damp": {s: {"v1_in": None, "ps_in": ("US", "VS", "WS"), "v1_out": None, "f1_in": None,
                      "f1_out": None} for s in (300, 450, 600)},
    "stokes_flume": {s: {"v1_in": None, "ps_in": ("US", "VS", "WS"), "v1_out": None, "f1_in": None,
                         "f1_out": None} for s in (100, 300, 600)},
    "cnoidal_flume": {s: {"v1_in": None, "ps_in": ("US", "VS", "WS"), "v1_out": None, "f1_in": None,
                          "f1_out": None} for s in (100, 300, 600)},
    "stream_flume": {s: {"v1_in": None, "ps_in": ("US", "VS", "WS"), "v1_out": None, "f1_in": None,
                         "f1_out": None} for s in (100, 300, 600)},
}


# cases whose reference run also writes its result file (FILE RSL) and its
# graphics file (FILE GRP, Source code/vf_iifile.f:148-238); the files' bytes
# are kept as <case>_rsl.npz / <case>_grp.npz
_RSL = {"tiny": "FILE RSL STEP 0 5 1\nFILE GRP STEP 0 5 1",
        "tdoor": "FILE RSL STEP 100 300 100\nFILE GRP STEP 100 300 100\nFILE GRP AREA 4 2 3 18 4 14\n"
                 + "\n".join("FILE TRN " + t for t in (
                     "STEP 0 300 10", "W-LEVEL 5 3", "W-LEVEL 11 2", "POINT U 5 3 2", "POINT V 3 2 2",
                     "POINT W 3 3 2", "POINT P 8 3 2", "POINT F 11 3 3", "MAX U 1 1 1 20 5 16",
                     "MIN P 1 1 1 20 5 16", "AV F 1 1 1 11 5 8", "INT P 1 1 1 11 5 8",
                     "AV W 1 1 1 20 5 16", "INT U 2 1 1 10 5 6", "MIN F 1 1 1 20 5 16",
                     "FORCE X- 1 1 1 20 5 16", "FORCE X+ 1 1 1 20 5 16", "FORCE Z- 1 1 1 20 5 16",
                     "FORCE Z+ 1 1 1 20 5 16", "FORCE Y- 1 1 1 20 5 16", "FORCE Y+ 1 1 1 20 5 16"))}
# result files kept from those runs (the .tran only where FILE TRN is set)
_RSL["wave_flume"] = "\n".join("FILE TRN " + t for t in (
    "STEP 0 400 20", "W-LEVEL ANS X-", "W-LEVEL 2 2", "W-LEVEL 30 3", "POINT U 1 2 5", "POINT U 31 2 8",
    "POINT P 10 2 5", "AV F 1 1 1 60 4 20"))
_OUT = {"tiny": ("rsl", "grp"), "tdoor": ("rsl", "grp", "tran"), "wave_flume": ("tran",)}

_STEP_RE = re.compile(
    r"STEP=\s*(\d+)\s*:\s*TIME=\s*(\S+)\s*:\s*DT\s*=\s*(\S+)\s*:\s*FSUM=\s*(\S+)\s*:\s*FCUT=\s*(\S+)"
    r"\s*:\s*!VD!=\s*(\S+)\s*\n\s*:\s*!B!\s*=\s*(\S+)\s*:\s*!R!\s*=\s*(\S+)\s*:\s*ITR\s*=\s*(\S+)")


def run_case(name, factory, keep_steps):
    case = factory()
    if name in _RSL:
        case.deck = case.deck + _RSL[name] + "\n"
    with tempfile.TemporaryDirectory(prefix=f"golden_{name}_") as td:
        td = Path(td)
        (td / "dump").mkdir()
        (td / f"{name}.in").write_text(case.deck)
        for ext, text in case.meta.get("files", {}).items():
            (td / f"{name}.{ext}").write_text(text)
        (td / "data.env").write_text(f"1\n{name} 1 0\n")
        env = dict(os.environ, CADMAS_DUMP_DIR=str(td / "dump"),
                   CADMAS_DUMP_STEPS=",".join(str(s) for s in keep_steps))
        r = subprocess.run([str(REF_EXE)], cwd=td, env=env, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference run {name} failed rc={r.returncode}\n{r.stdout[-2000:]}")
        listing = (td / f"{name}.list").read_text()
        lines = listing.splitlines()
        list_lines = [a + "\n" + b for a, b in zip(lines, lines[1:]) if a.startswith(" STEP=")]
        if name in _RSL:
            for ext in _OUT[name]:
                raw = (td / f"{name}.{ext}").read_bytes()
                np.savez_compressed(HERE / f"{name}_{ext}.npz",
                                    **{ext.upper(): np.frombuffer(raw, dtype=np.uint8)})
        traj = {k: [] for k in ("STEP", "TIME", "DT", "FSUM", "FCUT", "DV", "BNORM", "RNORM", "ITR")}
        for m in _STEP_RE.finditer(listing):
            vals = m.groups()
            traj["STEP"].append(int(vals[0]))
            for key, v in zip(("TIME", "DT", "FSUM", "FCUT", "DV", "BNORM", "RNORM"), vals[1:8]):
                traj[key].append(float(v))
            traj["ITR"].append(int(vals[8]))
        # full-precision per-step scalars from the recorder
        dv, dtf, tnf = {}, {}, {}
        for f in sorted((td / "dump").glob("s*_dv.bin")):
            d = read_dump(f)
            s = int(d["NNOW"][0])
            dv[s] = float(d["DV"][0])
            if "DTNOW" in d:
                dtf[s] = float(d["DTNOW"][0])
                tnf[s] = float(d["TNOW"][0])
        steps_dv = sorted(dv)
        extra = {}
        if len(dtf) == len(dv):
            extra = dict(DT_FULL=np.array([dtf[s] for s in steps_dv]),
                         TNOW_FULL=np.array([tnf[s] for s in steps_dv]))
        np.savez_compressed(HERE / f"{name}_trajectory.npz",
                            **{k: np.array(v) for k, v in traj.items()},
                            DV_STEPS=np.array(steps_dv), DV_FULL=np.array([dv[s] for s in steps_dv]),
                            DECK=np.array(case.deck), LIST=np.array(list_lines), **extra)
        static = read_dump(td / "dump" / "static.bin")
        extra_files = {f"FILE_{k.upper()}": np.array(v) for k, v in case.meta.get("files", {}).items()}
        np.savez_compressed(HERE / f"{name}_static.npz", **static, DECK=np.array(case.deck),
                            **extra_files)
        for s in keep_steps:
            rec = {}
            keep = _KEEP.get(name, {}).get(s)
            tags = ["v1_in", "ps_in", "ps_out", "v1_out", "f1_in", "f1_out"]
            # sub-loop passes 2.. of VF_V1CAL (OPTION SUB-LOOP)
            tags += [f"{t}_l{l}" for l in range(2, 5) for t in ("v1_in", "ps_in", "ps_out", "v1_out")]
            for tag in tags:
                p = td / "dump" / f"s{s:05d}_{tag}.bin"
                if keep is not None and tag not in keep:
                    continue
                if p.exists():
                    for k, v in read_dump(p).items():
                        if keep is not None and keep[tag] is not None and k not in keep[tag]:
                            continue
                        rec[f"{tag}:{k}"] = v
            np.savez_compressed(HERE / f"{name}_s{s:05d}.npz", **rec)
        print(f"{name}: {len(traj['STEP'])} list steps, kept {keep_steps}")


def main(argv):
    if not REF_EXE.exists():
        subprocess.run([str(ROOT / "oracle" / "build_ref.sh")], check=True)
    cases = _cases()
    names = argv or list(cases)
    for n in names:
        fac, keep = cases[n]
        run_case(n, fac, keep)


if __name__ == "__main__":
    main(sys.argv[1:])

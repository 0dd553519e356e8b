"""Benchmark of the gfx950 VOF-SMAC step (one JSON line on stdout).

Workload (default): the 512x256x128-cell tank (16.8 M cells, 0.02 m cells) with
a 1:20 plane beach from x=6 m, 40 seeded rectangular buildings on the beach
half and a Boussinesq solitary wave H/h=0.15 on h=0.8 m, plain FP64 form, MILU
preconditioned BiCGSTAB stopped at a relative preconditioned residual of 1e-8
and a true relative residual of 1e-8.  A "step" is one SMAC time step (predictor, Poisson solve, correction,
VOF advection and flagging) over the whole grid; fields stay resident in HBM.

The forms run on the same workload from the same initial state (--form
selects them; the default "all" runs the learned form when its weights file
exists): the plain FP64 form (the reference's MILU-BiCGSTAB, stopped on its
preconditioned residual and held to a true relative residual of 1e-8) gives
value, roofline and the step figures; the mixed form (FP64 iterative
refinement around an FP32 BiCGSTAB with FP32 stencil matvecs and an FP32
aggregation-multigrid preconditioner, FP64 dots, stopped on the true residual
||BB - A PT|| <= 1e-8 ||BB||) and the learned form (FP64 flexible GMRES
preconditioned by one FP32 V-cycle and the BF16 3-D U-Net, MILU-BiCGSTAB as
the fallback, same stop) are reported under "mixed" and "learned" with their
speed-ups and their fields' differences from the plain form's after the same
steps.

value  = interior cells x timed steps / wall time of the timed steps (each step
         timed on its own, as its two halves VF_V1CAL and VF_F1CAL; the true
         residual of its solve, read between them before VF_F1CAL rebuilds the
         activity mask, is outside the timer).  All forms are held to a true
         relative residual ||BB - A PT|| / ||BB|| <= 1e-8; the worst over the
         timed steps is reported.
roofline = (plain form) the MILU preconditioner application M^-1 (forward + backward
           triangular sweeps, the solver's dominant kernel): 12 FP64 streams per
           active cell (DD, ALI, ALJ, ALK, y, x forward; DD, AUI, AUJ, AUK, x,
           x backward) = 96 B per active cell per application, divided by its
           average duration over every fourth of its launches inside the timed
           steps' solves, each bracketed by HIP events on the library's stream
           (cad_set_kernel_timing; "launches" counts the timed ones); isolated_avg_ms is the same application
           repeated back to back after the steps (cad_time_minv).
mixed.roofline = the FP32 multigrid V-cycle, timed the same way inside the
           timed steps, against its algorithmic bytes from the level sizes and
           each solve's K range (averaged over the launches); isolated_avg_ms
           from cad_time_precond; beside it the same V-cycle in FP64
           (cad_time_vcycle), what FP32 buys.
learned.roofline = the U-Net's forward pass, timed the same way inside the
           timed steps, against the BF16 dense MFMA peak (2.5 PFLOP/s), 2 FLOP
           per multiply-add on each solve's box (averaged over the launches);
           isolated_avg_ms from cad_time_learned.
cpu_baseline = the reference itself (oracle/_ref/cadmas_ref.x, compiled from
           its Fortran sources) on one host core: one SMAC step of the
           512x64x128 slab of the same tank (same beach, wave and cell size, 10
           buildings) restarted from the slab's initial state, about 15 s, as
           cell-updates/s; cpu_port = the C restatement (oracle/) on the same
           slab, three steps.  Without the executable, cpu_baseline is the C
           restatement's figure.

--calibrate-copy runs one device-to-device copy of 1 GiB (torch) and nothing
else: its FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc scale the counters of
the sweep kernels to bytes.
"""
import argparse
import importlib.util
import re
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG_DIR = ROOT / "multiscale-and-multiphysics-integrated-simulator-for-tsunami_amd"
METRIC = "SMAC cell-updates/s + Poisson ms/step at 1e-8 resid: learned/mixed vs plain FP64"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# the dominant kernels' in-solve timing takes every fourth launch: its events
# then cost the timed steps little
KT_EVERY = 4
TRUE_RESIDUAL = 1.0e-8


def _pkg():
    name = "cadmas_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(pkg, name):
    cfg = pkg.configs
    if name == "buildings_c":
        return cfg.buildings_tank(eps_r=1.0e-8)
    if name == "runup_b":
        return cfg.solitary_runup(eps_r=1.0e-8)
    if name == "dam_a":
        return cfg.dam_break(eps_r=1.0e-8)
    if name == "bore_e":
        return cfg.bore_tank(eps_r=1.0e-8)
    if name == "bore_e_slab":
        return cfg.bore_tank(1024, 32, 192, nbuild=8, eps_r=1.0e-8, name="bore_e_slab")
    if name == "buildings_c_slab":
        return cfg.buildings_tank(512, 64, 128, nbuild=10, eps_r=1.0e-8, name="buildings_c_slab")
    if name.startswith("buildings_small"):
        return cfg.buildings_tank(128, 64, 48, eps_r=1.0e-8, nbuild=8, name=name)
    raise SystemExit(f"unknown workload {name}")


def prepare(pkg, case):
    lay = pkg.layout.from_text(case.deck)
    state = lay.raw_state()
    flags = pkg.capi.INIT_VELOCITY | pkg.capi.INIT_PRESSURE
    if case.ic is not None:
        # initial condition the deck cannot express, applied like a restart
        case.ic(lay.XX, lay.ZZ, lay.NF0, state)
        flags = 0
    return lay, state, flags


def measured_traffic(workload_name, key="traffic_bytes_per_minv",
                     summary="sweep_traffic_buildings_512x256x128.md"):
    """HBM bytes of one application of the dominant kernel (the plain form's
    M^-1, with key="traffic_bytes_per_vcycle" the mixed form's V-cycle, with
    key="traffic_bytes_per_unet_forward" the learned form's network pass)
    measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    of known size) for the default workload, from the committed profile
    summary; None for other workloads."""
    if workload_name != "buildings_c":
        return None
    f = ROOT / "profiles" / summary
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(key + ":"):
            return int(line.split(":")[1])
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(pkg, lay, state, flags, steps=1):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    arr = dict(lay.static_arrays())
    arr.update(state)
    F = po.Fields(arr, lay.params())
    rc = F.init_state(lay.deck.wvlvl, lay.deck.uini, flags)
    if rc != 0:
        raise RuntimeError(f"oracle init failed {rc}")
    t0 = time.perf_counter()
    for _ in range(steps):
        rc, _ = F.step()
        if rc != 0:
            raise RuntimeError(f"oracle step failed {rc}")
    dt = time.perf_counter() - t0
    cells = np.prod(lay.deck.ncells)
    return cells * steps / dt, dt, F.st.icgitr


def run_form(pkg, lay, state, flags, form, args):
    """Warm-up then timed steps of one form on a fresh context; the figures of
    the bench line for that form."""
    ni, nj, nk = lay.shape
    cells = int(np.prod(lay.deck.ncells))
    ctx = pkg.capi.Context(ni, nj, nk, lay.numb)
    p = lay.params()
    ctx.set_params(**{k: v for k, v in p.items() if k in ctx.params})
    if form == "mixed":
        ctx.set_params(precision=pkg.capi.PRECISION_MIXED)
    elif form == "learned":
        ctx.set_params(precision=pkg.capi.PRECISION_LEARNED)
        pkg.learned.set_weights(ctx, pkg.learned.load_weights(args.weights))
    # every form stops only at a true relative residual of 1e-8: the plain form
    # rechecks it whenever the reference's preconditioned test passes, and the
    # mixed and learned forms' true-residual test (CGEPSR ||BB|| + CGEPSA) may
    # not exceed it
    ctx.set_true_residual_stop(TRUE_RESIDUAL)
    ctx.upload_static(lay.static_arrays())
    ctx.upload_state(state)
    ctx.init_state(lay.deck.wvlvl, lay.deck.uini, flags)
    log(f"bench: state resident on device, form {form}")
    for w in range(args.warmup):
        info = ctx.step()
        log(f"bench: {form} warmup {w} ITR={info['icgitr']} poisson={info['ms_poisson']:.1f}ms "
            f"true residual {ctx.poisson_residual():.3e}")
    itr, poisson, pred, corr, vof, res = [], [], [], [], [], []
    wall = 0.0
    loops = ctx.get_params()["loops"]
    # HIP events around every KT_EVERY-th launch of the form's dominant
    # kernel inside the timed steps' solves (read after them)
    ctx.set_kernel_timing(KT_EVERY)
    for s in range(args.steps):
        # each phase returns after its work has finished on the device.  With
        # one pass a step, the step is run as its two halves, VF_V1CAL then
        # VF_F1CAL (what cad_step runs, TNOW advanced first as it does), and
        # the true residual of the solve is read between them with the timer
        # stopped: VF_F1CAL rebuilds the activity mask for the next step.
        if loops == 1:
            ctx.set_params(tnow=ctx.get_params()["tnow"] + ctx.get_params()["dtnow"])
            t0 = time.perf_counter()
            info = ctx.v1cal()
            wall += time.perf_counter() - t0
            solve_res = ctx.poisson_residual()
            t0 = time.perf_counter()
            info["ms_vof"] = ctx.f1cal()["ms_vof"]
            wall += time.perf_counter() - t0
        else:
            t0 = time.perf_counter()
            info = ctx.step()
            wall += time.perf_counter() - t0
            solve_res = ctx.poisson_residual()
        itr.append(info["icgitr"])
        poisson.append(info["ms_poisson"])
        pred.append(info["ms_predictor"])
        corr.append(info["ms_correction"])
        vof.append(info["ms_vof"])
        res.append(solve_res)
        log(f"bench: {form} step {s} ITR={info['icgitr']} poisson={info['ms_poisson']:.1f}ms "
            f"true residual {res[-1]:.3e}")
    ctx.set_kernel_timing(False)
    true_res = max(res)
    log(f"bench: {form} timed {args.steps} steps in {wall:.3f}s ITR={itr} worst true residual "
        f"{true_res:.3e}")
    end_state = ctx.download_state() if args.form in ("both", "all") else None
    names = {"plain": "plain FP64", "mixed": "mixed FP64/FP32", "learned": "learned BF16/FP32/FP64"}
    texts = {"plain": "plain FP64 form, MILU-BiCGSTAB to 1e-8 relative preconditioned and true residual",
             "mixed": "mixed form, FP64 iterative refinement around FP32 BiCGSTAB (FP32 stencil "
                      "matvecs, FP32 aggregation-multigrid preconditioner, FP64 dots) to 1e-8 relative "
                      "true residual",
             "learned": "learned form, FP64 flexible GMRES preconditioned by an FP32 V-cycle and the "
                        "BF16 3-D U-Net on the matrix cores, MILU-BiCGSTAB fallback, to 1e-8 relative "
                        "true residual"}
    r = {"form": names[form], "form_text": texts[form],
         "value": round(cells * args.steps / wall, 1), "ms_per_step": round(1e3 * wall / args.steps, 3),
         "poisson_ms_per_step": round(float(np.mean(poisson)), 3),
         "poisson_true_residual": float(f"{true_res:.3e}"),
         "poisson_true_residual_note": "worst ||BB - A PT|| / ||BB|| over the timed steps, "
                                       "right after each solve",
         "itr_per_step": float(np.mean(itr)),
         "phase_ms": {"predictor": round(float(np.mean(pred)), 3),
                      "poisson": round(float(np.mean(poisson)), 3),
                      "correction": round(float(np.mean(corr)), 3),
                      "vof": round(float(np.mean(vof)), 3)}}
    if form == "plain":
        # dominant kernel: M^-1, its launches inside the timed steps' solves;
        # beside it the same application repeated back to back on the last
        # step's factorisation
        minv_ms, _, launches = ctx.kernel_timing(pkg.capi.KERNEL_MINV)
        iso_ms = ctx.time_minv(args.minv_reps)
        active = ctx.count_active()
        minv_bytes = 96.0 * active
        achieved = minv_bytes / (minv_ms * 1e-3) / 1e9
        r["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                         "traffic": measured_traffic(args.workload),
                         "kernel": "MILU M^-1 sweeps (forward+backward), 96 B/active cell",
                         "avg_ms": round(minv_ms, 4), "launches": launches,
                         "isolated_avg_ms": round(iso_ms, 4), "active_cells": active}
    elif form == "learned":
        # the learned form's network: one forward pass against the BF16 MFMA peak
        st = pkg.learned.stats(ctx)
        net_ms, net_flop, launches = ctx.kernel_timing(pkg.capi.KERNEL_UNET)
        iso_ms, _ = ctx.time_learned(args.minv_reps)
        achieved = net_flop / (net_ms * 1e-3) / 1e12
        r["roofline"] = {"bound": "mfma", "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFLOPS, 5),
                         "traffic": measured_traffic(args.workload, "traffic_bytes_per_unet_forward",
                                                     "phases_and_unet_traffic_buildings_512x256x128.md"),
                         "kernel": "3-D U-Net forward pass (BF16 MFMA, FP32 accumulation), "
                                   f"{pkg.unet.param_count()} parameters, box depth {st['depth']}",
                         "avg_ms": round(net_ms, 4), "launches": launches,
                         "isolated_avg_ms": round(iso_ms, 4), "flop": round(net_flop)}
        r["fallbacks"] = st["fallbacks"]
        r["fgmres_iters_total"] = st["total_iters"]
    else:
        # dominant kernel of the mixed form: one multigrid V-cycle
        vc_ms, vc_bytes, launches = ctx.kernel_timing(pkg.capi.KERNEL_VCYCLE)
        iso_ms, _ = ctx.time_precond(args.minv_reps)
        achieved = vc_bytes / (vc_ms * 1e-3) / 1e9
        r["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                         "traffic": measured_traffic(args.workload, "traffic_bytes_per_vcycle",
                                                     "multigrid_traffic_buildings_512x256x128.md"),
                         "kernel": "FP32 multigrid V-cycle (the mixed form's preconditioner)",
                         "avg_ms": round(vc_ms, 4), "launches": launches,
                         "isolated_avg_ms": round(iso_ms, 4), "bytes": round(vc_bytes)}
        # the same V-cycle with every level in FP64: what FP32 buys
        v64_ms, v64_bytes = ctx.time_vcycle(args.minv_reps, fp64=True)
        r["vcycle_fp64"] = {"avg_ms": round(v64_ms, 4), "bytes": round(v64_bytes),
                            "fp32_speedup": round(v64_ms / iso_ms, 3)}
    ctx.close()
    r["state"] = end_state
    return r


def field_errors(a, b):
    """Relative differences of state a against b after the same steps: the
    velocity components against b's velocity scale, P and F against their own
    scale, and the number of cells whose flag differs."""
    vs = max(float(np.max(np.abs(b[k]))) for k in ("UU", "VV", "WW")) or 1.0
    e = {"velocity": max(float(np.max(np.abs(a[k] - b[k]))) for k in ("UU", "VV", "WW")) / vs}
    for k, n in (("PP", "pressure"), ("FF", "F")):
        sc = float(np.max(np.abs(b[k]))) or 1.0
        e[n] = float(np.max(np.abs(a[k] - b[k]))) / sc
    out = {k: float(f"{v:.3e}") for k, v in e.items()}
    out["flags_differing"] = int(np.count_nonzero(a["NF"] != b["NF"]))
    return out


REF_EXE = ROOT / "oracle" / "_ref" / "cadmas_ref.x"


def cpu_baseline_reference(pkg, case, lay, state, flags, steps=1, timeout=600):
    """The reference itself (compiled from its Fortran sources by
    oracle/build_ref.sh; one rank, one core) on the sample: it restarts from
    the sample's initial state, written as <name>.res by rsl.Writer from the
    C restatement's VF_CINIT, and runs `steps` steps; a run of 0 steps from
    the same file is subtracted, so reading the deck and the restart file are
    not timed.  None when the executable is not there or fails."""
    import shutil
    import subprocess
    import tempfile
    if not REF_EXE.exists():
        return None
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    arr = dict(lay.static_arrays())
    arr.update(state)
    F = po.Fields(arr, lay.params())
    if F.init_state(lay.deck.wvlvl, lay.deck.uini, flags) != 0:
        return None
    name = "sample"
    keys = ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF")
    info = {"fsum": 0.0, "fcut": 0.0, "cgbnrm": 0.0, "cgxnrm": 0.0, "icgitr": 0}
    body = "\n".join(ln for ln in case.deck.splitlines() if not ln.startswith(("TIME END", "FILE ")))
    with tempfile.TemporaryDirectory(prefix="cadmas_ref_") as top:
        top = Path(top)
        with pkg.rsl.Writer(top / f"{name}.res", lay.shape, lay.numb, lay.static_arrays()["GGV"],
                            idrop=lay.deck.idrop) as w:
            w.write_step(0, 0.0, lay.params()["dtnow"], info, {k: F[k] for k in keys},
                         {"TBUB": np.zeros(lay.shape[2])})
        times, itr = {}, None
        for n in (0, steps):
            run = top / f"run{n}"
            run.mkdir()
            shutil.copy(top / f"{name}.res", run / f"{name}.res")
            (run / f"{name}.in").write_text(body + f"\nTIME END {n} 1.0D9\nFILE RES 0\n")
            (run / "data.env").write_text(f"1\n{name} 1 0\n")
            env = {k: v for k, v in os.environ.items() if not k.startswith("CADMAS_DUMP")}
            env["OMP_NUM_THREADS"] = "1"
            t0 = time.perf_counter()
            r = subprocess.run([str(REF_EXE)], cwd=run, env=env, capture_output=True, text=True,
                               timeout=timeout)
            times[n] = time.perf_counter() - t0
            if r.returncode != 0 or "NORMAL END" not in r.stdout:
                log(f"bench: reference run of {n} steps failed (rc {r.returncode})")
                return None
            if n:
                lst = (run / f"{name}.list").read_text()
                itr = [int(x) for x in re.findall(r"ITR\s*=\s*(\d+)", lst)][1:]
    dt = times[steps] - times[0]
    cells = int(np.prod(lay.deck.ncells))
    return cells * steps / dt, dt, itr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=os.environ.get("CADMAS_BENCH_WORKLOAD", "buildings_c"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--minv-reps", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--form", choices=("plain", "mixed", "learned", "both", "all"), default="all",
                    help="both: plain and mixed; all: also the learned form when its weights exist")
    ap.add_argument("--weights", default=None, help="the learned form's network parameters (.npy)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 1 << 27                                    # 2^27 doubles = 1 GiB
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": n * 8, "bytes_written": n * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs one grid on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not an implementation).", file=sys.stderr)
        sys.exit(2)
    pkg = _pkg()
    case = workload(pkg, args.workload)
    t0 = time.perf_counter()
    lay, state, flags = prepare(pkg, case)
    log(f"bench: {case.name} {lay.deck.ncells} layout {time.perf_counter() - t0:.1f}s")
    cells = int(np.prod(lay.deck.ncells))
    if args.weights is None:
        args.weights = str(pkg.learned.DEFAULT_WEIGHTS)
    if args.form == "all":
        forms = ("plain", "mixed") + (("learned",) if Path(args.weights).exists() else ())
    elif args.form == "both":
        forms = ("plain", "mixed")
    else:
        forms = (args.form,)
    res = {f: run_form(pkg, lay, state, flags, f, args) for f in forms}
    head = res["plain"] if "plain" in res else res[forms[0]]
    out = {
        "metric": METRIC, "value": head["value"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head is res.get("plain") else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"{case.name}: {head['form_text']}", "cells": cells,
                   "grid": list(lay.deck.ncells), "form": head["form"],
                   "parallelism": "single GPU"},
        "poisson_ms_per_step": head["poisson_ms_per_step"],
        "poisson_true_residual": head["poisson_true_residual"],
        "itr_per_step": head["itr_per_step"],
        "phase_ms": head["phase_ms"],
    }
    if "plain" in res:
        pl = res["plain"]
        # the step's byte model over the cells each part touches: predictor,
        # assembly, correction and VOF (716 B) over the interior cells; MILU
        # factorisation and solver set-up (396 B) and each BiCGSTAB iteration
        # (496 B) over the active cells the solver's vectors cover
        active = pl["roofline"]["active_cells"]
        step_bytes = 716.0 * cells + (396.0 + 496.0 * pl["itr_per_step"]) * active
        out["step_algorithmic_GBps"] = round(step_bytes / (pl["ms_per_step"] * 1e-3) / 1e9, 2)
        out["step_algorithmic_frac"] = round(step_bytes / (pl["ms_per_step"] * 1e-3) / 1e9
                                             / HBM_PEAK_GBS, 4)
        out["roofline"] = pl["roofline"]
    if "mixed" in res:
        mx = res["mixed"]
        out["mixed"] = {k: mx[k] for k in ("value", "ms_per_step", "poisson_ms_per_step",
                                           "poisson_true_residual", "itr_per_step", "phase_ms",
                                           "form_text", "roofline", "vcycle_fp64")}
        if "plain" in res:
            out["mixed"]["field_error_vs_plain"] = field_errors(mx["state"], res["plain"]["state"])
            out["mixed"]["speedup_vs_plain"] = round(mx["value"] / res["plain"]["value"], 3)
            out["mixed"]["poisson_speedup_vs_plain"] = round(
                res["plain"]["poisson_ms_per_step"] / mx["poisson_ms_per_step"], 3)
        if "roofline" not in out:
            out["roofline"] = None
    if "learned" in res:
        ln = res["learned"]
        out["learned"] = {k: ln[k] for k in ("value", "ms_per_step", "poisson_ms_per_step",
                                             "poisson_true_residual", "itr_per_step", "phase_ms",
                                             "form_text", "roofline", "fallbacks", "fgmres_iters_total")}
        out["learned"]["weights"] = Path(args.weights).name
        if "plain" in res:
            out["learned"]["field_error_vs_plain"] = field_errors(ln["state"], res["plain"]["state"])
            out["learned"]["speedup_vs_plain"] = round(ln["value"] / res["plain"]["value"], 3)
            out["learned"]["poisson_speedup_vs_plain"] = round(
                res["plain"]["poisson_ms_per_step"] / ln["poisson_ms_per_step"], 3)
        if "roofline" not in out:
            out["roofline"] = None
    if not args.no_cpu_baseline:
        slabs = {"buildings_c": "buildings_c_slab", "bore_e": "bore_e_slab"}
        sample = workload(pkg, slabs[args.workload]) if args.workload in slabs else case
        slay, sstate, sflags = prepare(pkg, sample)
        scells = int(np.prod(slay.deck.ncells))
        v, dt, it = cpu_baseline(pkg, slay, sstate, sflags, steps=args.cpu_steps)
        port = {"value": round(v, 1), "unit": "cell-updates/s", "cores": 1, "kind": "port",
                "sample": f"{args.cpu_steps} SMAC steps of {sample.name} {list(slay.deck.ncells)} "
                          f"({scells} cells, {dt:.1f} s, ITR={it}) with the C restatement in "
                          "oracle/, one thread"}
        ref = cpu_baseline_reference(pkg, sample, slay, sstate, sflags)
        if ref is None:
            out["cpu_baseline"] = port
        else:
            rv, rdt, ritr = ref
            out["cpu_baseline"] = {
                "value": round(rv, 1), "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": f"1 SMAC step of {sample.name} {list(slay.deck.ncells)} ({scells} cells, "
                          f"{rdt:.1f} s, ITR={ritr}) by the reference's own Fortran (oracle/_ref, "
                          "flang -O2, one MPI rank), restarted from the same initial state; the "
                          "time of a 0-step run from the same restart file is subtracted"}
            out["cpu_port"] = port
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark of the gfx950 VOF-SMAC step (one JSON line on stdout).

Workload (default): the 512x256x128-cell tank (16.8 M cells, 0.02 m cells) with
a 1:20 plane beach from x=6 m, 40 seeded rectangular buildings on the beach
half and a Boussinesq solitary wave H/h=0.15 on h=0.8 m, plain FP64 form, MILU
preconditioned BiCGSTAB stopped at a relative (preconditioned) residual of
1e-8.  A "step" is one SMAC time step (predictor, Poisson solve, correction,
VOF advection and flagging) over the whole grid; fields stay resident in HBM.

value  = interior cells x timed steps / wall time of the timed steps.
roofline = the MILU preconditioner application M^-1 (forward + backward
           triangular sweeps, the solver's dominant kernel): 12 FP64 streams per
           active cell (DD, ALI, ALJ, ALK, y, x forward; DD, AUI, AUJ, AUK, x,
           x backward) = 96 B per active cell per application, divided by its
           average duration timed with HIP events on the library's stream.
cpu_baseline = the C restatement (oracle/) on one host core: three SMAC steps
           of the 512x64x128 slab of the same tank (same beach, wave and cell
           size, 10 buildings), a bounded sample of about 15 s, as
           cell-updates/s.

--calibrate-copy runs one device-to-device copy of 1 GiB (torch) and nothing
else: its FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc scale the counters of
the sweep kernels to bytes.
"""
import argparse
import importlib.util
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


def measured_traffic(workload_name):
    """HBM bytes of one M^-1 application measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size) for the default
    workload, from the committed profile summary; None for other workloads."""
    if workload_name != "buildings_c":
        return None
    f = ROOT / "profiles" / "sweep_traffic_buildings_512x256x128.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_minv:"):
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=os.environ.get("CADMAS_BENCH_WORKLOAD", "buildings_c"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--minv-reps", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--form", choices=("plain", "mixed"), default="plain")
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
    ni, nj, nk = lay.shape
    cells = int(np.prod(lay.deck.ncells))
    ctx = pkg.capi.Context(ni, nj, nk, lay.numb)
    p = lay.params()
    ctx.set_params(**{k: v for k, v in p.items() if k in ctx.params})
    if args.form == "mixed":
        ctx.set_params(precision=pkg.capi.PRECISION_MIXED)
    ctx.upload_static(lay.static_arrays())
    ctx.upload_state(state)
    ctx.init_state(lay.deck.wvlvl, lay.deck.uini, flags)
    log(f"bench: state resident on device, form {args.form}")
    for w in range(args.warmup):
        info = ctx.step()
        log(f"bench: warmup {w} ITR={info['icgitr']} poisson={info['ms_poisson']:.1f}ms "
            f"true residual {ctx.poisson_residual():.3e}")
    itr, poisson, pred, corr, vof = [], [], [], [], []
    t0 = time.perf_counter()
    for s in range(args.steps):
        info = ctx.step()
        itr.append(info["icgitr"])
        poisson.append(info["ms_poisson"])
        pred.append(info["ms_predictor"])
        corr.append(info["ms_correction"])
        vof.append(info["ms_vof"])
        log(f"bench: step {s} ITR={info['icgitr']} poisson={info['ms_poisson']:.1f}ms")
    wall = time.perf_counter() - t0
    true_res = ctx.poisson_residual()
    log(f"bench: timed {args.steps} steps in {wall:.3f}s ITR={itr} true residual {true_res:.3e}")
    ms_step = 1e3 * wall / args.steps
    value = cells * args.steps / wall
    # dominant kernel: M^-1 application on the last step's factorisation
    minv_ms = ctx.time_minv(args.minv_reps)
    active = ctx.count_active()
    minv_bytes = 96.0 * active
    achieved = minv_bytes / (minv_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(args.workload),
                "kernel": "MILU M^-1 sweeps (forward+backward), 96 B/active cell",
                "avg_ms": round(minv_ms, 4), "active_cells": active}
    mean_itr = float(np.mean(itr))
    # step-level algorithmic traffic: 1112 + 496*ITR bytes per cell
    step_bytes = (1112.0 + 496.0 * mean_itr) * cells
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{case.name} plain FP64 form, MILU-BiCGSTAB to 1e-8 relative "
                               "(preconditioned) residual", "cells": cells,
                   "grid": list(lay.deck.ncells), "form": "plain FP64",
                   "parallelism": "single GPU"},
        "poisson_ms_per_step": round(float(np.mean(poisson)), 3),
        "poisson_true_residual": float(f"{true_res:.3e}"),
        "itr_per_step": mean_itr,
        "phase_ms": {"predictor": round(float(np.mean(pred)), 3),
                     "poisson": round(float(np.mean(poisson)), 3),
                     "correction": round(float(np.mean(corr)), 3),
                     "vof": round(float(np.mean(vof)), 3)},
        "step_algorithmic_GBps": round(step_bytes / (ms_step * 1e-3) / 1e9, 2),
        "roofline": roofline,
    }
    ctx.close()
    if not args.no_cpu_baseline:
        sample = workload(pkg, "buildings_c_slab") if args.workload == "buildings_c" else case
        slay, sstate, sflags = prepare(pkg, sample)
        v, dt, it = cpu_baseline(pkg, slay, sstate, sflags, steps=args.cpu_steps)
        scells = int(np.prod(slay.deck.ncells))
        out["cpu_baseline"] = {"value": round(v, 1), "unit": "cell-updates/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{args.cpu_steps} SMAC steps of {sample.name} {list(slay.deck.ncells)} "
                                         f"({scells} cells, {dt:.1f} s, ITR={it}) with the C restatement in "
                                         "oracle/, one thread"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

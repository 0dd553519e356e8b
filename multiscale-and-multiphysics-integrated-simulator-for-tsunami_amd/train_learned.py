"""Trains the learned form's network and writes its weights.

    python multiscale-and-multiphysics-integrated-simulator-for-tsunami_amd/train_learned.py \
        --out <weights.npy> [--steps 3000] [--loss residual|error] [--lr 1e-3]

Pairs from the plain solver on the 512x256x128 buildings tank with seeded
layouts: seeds 1..8 train, seeds 9..12 are held out (learned.py).  Writes the
FP32 parameter vector (.npy) and a .json beside it with the losses, the pair
counts and the held-out layouts' FGMRES iterations with and without the
network.  Runs on one GPU.
"""
import argparse
import importlib.util
import sys
import time
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent


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


def log(msg):
    print(msg, flush=True)


def solve_iters(pkg, seed, p, steps, size, nbuild):
    """FGMRES iterations per step of the learned form on one layout (after
    `warm` mixed steps), with parameters p."""
    L = pkg.learned
    case = pkg.configs.buildings_tank(*size, nbuild=nbuild, seed=seed, eps_r=1.0e-8)
    out = []
    with L._context(case, pkg.capi.PRECISION_MIXED) as ctx:
        for _ in range(10):
            ctx.step()
        L.set_weights(ctx, p)
        ctx.set_params(precision=pkg.capi.PRECISION_LEARNED)
        for _ in range(steps):
            info = ctx.step()
            st = L.stats(ctx)
            out.append((st["iters"], st["fallback_iters"], info["cgxnrm"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--samples", type=int, default=10)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--crops", type=int, default=40)
    ap.add_argument("--train-seeds", default="1,2,3,4,5,6,7,8")
    ap.add_argument("--test-seeds", default="9,10,11,12")
    ap.add_argument("--size", default="512,256,128")
    ap.add_argument("--nbuild", type=int, default=40)
    ap.add_argument("--loss", choices=("error", "residual"), default="residual")
    ap.add_argument("--lr", type=float, default=1.0e-3)
    args = ap.parse_args()
    pkg = _pkg()
    L = pkg.learned
    size = tuple(int(v) for v in args.size.split(","))
    tr = [int(s) for s in args.train_seeds.split(",")]
    te = [int(s) for s in args.test_seeds.split(",")]
    t0 = time.time()
    ps = L.generate(tr, samples=args.samples, every=args.every, crops_per_pair=args.crops,
                    capacity=len(tr) * args.samples * len(L.ITERS) * args.crops, size=size,
                    nbuild=args.nbuild, log=log, seed=1, loss=args.loss)
    test = L.generate(te, samples=max(2, args.samples // 4), every=args.every, crops_per_pair=10,
                      capacity=len(te) * max(2, args.samples // 4) * len(L.ITERS) * 10, size=size,
                      nbuild=args.nbuild, log=log, seed=2, loss=args.loss)
    t_gen = time.time() - t0
    log(f"train: {ps.n} training crops, {test.n} held-out crops in {t_gen:.1f} s")
    t1 = time.time()
    p, hist = L.train(ps, steps=args.steps, lr=args.lr, log=log)
    t_train = time.time() - t1
    train_loss, test_loss = L.evaluate(ps, p), L.evaluate(test, p)
    zero = np.zeros_like(p)
    log(f"train: loss train {train_loss:.4f} held-out {test_loss:.4f} (zero network: "
        f"{L.evaluate(test, zero):.4f}); {t_train:.1f} s")
    held = {}
    for sd in te[:2]:
        with_net = solve_iters(pkg, sd, p, 3, size, args.nbuild)
        without = solve_iters(pkg, sd, zero, 3, size, args.nbuild)
        held[sd] = {"with_network": with_net, "vcycle_only": without}
        log(f"train: held-out layout {sd}: FGMRES (iters, fallback iters, true residual) "
            f"with the network {with_net}, V-cycle only {without}")
    meta = {"pairs": "plain-solver iterates " + str(list(L.ITERS)) + " per system, after one V-cycle",
            "loss": args.loss, "lr": args.lr,
            "train_seeds": tr, "test_seeds": te, "grid": list(size), "buildings": args.nbuild,
            "train_crops": ps.n, "test_crops": test.n, "crop": list(L.CROP), "steps": args.steps,
            "loss_history": hist, "train_loss": train_loss, "test_loss": test_loss,
            "held_out_fgmres": held, "seconds_generate": round(t_gen, 1),
            "seconds_train": round(t_train, 1)}
    L.save(args.out, p, meta)
    log(f"train: wrote {args.out}")


if __name__ == "__main__":
    main()

"""Training of the learned form's preconditioner network (csrc/cad_unet.hip).

Pairs come from the plain solver itself (cad_learned_sample): on a step's
Poisson system the plain MILU-BiCGSTAB solve gives x*, and its iterates x_k
from 0 give pairs (r_k = BB - A x_k, e_k = x* - x_k); after the V-cycle that
precedes the network in the preconditioner, the pair is (r_k - A V(r_k),
e_k - V(r_k)).  The network sees the residual scaled by its RMS over the
active cells, GGV, F and the cell type; its target is the correction on the
same scale.

Two losses train it.  "error": the relative squared error of the correction
over the active voxels of a crop's interior.  "residual": the relative
squared residual that the correction leaves, sum m (rho - A y)^2 / sum m rho^2
with rho = A e (the pair's residual on the same scale) and A the Poisson
matrix on the crop (cad_learned_box_matrix), y taken as 0 on inactive voxels
as the preconditioner does.  The second asks the network for what a local
operator can give, the part of A^-1 that a crop sees, where the first also
asks for the smooth error spread over the tank.

Workload: the 512x256x128 buildings tank (configs.buildings_tank) with seeded
layouts -- seeds 1..8 for training, 9..12 held out.  Each layout's flow is
advanced with the mixed form and sampled every few steps; each sample gives
one pair per k, and each pair several crops of 64x32x32 voxels placed where
the water is.  Both losses are taken over the active voxels of a crop's
interior (a margin of 4 voxels at the crop's faces is left out: there the
network sees zero padding where the full box has data).

Device memory comes from torch; the network's kernels are the library's.
"""
import ctypes as C
import json
import time
from pathlib import Path

import numpy as np

from . import capi, configs, layout
from . import unet as U

CROP = (64, 32, 32)            # x, y, z voxels
MARGIN = 4
ITERS = (0, 1, 2, 4, 8, 16)    # plain-solver iterates sampled per system
LOSSES = ("error", "residual")
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"
DEFAULT_WEIGHTS = WEIGHTS_DIR / "unet_buildings_512x256x128.npy"


def _bind():
    L = capi.lib()
    if getattr(L, "_learned_bound", False):
        return L
    vp = C.c_void_p
    L.cad_set_learned_weights.argtypes = [vp, C.POINTER(C.c_float), C.c_int]
    L.cad_set_learned_mode.argtypes = [vp, C.c_int, C.c_int]
    L.cad_learned_stats.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.cad_learned_box.argtypes = [vp, C.POINTER(C.c_int)]
    L.cad_learned_sample.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_int, vp, vp, vp]
    L.cad_learned_box_matrix.argtypes = [vp, vp]
    L._learned_bound = True
    return L


def set_weights(ctx, p):
    p = np.ascontiguousarray(p, np.float32)
    capi._check(_bind().cad_set_learned_weights(ctx._h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))


def set_mode(ctx, mode=1, restart=30):
    capi._check(_bind().cad_set_learned_mode(ctx._h, int(mode), int(restart)))


def stats(ctx):
    """last solve's FGMRES iterations, its fallback BiCGSTAB iterations,
    solves, solves that fell back, all FGMRES iterations, the box depth."""
    a = (C.c_longlong * 6)()
    capi._check(_bind().cad_learned_stats(ctx._h, a))
    keys = ("iters", "fallback_iters", "solves", "fallbacks", "total_iters", "depth")
    return dict(zip(keys, [int(v) for v in a]))


def box(ctx):
    b = (C.c_int * 4)()
    capi._check(_bind().cad_learned_box(ctx._h, b))
    return tuple(int(v) for v in b)


def load_weights(path=DEFAULT_WEIGHTS):
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.size != U.param_count():
        raise ValueError(f"{path}: {p.size} {p.dtype} values, the network has {U.param_count()} float32")
    return p


def sample(ctx, mode=1, iters=ITERS):
    """Pairs of the context's last Poisson system: (features, target, mask)
    as torch CUDA tensors of shapes (n, nz+2, ny+2, nx+2, 8) BF16 and
    (n, nz, ny, nx) FP32, plus the box (nx, ny, nz, kb)."""
    import torch
    nx, ny, nz, kb = box(ctx)
    n = len(iters)
    feat = torch.zeros((n, nz + 2, ny + 2, nx + 2, 8), dtype=torch.bfloat16, device="cuda")
    tgt = torch.zeros((n, nz, ny, nx), dtype=torch.float32, device="cuda")
    msk = torch.zeros((n, nz, ny, nx), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    it = (C.c_int * n)(*iters)
    capi._check(_bind().cad_learned_sample(ctx._h, int(mode), it, n, C.c_void_p(feat.data_ptr()),
                                           C.c_void_p(tgt.data_ptr()), C.c_void_p(msk.data_ptr())))
    return feat, tgt, msk, (nx, ny, nz, kb)


def box_matrix(ctx):
    """The context's current Poisson matrix on the network's box: torch FP32
    CUDA tensor (7, nz, ny, nx) of the rows AD, ALI, AUI, ALJ, AUJ, ALK, AUK."""
    import torch
    nx, ny, nz, _ = box(ctx)
    a = torch.empty((7, nz, ny, nx), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    capi._check(_bind().cad_learned_box_matrix(ctx._h, C.c_void_p(a.data_ptr())))
    return a


def apply_matrix(a, v):
    """A v on crops: a (b, 7, z, y, x) the rows, v (b, z, y, x); values
    beyond a crop are taken as 0 (only its faces' voxels see them)."""
    import torch.nn.functional as Fn
    p = Fn.pad(v, (1, 1, 1, 1, 1, 1))
    c = p[:, 1:-1, 1:-1, 1:-1]
    return (a[:, 0] * c + a[:, 1] * p[:, 1:-1, 1:-1, :-2] + a[:, 2] * p[:, 1:-1, 1:-1, 2:]
            + a[:, 3] * p[:, 1:-1, :-2, 1:-1] + a[:, 4] * p[:, 1:-1, 2:, 1:-1]
            + a[:, 5] * p[:, :-2, 1:-1, 1:-1] + a[:, 6] * p[:, 2:, 1:-1, 1:-1])


def residual_loss(a, act, e, m, y):
    """Mean over crops of sum m (rho - A (act y))^2 / sum m rho^2, rho = A e."""
    import torch
    with torch.no_grad():
        rho = apply_matrix(a, e)
        den = (m * rho * rho).sum(dim=(1, 2, 3)).clamp_min(1e-30)
    res = m * (rho - apply_matrix(a, act * y))
    return ((res * res).sum(dim=(1, 2, 3)) / den).mean()


def _crop_origins(msk, count, rng, min_active=0.25):
    """Up to count crop origins (x0, y0, z0) whose interior holds at least
    min_active active voxels (fraction), drawn at random."""
    import torch
    n_z, n_y, n_x = msk.shape
    cx, cy, cz = CROP
    if n_x < cx or n_y < cy or n_z < cz:
        return []
    out = []
    for _ in range(count * 20):
        if len(out) == count:
            break
        x0 = int(rng.integers(0, n_x - cx + 1))
        y0 = int(rng.integers(0, n_y - cy + 1))
        z0 = int(rng.integers(0, n_z - cz + 1))
        frac = float(torch.mean(msk[z0:z0 + cz, y0:y0 + cy, x0:x0 + cx]))
        if frac >= min_active:
            out.append((x0, y0, z0))
    return out


class PairSet:
    """Crops of pairs, resident on the GPU; with the residual loss also the
    matrix rows (7 FP32 per voxel) and the activity of each crop."""

    def __init__(self, capacity, loss="error"):
        import torch
        cx, cy, cz = CROP
        self.loss = loss
        self.x = torch.zeros((capacity, cz + 2, cy + 2, cx + 2, 8), dtype=torch.bfloat16, device="cuda")
        self.t = torch.zeros((capacity, cz, cy, cx), dtype=torch.float32, device="cuda")
        self.m = torch.zeros((capacity, cz, cy, cx), dtype=torch.float32, device="cuda")
        if loss == "residual":
            self.a = torch.zeros((capacity, 7, cz, cy, cx), dtype=torch.float32, device="cuda")
            self.act = torch.zeros((capacity, cz, cy, cx), dtype=torch.float32, device="cuda")
        self.n = 0
        inner = torch.zeros((cz, cy, cx), dtype=torch.float32, device="cuda")
        inner[MARGIN:cz - MARGIN, MARGIN:cy - MARGIN, MARGIN:cx - MARGIN] = 1.0
        self.inner = inner

    def add(self, feat, tgt, msk, crops_per_pair, rng, a=None):
        cx, cy, cz = CROP
        for q in range(feat.shape[0]):
            for (x0, y0, z0) in _crop_origins(msk[q], crops_per_pair, rng):
                if self.n >= self.x.shape[0]:
                    return
                i = self.n
                self.x[i, 1:-1, 1:-1, 1:-1, :] = feat[q, 1 + z0:1 + z0 + cz, 1 + y0:1 + y0 + cy,
                                                      1 + x0:1 + x0 + cx, :]
                self.t[i] = tgt[q, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx]
                self.m[i] = msk[q, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx] * self.inner
                if self.loss == "residual":
                    self.a[i] = a[:, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx]
                    self.act[i] = msk[q, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx]
                self.n += 1


def _context(case, precision):
    lay = layout.from_text(case.deck)
    state = lay.raw_state()
    flags = capi.INIT_VELOCITY | capi.INIT_PRESSURE
    if case.ic is not None:
        case.ic(lay.XX, lay.ZZ, lay.NF0, state)
        flags = 0
    ni, nj, nk = lay.shape
    ctx = capi.Context(ni, nj, nk, lay.numb)
    p = lay.params()
    ctx.set_params(**{k: v for k, v in p.items() if k in ctx.params})
    ctx.set_params(precision=precision, cgepsr=1.0e-8)
    ctx.upload_static(lay.static_arrays())
    ctx.upload_state(state)
    ctx.init_state(lay.deck.wvlvl, lay.deck.uini, flags)
    return ctx


def generate(seeds, samples=10, every=10, crops_per_pair=40, capacity=20000, size=(512, 256, 128),
             nbuild=40, mode=1, log=print, seed=0, loss="error"):
    """The crop set from the layouts of the given seeds: per layout `samples`
    systems, `every` mixed-form steps apart."""
    rng = np.random.default_rng(seed)
    ps = PairSet(capacity, loss)
    for sd in seeds:
        t0 = time.time()
        case = configs.buildings_tank(*size, nbuild=nbuild, seed=sd, eps_r=1.0e-8)
        with _context(case, capi.PRECISION_MIXED) as ctx:
            for s in range(samples):
                for _ in range(every):
                    ctx.step()
                feat, tgt, msk, bx = sample(ctx, mode)
                a = box_matrix(ctx) if loss == "residual" else None
                ps.add(feat, tgt, msk, crops_per_pair, rng, a)
                del feat, tgt, msk, a
        log(f"learned: layout seed {sd}: {ps.n} crops so far ({time.time() - t0:.1f} s, box {bx})")
    return ps


def train(ps, steps=3000, batch=8, lr=1.0e-3, lr_min=1.0e-5, seed=0, p0=None, log=print):
    """Adam on random batches of the crop set with the set's loss; cosine
    learning-rate decay.  Returns the parameters and the loss history (every
    50 steps)."""
    import torch
    rng = np.random.default_rng(seed)
    cx, cy, cz = CROP
    p = U.init_params(seed) if p0 is None else p0
    hist = []
    with U.UNet(cx, cy, cz, batch, train=True) as net:
        net.set_params(p)
        xb = torch.empty((batch, cz + 2, cy + 2, cx + 2, 8), dtype=torch.bfloat16, device="cuda")
        tb = torch.empty((batch, cz, cy, cx), dtype=torch.float32, device="cuda")
        mb = torch.empty((batch, cz, cy, cx), dtype=torch.float32, device="cuda")
        yb = torch.empty((batch, cz, cy, cx), dtype=torch.float32, device="cuda")
        t0 = time.time()
        run = []
        for s in range(steps):
            idx = torch.from_numpy(rng.integers(0, ps.n, size=batch)).cuda()
            xb.copy_(ps.x[idx])
            tb.copy_(ps.t[idx])
            mb.copy_(ps.m[idx])
            torch.cuda.synchronize()
            net.set_input(xb)
            a = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * s / steps))
            if ps.loss == "residual":
                net.forward(yb)
                y = yb.clone().requires_grad_(True)
                loss = residual_loss(ps.a[idx], ps.act[idx], tb, mb, y)
                loss.backward()
                torch.cuda.synchronize()
                net.train_step_dy(y.grad.contiguous(), a)
                run.append(float(loss))
            else:
                run.append(net.train_step(tb, mb, a))
            if (s + 1) % 50 == 0:
                hist.append(float(np.mean(run)))
                run = []
                if (s + 1) % 500 == 0:
                    log(f"learned: step {s + 1} loss {hist[-1]:.4f} ({time.time() - t0:.1f} s)")
        p = net.get_params()
    return p, hist


def evaluate(ps, p, batch=8):
    """Mean loss (the set's) of the network over a crop set (no training)."""
    import torch
    cx, cy, cz = CROP
    losses = []
    with U.UNet(cx, cy, cz, batch, train=True) as net:
        net.set_params(p)
        yb = torch.empty((batch, cz, cy, cx), dtype=torch.float32, device="cuda")
        for i in range(0, ps.n - batch + 1, batch):
            net.set_input(ps.x[i:i + batch].contiguous())
            if ps.loss == "residual":
                net.forward(yb)
                with torch.no_grad():
                    losses.append(float(residual_loss(ps.a[i:i + batch], ps.act[i:i + batch],
                                                      ps.t[i:i + batch], ps.m[i:i + batch], yb)))
            else:
                loss, _ = net.gradients(ps.t[i:i + batch].contiguous(), ps.m[i:i + batch].contiguous())
                losses.append(loss)
    return float(np.mean(losses)) if losses else float("nan")


def save(path, p, meta):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.save(path, np.ascontiguousarray(p, np.float32), allow_pickle=False)
    path.with_suffix(".json").write_text(json.dumps(meta, indent=1) + "\n")

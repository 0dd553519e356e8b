"""A plain float32 PyTorch statement of the U-Net of csrc/cad_unet.hip (CPU),
the reference its kernels are checked against: same layers, same parameter
vector, same loss."""
import numpy as np
import torch
import torch.nn.functional as F


def _w(p, sl, ci, co):
    # W[tap][ci][co], tap = kx + 3 ky + 9 kz  ->  torch [co][ci][kz][ky][kx]
    return torch.from_numpy(p[sl].reshape(3, 3, 3, ci, co)).permute(4, 3, 0, 1, 2).contiguous()


def forward(unet_mod, p, x):
    """x: float32 (nb, 8, nz, ny, nx) -> y (nb, nz, ny, nx); p: torch-free
    numpy parameters or a dict of torch leaf tensors (for autograd)."""
    sl = unet_mod.layer_slices()
    prm = {}
    for name in unet_mod.NAMES:
        w, b, (ci, co) = sl[name]
        prm[name] = (_w(p, w, ci, co), torch.from_numpy(p[b].copy()))
    w, b, _ = sl["out"]
    prm["out"] = (torch.from_numpy(p[w].copy()), torch.from_numpy(p[b].copy()))
    return prm


def run(prm, x):
    def c(name, h):
        w, b = prm[name]
        return torch.relu(F.conv3d(h, w, b, padding=1))
    h1a = c("e1a", x)
    s1 = c("e1b", h1a)
    p1 = F.avg_pool3d(s1, 2)
    h2a = c("e2a", p1)
    s2 = c("e2b", h2a)
    p2 = F.avg_pool3d(s2, 2)
    h3a = c("b3a", p2)
    h3b = c("b3b", h3a)
    h3c = c("u3", h3b)
    cat2 = torch.cat([F.interpolate(h3c, scale_factor=2, mode="nearest"), s2], dim=1)
    h2d = c("d2a", cat2)
    h2e = c("d2b", h2d)
    h2f = c("u2", h2e)
    cat1 = torch.cat([F.interpolate(h2f, scale_factor=2, mode="nearest"), s1], dim=1)
    h1d = c("d1a", cat1)
    h1e = c("d1b", h1d)
    w, b = prm["out"]
    return (h1e * w.view(1, -1, 1, 1, 1)).sum(1) + b


def loss(y, e, m):
    nb = y.shape[0]
    num = ((m * (y - e)) ** 2).reshape(nb, -1).sum(1)
    den = (m * e * e).reshape(nb, -1).sum(1)
    return (num / den).sum() / nb


def grad_vector(unet_mod, prm, n):
    g = np.zeros(n, np.float32)
    sl = unet_mod.layer_slices()
    for name in unet_mod.NAMES:
        w, b, (ci, co) = sl[name]
        gw = prm[name][0].grad.permute(2, 3, 4, 1, 0).contiguous().numpy().reshape(-1)
        g[w] = gw
        g[b] = prm[name][1].grad.numpy()
    w, b, _ = sl["out"]
    g[w] = prm["out"][0].grad.numpy()
    g[b] = prm["out"][1].grad.numpy()
    return g

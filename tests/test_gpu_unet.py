"""The U-Net kernels (BF16 on the matrix cores, FP32 accumulation) against a
plain float32 PyTorch statement of the same network on the CPU
(tests/unet_ref.py): the output, the loss and every layer's parameter
gradient, on random inputs and parameters (biases non-zero), for box sizes
that take every tiling of the convolution: one wave per x row with idle
16-voxel groups, two waves (the x shift crosses waves through LDS), eight
and sixteen waves, and output channels split over two launches (1024 voxels
along x).
The bar is BF16's: relative L2 error of the output within 3e-2, of each
layer's gradient within 1e-1 with cosine similarity above 0.99 -- a wrong tap,
a transposed weight block or a misplaced channel slice gives errors of order
one.  A few Adam steps on a fixed batch lower the loss."""
import numpy as np
import pytest

import unet_ref

pytestmark = pytest.mark.gpu


def _setup(pkg, box, nb, seed):
    import torch
    nx, ny, nz = box
    rng = np.random.default_rng(seed)
    U = pkg.unet
    p = U.init_params(seed)
    sl = U.layer_slices()
    for name in U.NAMES + ("out",):
        p[sl[name][1]] = rng.normal(0.0, 0.1, sl[name][1].stop - sl[name][1].start)
    x = rng.standard_normal((nb, 8, nz, ny, nx)).astype(np.float32)
    xb = torch.from_numpy(x).to(torch.bfloat16)
    xp = torch.zeros((nb, nz + 2, ny + 2, nx + 2, 8), dtype=torch.bfloat16)
    xp[:, 1:-1, 1:-1, 1:-1, :] = xb.permute(0, 2, 3, 4, 1)
    # the kernels' weights are BF16: the reference takes the same rounded values
    pr = p.copy()
    for name in U.NAMES + ("out",):
        w = sl[name][0]
        pr[w] = torch.from_numpy(p[w]).to(torch.bfloat16).float().numpy()
    e = rng.standard_normal((nb, nz, ny, nx)).astype(np.float32)
    m = (rng.random((nb, nz, ny, nx)) < 0.7).astype(np.float32)
    return p, pr, xb.float(), xp, e, m


@pytest.mark.parametrize("box,nb", [((64, 8, 8), 2), ((128, 4, 8), 1), ((512, 4, 4), 1),
                                    ((1024, 4, 4), 1)])
def test_unet_forward_and_gradients_match_torch(pkg, box, nb):
    import torch
    U = pkg.unet
    nx, ny, nz = box
    p, pr, xf, xp, e, m = _setup(pkg, box, nb, 3)
    print(f"unet box {box} batch {nb}", flush=True)
    with U.UNet(nx, ny, nz, nb, train=True) as net:
        net.set_params(p)
        net.set_input(xp.cuda())
        y = torch.empty(nb * nz * ny * nx, dtype=torch.float32, device="cuda")
        net.forward(y)
        y = y.cpu().numpy().reshape(nb, nz, ny, nx)
        loss, g = net.gradients(torch.from_numpy(e.ravel()).cuda(), torch.from_numpy(m.ravel()).cuda())
    # an inference network (which stores neither the coarse outputs of the
    # layers it upsamples nor the last hidden layer, only their fused
    # products) gives the same output bit for bit
    with U.UNet(nx, ny, nz, nb, train=False) as net:
        net.set_params(p)
        net.set_input(xp.cuda())
        yi = torch.empty(nb * nz * ny * nx, dtype=torch.float32, device="cuda")
        net.forward(yi)
        yi = yi.cpu().numpy().reshape(nb, nz, ny, nx)
    assert np.array_equal(yi, y)
    prm = unet_ref.forward(U, pr, xf)
    for v in prm.values():
        v[0].requires_grad_(True)
        v[1].requires_grad_(True)
    yr = unet_ref.run(prm, xf)
    ey = np.linalg.norm(y - yr.detach().numpy()) / np.linalg.norm(yr.detach().numpy())
    print(f"output rel L2 {ey:.3e}", flush=True)
    assert ey <= 3e-2
    lr = unet_ref.loss(yr, torch.from_numpy(e), torch.from_numpy(m))
    lr.backward()
    assert abs(loss - lr.item()) <= 3e-2 * lr.item(), (loss, lr.item())
    gr = unet_ref.grad_vector(U, prm, U.param_count())
    sl = U.layer_slices()
    for name in U.NAMES + ("out",):
        for part in (0, 1):
            a, b = g[sl[name][part]], gr[sl[name][part]]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
            print(f"{name} {'wb'[part]}: rel {rel:.3e} cos {cos:.5f}", flush=True)
            assert rel <= 1e-1 and cos >= 0.99, (name, part, rel, cos)


def test_unet_gradients_from_a_given_output_gradient_match_torch(pkg):
    """The backward pass from a caller's output gradient (the residual loss's
    path, cad_unet_gradients_dy): with dy random, every layer's gradient
    against the reference's gradient of sum(dy * y).  This checks the path of
    dy into the backward pass: a dy read in the wrong order or from the wrong
    buffer gives gradients unrelated to the reference's (cosine near 0).  An
    unmasked random dy sums to small gradients through much cancellation,
    which BF16 activations widen, so the bar is relative error 2e-1 and
    cosine 0.98 (measured worst: b3b's bias, 0.146 and 0.9896)."""
    import torch
    U = pkg.unet
    box, nb = (64, 8, 8), 2
    nx, ny, nz = box
    p, pr, xf, xp, e, m = _setup(pkg, box, nb, 5)
    dy = np.random.default_rng(9).standard_normal((nb, nz, ny, nx)).astype(np.float32)
    with U.UNet(nx, ny, nz, nb, train=True) as net:
        net.set_params(p)
        net.set_input(xp.cuda())
        y = torch.empty(nb * nz * ny * nx, dtype=torch.float32, device="cuda")
        net.forward(y)
        g = net.gradients_dy(torch.from_numpy(dy.ravel()).cuda())
    prm = unet_ref.forward(U, pr, xf)
    for v in prm.values():
        v[0].requires_grad_(True)
        v[1].requires_grad_(True)
    yr = unet_ref.run(prm, xf)
    (yr * torch.from_numpy(dy)).sum().backward()
    gr = unet_ref.grad_vector(U, prm, U.param_count())
    sl = U.layer_slices()
    for name in U.NAMES + ("out",):
        for part in (0, 1):
            a, b = g[sl[name][part]], gr[sl[name][part]]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
            print(f"{name} {'wb'[part]}: rel {rel:.3e} cos {cos:.5f}", flush=True)
            assert rel <= 2e-1 and cos >= 0.98, (name, part, rel, cos)


def test_unet_adam_steps_lower_the_loss(pkg):
    import torch
    U = pkg.unet
    box, nb = (64, 8, 8), 2
    p, _, _, xp, e, m = _setup(pkg, box, nb, 5)
    nx, ny, nz = box
    with U.UNet(nx, ny, nz, nb, train=True) as net:
        net.set_params(p)
        net.set_input(xp.cuda())
        t, mk = torch.from_numpy(e.ravel()).cuda(), torch.from_numpy(m.ravel()).cuda()
        losses = [net.train_step(t, mk, 1e-3) for _ in range(40)]
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.8 * losses[0]
    assert np.all(np.isfinite(losses))

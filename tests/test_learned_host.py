"""Host side of the learned form's residual loss: the matrix product on
crops (learned.apply_matrix) against the oracle's VF_MZAX (vf_mzax.f:42-56)
on the same rows, and the loss's two fixed points (1 for a zero correction,
0 for the exact one)."""
import numpy as np
import pytest

import pyoracle as po

torch = pytest.importorskip("torch")


def test_crop_matrix_product_is_mzax(pkg):
    rng = np.random.default_rng(4)
    nx, ny, nz = 7, 5, 6
    # a grid whose interior is the crop: ghosts hold x = 0, as beyond a crop
    shape = (nx + 2, ny + 2, nz + 2)
    mat = {k: np.asfortranarray(rng.standard_normal(shape)) for k in ("AD", "ALI", "ALJ", "ALK", "AUI", "AUJ", "AUK")}
    x = np.zeros(shape, order="F")
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal((nx, ny, nz))
    indc = np.zeros(shape, dtype=np.int32, order="F")
    want = po.mzax(mat, x, indc)[1:-1, 1:-1, 1:-1]
    # crop order (z, y, x) and the rows AD, ALI, AUI, ALJ, AUJ, ALK, AUK
    order = ("AD", "ALI", "AUI", "ALJ", "AUJ", "ALK", "AUK")
    a = torch.from_numpy(np.stack([mat[k][1:-1, 1:-1, 1:-1].transpose(2, 1, 0) for k in order]))[None]
    v = torch.from_numpy(np.ascontiguousarray(x[1:-1, 1:-1, 1:-1].transpose(2, 1, 0)))[None]
    got = pkg.learned.apply_matrix(a, v)[0].numpy().transpose(2, 1, 0)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_residual_loss_fixed_points(pkg):
    rng = np.random.default_rng(5)
    b, nz, ny, nx = 2, 6, 5, 7
    a = torch.from_numpy(rng.standard_normal((b, 7, nz, ny, nx)))
    act = torch.from_numpy((rng.random((b, nz, ny, nx)) < 0.8).astype(np.float64))
    e = torch.from_numpy(rng.standard_normal((b, nz, ny, nx))) * act
    m = act.clone()
    m[:, 0] = 0.0
    m[:, -1] = 0.0
    L = pkg.learned
    assert float(L.residual_loss(a, act, e, m, torch.zeros_like(e))) == pytest.approx(1.0)
    assert float(L.residual_loss(a, act, e, m, e)) == pytest.approx(0.0, abs=1e-24)

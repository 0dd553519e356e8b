"""cad_m1bcgs_noalloc, the solve on device-resident arrays that allocates
nothing, against the solve drop-in vf_m1bcgs_ on the same recorded Poisson
systems: the solution bit for bit, BNORM, XNORM and ITER equal.  Two calls in
a row, the second on a smaller grid and on another stream, with the work
arrays filled with NaN (the solve must clear what it reads), and the
refusals of sizes beyond the limit (no device call is made for those)."""
import numpy as np
import pytest

import fixtures as fx

WORK = ("DD", "XW", "YY", "R0", "RR", "PP", "AP", "AE")


def _params(case):
    st = fx.static(case)
    p = fx.deck_params(st["DECK"])
    p.update(fx.solver_params(st))
    return {k: p[k] for k in ("cgpara", "cgepsa", "cgepsr", "cgdiv", "icgtyp", "icgmax")}


def _flat(a):
    return np.asfortranarray(a).ravel(order="F")


@pytest.mark.gpu
def test_noalloc_solve_equals_the_dropin(pkg):
    import torch
    cap = pkg.capi
    calls = [("corner3d", fx.steps("corner3d")[-1], None), ("tiny", fx.steps("tiny")[-1], "side")]
    for case, s, stream_kind in calls:
        ps = fx.step(case, s)["ps_in"]
        p = _params(case)
        shape = ps["AD"].shape
        x0 = np.zeros(shape, order="F")
        print(f"vf_m1bcgs_ {case} step {s} {shape}", flush=True)
        xr, bnr, xnr, itr = cap.m1bcgs(ps, ps["INDC"], x0=x0, **p)
        dev = {k: torch.from_numpy(_flat(ps[k]).copy()).cuda() for k in cap.MATRIX}
        dev["INDC"] = torch.from_numpy(_flat(ps["INDC"]).astype(np.int32)).cuda()
        dev["X0"] = torch.from_numpy(_flat(x0).copy()).cuda()
        for k in WORK:
            dev[k] = torch.full((int(np.prod(shape)),), float("nan"), dtype=torch.float64,
                                device="cuda")
        torch.cuda.synchronize()
        side = torch.cuda.Stream() if stream_kind == "side" else None
        ptrs = {k: v.data_ptr() for k, v in dev.items()}
        print(f"cad_m1bcgs_noalloc {case} step {s} stream={stream_kind}", flush=True)
        rc, bn, xn, it = cap.m1bcgs_noalloc(ptrs, shape, stream=None if side is None else
                                            side.cuda_stream, **p)
        assert rc == 0, cap.lib().cad_last_error().decode()
        x = dev["X0"].cpu().numpy().reshape(shape, order="F")
        assert it == itr, (case, it, itr)
        assert bn == bnr and xn == xnr, (case, bn, bnr, xn, xnr)
        assert np.array_equal(x, xr), (case, np.abs(x - xr).max())
        assert it == int(fx.step(case, s)["ps_out"]["ICGITR"][0])


def test_noalloc_refuses_grids_beyond_the_limit(pkg):
    cap = pkg.capi
    ptrs = {k: 4096 for k in list(cap.MATRIX) + ["X0", "INDC"] + list(WORK)}
    n = cap.NOALLOC_MAX_CELLS
    rc, *_ = cap.m1bcgs_noalloc(ptrs, (256, 128, n // (256 * 128) + 1))
    assert rc == -1
    assert "CAD_NOALLOC_MAX_CELLS" in cap.lib().cad_last_error().decode()
    rc, *_ = cap.m1bcgs_noalloc(ptrs, (3, 8, 8))
    assert rc == -1
    bad = dict(ptrs, DD=0)
    rc, *_ = cap.m1bcgs_noalloc(bad, (8, 8, 8))
    assert rc == -1 and "null" in cap.lib().cad_last_error().decode()

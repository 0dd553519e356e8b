"""The set-up arrays of the full-size benchmark tanks, built from their decks by
layout.py, are bit-identical to the ones the reference built for the same
decks (its VF_CSETUP, VF_CGLV, VF_CGLXYZ, VF_CINDX, VF_CINDB): the sha256 of
each array was recorded from the reference's run (make_golden_full.py)."""
import ast

import numpy as np
import pytest

import fixtures as fx
from make_golden_full import sha


@pytest.mark.parametrize("name", ["runup_b", "buildings_c"])
def test_full_size_setup_arrays_are_the_references(pkg, name):
    p = fx.HERE / f"{name}_full_final.npz"
    if not p.exists():
        pytest.skip("no full-size record")
    with np.load(fx.HERE / f"{name}_full_trajectory.npz") as z:
        fac, kw, n = str(z["FACTORY"]), ast.literal_eval(str(z["KWARGS"])), int(z["NSTEPS"])
    case = getattr(pkg.configs, fac)(steps=n, name=name, **kw)
    a = pkg.layout.from_text(case.deck).static_arrays()
    with np.load(p) as z:
        want = {k[4:]: str(z[k]) for k in z.files if k.startswith("SHA_")}
    assert len(want) >= 17
    bad = [k for k, h in want.items() if sha(a[k]) != h]
    assert not bad, bad

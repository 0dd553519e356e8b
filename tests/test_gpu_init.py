"""Device part of VF_CINIT (cad_init_state) against the reference's initial state.

Deck -> host layout (layout.py) -> upload -> cad_init_state must reproduce the
state the reference hands to its first VF_V1CAL, bit for bit.
"""
import numpy as np
import pytest

import fixtures as fx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["tiny", "beach_small", "corner3d", "dam_a"])
def test_gpu_init_matches_reference(pkg, case):
    st = fx.static(case)
    lay = pkg.layout.from_text(str(st["DECK"]))
    ni, nj, nk = lay.shape
    with pkg.capi.Context(ni, nj, nk, lay.numb) as c:
        p = lay.params()
        c.set_params(**{k: v for k, v in p.items() if k in c.params})
        c.upload_static(lay.static_arrays())
        c.upload_state(lay.raw_state())
        info = c.init_state(lay.deck.wvlvl, lay.deck.uini)
        out = c.download_state()
    r1 = fx.step(case, 1)
    ref = r1["v1_in"]
    for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "NF", "INDC"):
        assert np.array_equal(out[k], ref[k]), k
    assert np.array_equal(out["BCF"], r1["f1_in"]["BCF"])
    assert info["nums"] == int(ref["NUMS"][0])
    assert abs(info["fsum"] - ref["FSUM"][0]) <= 1e-12 * ref["FSUM"][0]

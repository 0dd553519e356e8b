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


def _random_f_state(lay, seed):
    """The corner3d layout with F drawn at random on every non-obstacle
    cell: 0 with probability 0.6, a fraction with 0.1, else 1.  A state full
    of surface cells with no fluid/gas pair across them (348 of 1322 on the
    first pass for seed 3), which VF_FNFINI's fix-up passes have to resolve."""
    rng = np.random.default_rng(seed)
    state = lay.raw_state()
    u = rng.random(lay.shape)
    f = np.where(u < 0.6, 0.0, np.where(u < 0.7, rng.random(lay.shape), 1.0))
    f[lay.NF0 == -1] = 0.0
    f[0, :, :] = f[-1, :, :] = 0.0
    f[:, 0, :] = f[:, -1, :] = 0.0
    f[:, :, 0] = f[:, :, -1] = 0.0
    state["FF"][...] = f
    return state


def test_flag_passes_on_a_random_state_match_the_oracle(pkg):
    """VF_FNFINI with many fix-up passes (random F): the device's flags and
    surface count equal the oracle's; with the pass cap at 1 the same state
    is refused (CAD_ERR_STATE) instead of going on with unresolved flags."""
    import pyoracle as po
    case = pkg.configs.corner3d(name="rnd")
    lay = pkg.layout.from_text(case.deck)
    ni, nj, nk = lay.shape
    state = _random_f_state(lay, 3)
    arr = dict(lay.static_arrays())
    arr.update({k: v.copy() for k, v in state.items()})
    F = po.Fields(arr, lay.params())
    assert F.init_state(lay.deck.wvlvl, lay.deck.uini, 0) == 0
    with pkg.capi.Context(ni, nj, nk, lay.numb) as c:
        p = lay.params()
        c.set_params(**{k: v for k, v in p.items() if k in c.params})
        c.upload_static(lay.static_arrays())
        c.upload_state(state)
        info = c.init_state(lay.deck.wvlvl, lay.deck.uini, 0)
        out = c.download_state()
        assert np.array_equal(out["NF"], F["NF"])
        assert info["nums"] == F.st.nums
        c.upload_state(state)
        c.set_flag_pass_limit(1)
        with pytest.raises(pkg.capi.CadError) as e:
            c.init_state(lay.deck.wvlvl, lay.deck.uini, 0)
        assert e.value.code == -3 and "flag passes" in str(e.value)

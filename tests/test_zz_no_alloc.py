"""sphe_mass_all_noalloc (the all-scene bookkeeping from the library's static
scratch, no device allocation) against sphe_mass_all, its allocating form:
bit for bit on two calls in a row on one GPU, the second with fewer scenes,
and the limits it reports instead of allocating."""
import numpy as np
import pytest

import sphe
from sphe import scenes

pytestmark = pytest.mark.gpu


def same(a, b):
    return all(np.array_equal(x.view(np.uint64), y.view(np.uint64)) for x, y in zip(a, b))


def test_noalloc_equals_mass_all_on_two_calls():
    big = sphe.Context()
    for s in scenes.cfg3(count=96):
        big.add_synthetic(s)
    big.step(40)
    small = sphe.Context()
    for k in range(5):
        small.add_synthetic(scenes.cfg1(seed_idx=k, n=800 + 500 * k))
    small.step(12)
    # first call: 96 scenes; second call, on the same GPU's scratch: 5 scenes
    got_big = big.mass_all_noalloc()
    got_small = small.mass_all_noalloc()
    assert len(got_big[0]) == 96 and len(got_small[0]) == 5
    assert same(got_big, big.mass_all())
    assert same(got_small, small.mass_all())
    assert got_big[1].sum() > 0 and got_small[1].sum() > 0
    # the scratch kept the second call's contents; the first context's
    # figures are recomputed in full
    assert same(big.mass_all_noalloc(), got_big)


def test_noalloc_limits():
    fresh = sphe.Context()
    fresh.add_synthetic(scenes.cfg1(n=500))
    with pytest.raises(sphe.SpheError, match="ESTATE"):
        fresh.mass_all_noalloc()  # no device state yet: building it would allocate
    fresh.step(1)
    t, s, u = fresh.mass_all_noalloc()
    assert same((t, s, u), fresh.mass_all())
    many = sphe.Context()
    p = sphe.default_params(dt=1e-3)
    for _ in range(4097):
        many.add_scene(p)
    many.step(1)
    with pytest.raises(sphe.SpheError, match="at most 4096 scenes"):
        many.mass_all_noalloc()

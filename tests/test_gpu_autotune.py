"""Autotuner (hc_autotune, hc_set_tuning): every launch shape it can choose
gives bit-identical results, and tuning leaves the state untouched (needs
an MI355X)."""
import numpy as np
import pytest

from hercules_amd import Channel
from hercules_amd.channel import HerculesError
from tests.helpers import small_case

pytestmark = pytest.mark.gpu

SHAPES = [dict(rgroup=1), dict(rgroup=2), dict(ytile=8), dict(ytile=16), dict(c2r_pts=512),
          dict(c2r_pts=2048), dict(col_remap=1), dict(rgroup=4, ytile=16, c2r_pts=2048, col_remap=1)]


def _run(prm, st, n, shape=None, tune=False):
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    chosen = None
    if shape:
        ch.set_tuning(**shape)
    if tune:
        before = {k: ch.get_field(k) for k in st}
        chosen = ch.autotune(steps_per_trial=1)
        for k, v in before.items():
            assert np.array_equal(ch.get_field(k), v), k
        assert ch.scalar("stps") == 0
    ch.steps(n)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    out = {k: ch.get_field(k) for k in names}
    ch.close()
    return out, chosen


@pytest.mark.parametrize("kind,kw", [("closed", dict(nx=64, ny=64, nz=65)),
                                     ("open_t", dict(nx=128, ny=32, nz=33)),
                                     ("open_t", dict(nx=64, ny=64, nz=33, precision=1))])
def test_every_launch_shape_is_bit_identical(kind, kw):
    prm, st = small_case(kind, **kw)
    ref, _ = _run(prm, st, 3)
    for shape in SHAPES:
        got, _ = _run(prm, st, 3, shape=shape)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (shape, k)
    tuned, chosen = _run(prm, st, 3, tune=True)
    print(kind, kw, "chosen", chosen)
    for k in ref:
        assert np.array_equal(tuned[k], ref[k]), (chosen, k)


def test_bad_launch_shapes_are_rejected():
    prm, _ = small_case("closed", nx=64, ny=64, nz=65)
    ch = Channel(prm)
    for shape in (dict(rgroup=3), dict(ytile=12), dict(c2r_pts=4096), dict(col_remap=2)):
        with pytest.raises(HerculesError, match="status 1"):
            ch.set_tuning(**shape)
    ch.close()

"""Autotuner (hc_autotune, hc_set_tuning): every launch shape it can choose
gives bit-identical results, except the x-y tiled momentum right-hand side
(rhs_xt 1-4), whose separate x transform and stencil kernel round
differently from the fused whole-row kernel: those shapes agree with each
other bit for bit and with the fused kernel to round-off.  Tuning leaves the
state untouched (needs an MI355X)."""
import numpy as np
import pytest

from hercules_amd import Channel
from hercules_amd.channel import HerculesError
from tests.helpers import small_case

pytestmark = pytest.mark.gpu

SHAPES = [dict(rgroup=1), dict(rgroup=2), dict(ytile=8), dict(ytile=16), dict(c2r_pts=512),
          dict(c2r_pts=2048), dict(col_remap=1), dict(rgroup=4, ytile=16, c2r_pts=2048, col_remap=1),
          dict(rhs_xt=1), dict(rhs_xt=2), dict(rhs_xt=3), dict(rhs_xt=4), dict(rhs_xt=2, rgroup=4, ytile=16)]


def _same(got, ref, shape, xt_ref, tol):
    """bit-identical, or for a tiled momentum shape equal to the first tiled
    run and within round-off (tol, relative to the field's maximum) of the
    fused kernel"""
    for k in ref:
        if shape.get("rhs_xt", 0):
            scale = np.abs(ref[k]).max()
            assert np.abs(got[k] - ref[k]).max() <= tol[k == "p"] * scale, (shape, k)
            if xt_ref is not None:
                assert np.array_equal(got[k], xt_ref[k]), (shape, k)
        else:
            assert np.array_equal(got[k], ref[k]), (shape, k)


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
                                     ("open_t", dict(nx=64, ny=64, nz=33, precision=1)),
                                     ("open_t", dict(nx=64, ny=64, nz=33, isdamp=1, nzdamp=8, cdamp=0.2))])
def test_every_launch_shape_is_bit_identical(kind, kw):
    prm, st = small_case(kind, **kw)
    ref, _ = _run(prm, st, 3)
    # (other fields, p): p of the FP32 variant, the solve of a Neumann problem
    # from FP32 sources, moves by ~1e-4 under a round-off perturbation, as
    # it differs from FP64 by that much
    tol = (1e-5, 1e-3) if prm.precision == 1 else (1e-12, 1e-9)
    xt_ref = None
    for shape in SHAPES:
        got, _ = _run(prm, st, 3, shape=shape)
        _same(got, ref, shape, xt_ref, tol)
        if shape.get("rhs_xt", 0) and xt_ref is None:
            xt_ref = got
    tuned, chosen = _run(prm, st, 3, tune=True)
    print(kind, kw, "chosen", chosen)
    _same(tuned, ref, chosen, xt_ref, tol)


def test_bad_launch_shapes_are_rejected():
    prm, _ = small_case("closed", nx=64, ny=64, nz=65)
    ch = Channel(prm)
    for shape in (dict(rgroup=3), dict(ytile=12), dict(c2r_pts=4096), dict(col_remap=2), dict(rhs_xt=5)):
        with pytest.raises(HerculesError, match="status 1"):
            ch.set_tuning(**shape)
    ch.close()

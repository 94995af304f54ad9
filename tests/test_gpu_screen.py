"""GPU divergence screen (screen_div_error, src/module_io.f90:85-206) against
the oracle: mean |div|, the reference's signed running div_max and the
div_hist_max kept across calls; the per-substep screen of the step and its
NaN stop (src/module_navier_stokes.f90:344-350)."""
import numpy as np
import pytest

from hercules_amd import Channel, cases
from hercules_amd.channel import HerculesError
from tests.helpers import make_oracle
from tests.test_oracle_screen import random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cds,shape", [(1, (32, 32, 33)), (2, (32, 32, 33)), (2, (64, 16, 20)), (2, (16, 64, 9))])
def test_signed_screen_equals_oracle(cds, shape):
    nx, ny, nz = shape
    prm = cases.closed_channel(nx=nx, ny=ny, nz=nz, cds=cds)
    st = random_state(prm, 11 + nx + nz)
    o = make_oracle(prm, st)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    for scale, fin in ((1.0, 1), (0.7, 1), (3.0, 0)):
        for k in ("u", "v", "w"):
            ch.set_field(k, scale * st[k])
            o.set(k, scale * st[k])
        g = ch.screen_div_error_ref(fin)
        r = o.screen(fin)
        print("screen", cds, shape, fin, g, r)
        assert g[0] == pytest.approx(r[0], rel=1e-12)
        assert g[1] == pytest.approx(r[1], rel=1e-12)
        assert g[2] == pytest.approx(r[2], rel=1e-12)
    ch.close()


def test_substep_screen_records_and_hist():
    prm = cases.closed_channel(nx=32, ny=32, nz=33)
    st = cases.laminar_noise(prm, seed=5)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.project()
    ch.init()
    ch.screen_substeps(True)
    ch.time_advancement_fd()
    recs = [ch.get_screen(s) for s in (1, 2, 3)]
    # the step's last record is the screen of the final fields
    again = ch.screen_div_error_ref(0)
    assert again[0] == recs[2][0] and again[1] == recs[2][1]
    # the history is the signed rule applied over the substeps
    hist = recs[0][2]
    assert abs(hist) >= abs(recs[0][1]) or hist == recs[0][1]
    for a, m, h in recs:
        assert np.isfinite(a) and abs(m) < 1e-9 and abs(h) < 1e-9
    u0 = ch.get_field("u")
    ch.screen_substeps(False)
    ch.close()
    # the screen does not change the step
    ch2 = Channel(prm)
    for k, v in st.items():
        ch2.set_field(k, v)
    ch2.project()
    ch2.init()
    ch2.time_advancement_fd()
    assert np.array_equal(ch2.get_field("u"), u0)
    ch2.close()


@pytest.mark.parametrize("entry", ["time_advancement_fd", "steps"])
def test_nan_stops_the_step(entry):
    prm = cases.closed_channel(nx=32, ny=32, nz=33)
    st = cases.laminar_noise(prm, seed=6)
    st["u"][4, 3, 2] = np.nan
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    ch.screen_substeps(True)
    with pytest.raises(HerculesError) as ei:
        if entry == "steps":
            ch.steps(3)
        else:
            ch.time_advancement_fd()
    assert ei.value.code == 4 and "NaN" in str(ei.value)
    if entry == "steps":
        assert ch.scalar("stps") == 1  # stopped after the first step
    ch.close()

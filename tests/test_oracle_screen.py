"""CPU pin of the oracle's screen_div_error (src/module_io.f90:85-206): the
reference keeps a SIGNED running maximum, replaced at each point (scan order
k, j, i) whose |d| exceeds the stored value, and a history maximum updated
by the same rule on every isfinal = 1 call.  The rule is run here as a plain
Python loop over the divergence formed independently with numpy stencils."""
import numpy as np
import pytest

from hercules_amd import cases
from tests.helpers import divergence, make_oracle


def reference_rule(d, hist0, isfinal):
    """screen_div_error's loop: returns (div_avg, div_max, div_hist_max)."""
    mx, hist, s = 0.0, hist0, 0.0
    for x in d.ravel():  # C order (k, j, i): i fastest, as the Fortran loop nest
        s += abs(x)
        if abs(x) > mx:
            mx = x
        if isfinal and abs(x) > hist:
            hist = x
    return s / d.size, mx, hist


def random_state(prm, seed):
    rng = np.random.default_rng(seed)
    shape = (prm.nz, prm.ny, prm.nx)
    u, v, w = (rng.standard_normal(shape) for _ in range(3))
    w[-1] = 0.0  # update_boundary_uvw: w(kend3) = 0
    return {"u": u, "v": v, "w": w, "p": np.zeros(shape), "t": np.zeros(shape)}


@pytest.mark.parametrize("cds", [1, 2])
def test_signed_running_maximum_and_history(cds):
    prm = cases.closed_channel(nx=16, ny=8, nz=9, cds=cds)
    st = random_state(prm, 3 + cds)
    o = make_oracle(prm, st)
    d, _ = divergence(st["u"], st["v"], st["w"], prm)
    avg, mx, hist = reference_rule(d, 0.0, 1)
    got = o.screen(1)
    assert got[0] == pytest.approx(avg, rel=1e-13)
    assert got[1] == pytest.approx(mx, rel=1e-12)
    assert got[2] == pytest.approx(hist, rel=1e-12)
    # the rule is order dependent: the stored value is not max |d| here
    assert abs(mx) < np.abs(d).max()
    # a second, smaller field: div_max restarts from 0, the history carries
    st2 = {k: 0.5 * v for k, v in st.items()}
    for k in ("u", "v", "w"):
        o.set(k, st2[k])
    d2, _ = divergence(st2["u"], st2["v"], st2["w"], prm)
    avg2, mx2, hist2 = reference_rule(d2, hist, 1)
    got2 = o.screen(1)
    assert got2[1] == pytest.approx(mx2, rel=1e-12)
    assert got2[2] == pytest.approx(hist2, rel=1e-12)
    # isfinal = 0 leaves the history alone
    got3 = o.screen(0)
    assert got3[2] == pytest.approx(hist2, rel=1e-12)

"""Mean statistics of the oracle: calculate_mean_fields (src/module_tools.f90:
362-404) and calculate_mean_fields_spec (src/module_spectral.f90:1012-1052),
pinned by the explicit sums of the fields after each step (CPU)."""
import numpy as np
import pytest

from tests.helpers import make_oracle, small_case


def _sums(o, prm, nsteps):
    s = {k: 0.0 for k in ("u", "v", "w", "t", "uu", "vv", "ww", "uw", "vw", "tt", "tw")}
    um = prm.u_mrf
    for _ in range(nsteps):
        o.steps(1)
        u, v, w, t = (o.get(k) for k in "uvwt")
        wc = 0.5 * (w + np.concatenate([np.zeros_like(w[:1]), w[:-1]], 0))
        if prm.cds == 3:
            uh, vh = u, v
        else:  # the interpolants left by the step's last substep
            g = o.get_ghosted("uhx")[2:-2, 2:-2, 2:-2]
            h = o.get_ghosted("vhy")[2:-2, 2:-2, 2:-2]
            uh, vh = g, h
        s["u"] += u + um
        s["v"] += v
        s["w"] += w
        s["t"] += t
        s["uu"] += (u + um) ** 2
        s["vv"] += v ** 2
        s["ww"] += w ** 2
        s["uw"] += wc * (uh + um)
        s["vw"] += wc * vh
        s["tt"] += t ** 2
        s["tw"] += t * wc
    return {k: v / nsteps for k, v in s.items()}


@pytest.mark.parametrize("kind,kw", [
    ("closed", {}),
    ("closed", {"cds": 1, "u_mrf": 0.3}),
    ("open_t", {}),
    ("open_t", {"cds": 3, "issk": 1}),
    ("closed", {"dts": 1}),
])
def test_oracle_means_are_the_step_sums(kind, kw):
    prm, st = small_case(kind, nx=16, ny=16, nz=17, **kw)
    o = make_oracle(prm, st)
    o.mean_enable()
    ref = _sums(o, prm, 3)
    for k, v in ref.items():
        got = o.get_mean(k)
        scale = max(np.abs(v).max(), 1e-300)
        assert np.abs(got - v).max() <= 1e-13 * scale, k


def test_oracle_fd_uw_takes_the_last_substep_interpolant():
    """The FD uw sum uses uhx from the start of the last substep, not the
    updated u: after one step of a developing flow the two differ."""
    prm, st = small_case("closed", nx=16, ny=16, nz=17, cds=1)
    o = make_oracle(prm, st)
    o.mean_enable()
    o.steps(1)
    u, w = o.get("u"), o.get("w")
    wc = 0.5 * (w + np.concatenate([np.zeros_like(w[:1]), w[:-1]], 0))
    fresh = wc * 0.5 * (u + np.roll(u, 1, axis=2))
    stale = wc * o.get_ghosted("uhx")[2:-2, 2:-2, 2:-2]
    got = o.get_mean("uw")
    assert np.abs(got - stale).max() < 1e-14 * np.abs(stale).max()
    assert np.abs(got - fresh).max() > 1e-6 * np.abs(stale).max()

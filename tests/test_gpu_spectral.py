"""Spectral path (cds=3, time_advancement_sp) on the GPU vs the CPU oracle,
through the C ABI (needs an MI355X)."""
import numpy as np
import pytest

from hercules_amd import Channel, cases
from tests.helpers import make_oracle, rel_l2, small_case

pytestmark = pytest.mark.gpu


def _pair(prm, st):
    o = make_oracle(prm, st)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    return o, ch


@pytest.mark.parametrize("kind,kw", [
    ("closed", {"issk": 1}),
    ("closed", {"issk": 0}),
    ("closed", {"issk": 1, "dp_opt": 2, "u_mrf": 0.3}),
    ("closed", {"issk": 1, "dp_opt": 3, "fc": 0.5, "ug": 1.0, "vg": 0.2}),
    ("closed", {"issk": 1, "dts": 1}),
    ("open_t", {"issk": 1}),
    ("open_t", {"issk": 0, "is_ri_var": 1, "ri_str": 0.05, "ri_end": 0.2, "imax": 10}),
    ("open_t", {"issk": 1, "tbc_top": 1, "dtdz_top": 0.5}),
])
def test_spectral_steps_match_oracle(kind, kw):
    prm, st = small_case(kind, cds=3, **kw)
    o, ch = _pair(prm, st)
    o.steps(3)
    ch.steps(3)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    for n in names:
        err = rel_l2(ch.get_field(n), o.get(n))
        print(f"spectral {kind} {kw} {n}: rel L2 {err:.3e}")
        assert err < 1e-10, (n, err)
    ga, gm = ch.screen_div_error()
    oa, om = o.div_stats()
    print(f"   div gpu {ga:.3e} {gm:.3e} oracle {oa:.3e} {om:.3e}")
    scale = np.abs(o.get("u")).max() / (prm.lz / prm.nz)
    assert gm < 1e-12 * scale and om < 1e-12 * scale
    ch.close()


@pytest.mark.parametrize("shape", [(16, 16, 17), (64, 32, 65), (192, 64, 33), (32, 32, 129), (16, 16, 257)])
def test_spectral_shapes_match_oracle(shape):
    nx, ny, nz = shape
    prm, st = small_case("closed", nx=nx, ny=ny, nz=nz, cds=3, issk=1)
    o, ch = _pair(prm, st)
    o.steps(2)
    ch.steps(2)
    for n in ("u", "v", "w", "p"):
        err = rel_l2(ch.get_field(n), o.get(n))
        print(f"spectral {shape} {n}: rel L2 {err:.3e}")
        assert err < 1e-10, (n, err)
    ch.close()


def test_spectral_start_up_truncates():
    """hc_init forms the spectra, truncates u, v, w, p, t by the 2/3 rule and
    returns the truncated fields (src/module_navier_stokes.f90:96-123)."""
    prm, st = small_case("open_t", cds=3, issk=1)
    o, ch = _pair(prm, st)
    for n in ("u", "v", "w", "p", "t"):
        err = rel_l2(ch.get_field(n), o.get(n))
        assert err < 1e-13, (n, err)
    ch.close()


@pytest.mark.parametrize("eq", ["u", "v", "w", "p", "t"])
def test_spectral_coefficient_solve_matches_oracle(eq):
    """poisson_solver_fft with the spectral (cds=3) coefficients."""
    prm = cases.closed_channel(nx=32, ny=32, nz=33, cds=3)
    rng = np.random.default_rng(5)
    src = rng.standard_normal((prm.nz, prm.ny, prm.nx))
    if eq == "p":
        src -= src.mean(axis=(1, 2), keepdims=True)
    ref = make_oracle(prm, {}).solve(eq, 8.0 / 15.0, src)
    with Channel(prm) as ch:
        got = ch.poisson_solver_fft(eq, 8.0 / 15.0, src)
    err = rel_l2(got, ref)
    print(f"spectral solve {eq}: rel L2 {err:.3e}")
    assert err < (1e-9 if eq == "p" else 1e-13)


@pytest.mark.parametrize("kind,kw", [
    ("closed", {"issk": 1, "isdamp": 1, "nzdamp": 6, "cdamp": 0.2}),
    ("open_t", {"issk": 0, "isdamp": 1, "nzdamp": 8, "cdamp": 0.1, "dp_opt": 3, "fc": 0.5, "ug": 1.0, "vg": 0.0}),
])
def test_spectral_damping_matches_oracle(kind, kw):
    """rayleigh_damping_spec of c_u, c_v, c_w, c_p, c_t at every substep
    (src/module_navier_stokes.f90:406-414), with the whole complex mean."""
    prm, st = small_case(kind, cds=3, **kw)
    o, ch = _pair(prm, st)
    o.steps(3)
    ch.steps(3)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    for n in names:
        err = rel_l2(ch.get_field(n), o.get(n))
        print(f"spectral damped {kind} {n}: rel L2 {err:.3e}")
        assert err < 1e-10, (n, err)
    ch.close()

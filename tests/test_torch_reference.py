"""The torch-ops restatement of the FD step (tests/torch_reference.py)
against the C oracle: the same fields after the same steps.  This pins the
reference whose autograd gradient checks the hand-written adjoint."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hercules_amd import cases  # noqa: E402
from tests.helpers import make_oracle, rel_l2  # noqa: E402
from tests.torch_reference import TorchChannel  # noqa: E402


def _case(kind, **kw):
    if kind == "closed":
        prm = cases.closed_channel(nx=16, ny=16, nz=17, wall_rows=1, **kw)
        st = cases.laminar_noise(prm, seed=4)
    else:
        prm = cases.stratified_open_channel(nx=16, ny=16, nz=17, wall_rows=1, isscalar=0, got=0.0,
                                            nu=1.0 / 180.0, **kw)
        st = cases.loglaw_noise(prm, seed=4)
    st.pop("t", None)
    return prm, st


@pytest.mark.parametrize("kind,kw", [("closed", {}), ("closed", {"cds": 1}), ("open", {})])
def test_torch_step_matches_oracle(kind, kw):
    prm, st = _case(kind, **kw)
    o = make_oracle(prm, st)
    o.steps(2)
    tc = TorchChannel(prm)
    out = tc.steps({k: torch.tensor(v) for k, v in st.items()}, 2)
    for n in ("u", "v", "w", "p"):
        err = rel_l2(out[n].numpy(), o.get(n))
        print(kind, kw, n, f"{err:.2e}")
        assert err < 1e-11, (n, err)


@pytest.mark.parametrize("kw", [{}, {"tbc_top": 1, "dtdz_top": 0.5}, {"cds": 1, "ochannel": 0}])
def test_torch_step_with_temperature_matches_oracle(kw):
    prm = cases.stratified_open_channel(nx=16, ny=16, nz=17, wall_rows=1, nu=1.0 / 180.0, **kw)
    st = cases.loglaw_noise(prm, seed=6)
    o = make_oracle(prm, st)
    o.steps(2)
    tc = TorchChannel(prm)
    out = tc.steps({k: torch.tensor(v) for k, v in st.items()}, 2)
    for n in ("u", "v", "w", "p", "t"):
        err = rel_l2(out[n].numpy(), o.get(n))
        print(kw, n, f"{err:.2e}")
        assert err < 1e-11, (n, err)


@pytest.mark.parametrize("kind,kw", [
    ("closed", {"issk": 1}),
    ("closed", {"issk": 0, "dp_opt": 3, "fc": 0.5, "ug": 1.0, "vg": 0.2}),
    ("open", {"issk": 1}),
    ("open", {"issk": 1, "tbc_top": 1, "dtdz_top": 0.5}),
    ("open", {"issk": 0}),
])
def test_torch_spectral_step_matches_oracle(kind, kw):
    """The torch restatement of the spectral substep (tests/torch_spectral.py)
    against the C oracle's time_advancement_sp: the same fields after two steps."""
    from tests.torch_spectral import TorchSpectral
    if kind == "closed":
        prm = cases.closed_channel(nx=16, ny=16, nz=17, wall_rows=1, cds=3, **kw)
        st = cases.laminar_noise(prm, seed=4)
        st.pop("t", None)
    else:
        prm = cases.stratified_open_channel(nx=16, ny=16, nz=17, wall_rows=1, cds=3, nu=1.0 / 180.0, **kw)
        st = cases.loglaw_noise(prm, seed=4)
    o = make_oracle(prm, st)
    o.steps(2)
    ref = TorchSpectral(prm).steps(st, 2)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    for n in names:
        err = rel_l2(ref[n].numpy(), o.get(n))
        print(kind, kw, n, f"{err:.2e}")
        assert err < 1e-11, (n, err)

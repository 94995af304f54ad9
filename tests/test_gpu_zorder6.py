"""6th-order z operators (hc_params.zorder = 6) on the GPU, through the C
ABI, against the FP64 torch-ops restatement of the step with the dense
operators of tests/zops6.py (tests/torch_reference.py; those operators are
pinned on the CPU by tests/test_zorder6.py).  The reference is 2nd order in
z, so there is no oracle for this extension; needs an MI355X."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hercules_amd import Channel, cases  # noqa: E402
from hercules_amd.channel import HerculesError  # noqa: E402
from tests.helpers import rel_l2  # noqa: E402
from tests.torch_reference import TorchChannel  # noqa: E402

pytestmark = pytest.mark.gpu


def _case(kind, nz, zorder=6, **kw):
    if kind == "closed":
        prm = cases.closed_channel(nx=16, ny=16, nz=nz, wall_rows=1, zorder=zorder, **kw)
        st = cases.laminar_noise(prm, seed=5)
        st.pop("t", None)
    else:
        prm = cases.stratified_open_channel(nx=16, ny=16, nz=nz, wall_rows=1, zorder=zorder, nu=1.0 / 180.0, **kw)
        st = cases.loglaw_noise(prm, seed=5)
        if not prm.isscalar:
            st.pop("t", None)
    return prm, st


CASES = [
    ("closed", 33, {}),
    ("closed", 65, {"cds": 1}),
    ("closed", 129, {}),
    ("open", 33, {"isscalar": 0, "got": 0.0}),
    ("open", 65, {}),
    ("open", 33, {"tbc_top": 1, "dtdz_top": 0.5}),
    ("open", 129, {"cds": 1, "tbc_top": 1, "dtdz_top": 1.0}),
    # nz = 257: five levels per lane in the column kernels
    ("closed", 257, {}),
    # 4th-order z operators (zorder = 4): 5- and 4-point stencils
    ("closed", 65, {"zorder": 4}),
    ("open", 65, {"zorder": 4}),
    ("open", 33, {"zorder": 4, "cds": 1, "tbc_top": 1, "dtdz_top": 0.5}),
]


@pytest.mark.parametrize("kind,nz,kw", CASES)
def test_z6_steps_match_torch_reference(kind, nz, kw):
    prm, st = _case(kind, nz, **kw)
    tc = TorchChannel(prm)
    ref = tc.steps({k: torch.tensor(v) for k, v in st.items()}, 2)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    ch.steps(2)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    # p: its (0,0) mode solves the Neumann pressure operator with the delnull
    # row, whose condition number grows as 1/dz_min^2 (1.4e6 at nz = 129):
    # two exact solvers agree to ~1e-16 times that, relative to p
    tol = {"p": 1e-9}
    for n in names:
        err = rel_l2(ch.get_field(n), ref[n].numpy())
        print(kind, nz, kw, n, f"{err:.2e}")
        assert err < tol.get(n, 1e-12), (n, err)
    # divergence with the 6th-order z divergence: round-off of the projection
    _, dmax = ch.screen_div_error()
    scale = np.abs(ref["u"].numpy()).max() / float(tc.dz.min())
    print("   max |div|", dmax, "scale", scale)
    assert dmax < 1e-12 * scale
    ch.close()


@pytest.mark.parametrize("eq", ["u", "v", "w", "p", "t"])
@pytest.mark.parametrize("nz,zorder", [(33, 6), (257, 6), (65, 4)])
def test_z6_standalone_solve_matches_dense(eq, nz, zorder):
    """poisson_solver_fft with the 6th-order operators against the dense
    per-mode solve: (I - beta (lap_xy + Z)) x = q (+ wall data of u, t in the
    (0,0) mode), rc (lap_xy + D G) phi = s with phi(0) = 0 in the (0,0) mode."""
    prm = cases.stratified_open_channel(nx=16, ny=32, nz=nz, wall_rows=1, zorder=zorder, nu=1.0 / 180.0,
                                        tbc_top=1 if eq == "t" else 0, dtdz_top=0.7)
    rng = np.random.default_rng(9)
    src = rng.standard_normal((prm.nz, prm.ny, prm.nx))
    rk_c = 2.0 / 15.0
    ch = Channel(prm)
    got = ch.poisson_solver_fft(eq, rk_c, src)
    ch.close()
    tc = TorchChannel(prm)
    s = torch.tensor(src)
    if eq == "p":
        ref = tc.poisson(s / rk_c)
    elif eq == "t":
        beta = rk_c * 0.5 * prm.dt * prm.alpha
        ref = tc.helm(s + beta * tc.col(tc.bt), tc.Lt, beta)
    else:
        beta = rk_c * 0.5 * prm.dt * prm.nu
        ref = tc.helm(s, tc.Lw if eq == "w" else tc.Lu, beta, trivial_top=eq == "w")
    err = rel_l2(got, ref.numpy())
    print(eq, nz, f"{err:.2e}")
    # p: the conditioning of the (0,0)-mode pressure operator (6e6 at nz = 257)
    assert err < (1e-9 if eq == "p" else 1e-11), err


@pytest.mark.parametrize("zorder", [6, 4])
def test_z6_projection_is_divergence_free(zorder):
    prm, st = _case("closed", 129, zorder=zorder)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    before = ch.screen_div_error()[1]
    ch.project()
    after = ch.screen_div_error()[1]
    u = ch.get_field("u")
    scale = np.abs(u).max() / (prm.lz / prm.nz)
    print("div before", before, "after", after)
    assert before > 1e-6 * scale and after < 1e-12 * scale
    ch.close()


@pytest.mark.parametrize("kw", [dict(isdamp=1, nzdamp=8, cdamp=0.1), dict(cds=3, isdamp=1, nzdamp=8, cdamp=0.1),
                                dict(nz=9), dict(nz=1025), dict(zorder=3)])
def test_z6_rejects_unsupported(kw):
    base = dict(nx=16, ny=16, nz=33, zorder=6)
    base.update(kw)
    prm = cases.closed_channel(**base)
    with pytest.raises(HerculesError):
        Channel(prm)


def test_z6_adjoint_rejected():
    prm, st = _case("closed", 33)
    ch = Channel(prm)
    with pytest.raises(HerculesError, match="status 2"):
        ch.wall_shear_gradient(2, 1)
    ch.close()


SPECTRAL_CASES = [
    ("closed", 33, {"issk": 1}),
    ("closed", 65, {"issk": 0, "dp_opt": 3, "fc": 0.5, "ug": 1.0, "vg": 0.2}),
    ("open", 33, {"issk": 1}),
    ("open", 65, {"issk": 1, "tbc_top": 1, "dtdz_top": 0.5}),
    ("open", 129, {"issk": 0}),
    ("open", 257, {"issk": 1, "tbc_top": 1, "dtdz_top": 0.5}),
    ("open", 65, {"issk": 1, "zorder": 4}),
    ("closed", 33, {"issk": 0, "zorder": 4}),
]


@pytest.mark.parametrize("kind,nz,kw", SPECTRAL_CASES)
def test_z6_spectral_steps_match_torch_reference(kind, nz, kw):
    """Spectral path (cds = 3) with the 6th-order z operators against the
    torch restatement of time_advancement_sp (tests/torch_spectral.py, equal
    to the C oracle at 2nd order) with the same dense operators."""
    from tests.torch_spectral import TorchSpectral
    prm, st = _case(kind, nz, cds=3, **kw)
    ref = TorchSpectral(prm).steps(st, 2)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    ch.steps(2)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    tol = {"p": 1e-9}
    for n in names:
        err = rel_l2(ch.get_field(n), ref[n].numpy())
        print(kind, nz, kw, n, f"{err:.2e}")
        assert err < tol.get(n, 1e-12), (n, err)
    _, dmax = ch.screen_div_error()
    scale = np.abs(ref["u"].numpy()).max() / (prm.lz / prm.nz)
    assert dmax < 1e-12 * scale, dmax
    ch.close()


def test_z6_ensemble_members_equal_single_runs():
    """6th-order members inside a replayed ensemble graph (the interpolant
    pre-pass and the Chebyshev columns captured) equal their single runs."""
    from hercules_amd import Ensemble
    specs = [("closed", 33, {}), ("open", 65, {"cds": 3, "issk": 1})]
    cs = [_case(kind, nz, **kw) for kind, nz, kw in specs]

    def make(prm, st):
        ch = Channel(prm)
        for k, v in st.items():
            ch.set_field(k, v)
        ch.init()
        return ch

    singles = [make(p, s) for p, s in cs]
    for ch in singles:
        ch.steps(4)
    members = [make(p, s) for p, s in cs]
    with Ensemble(members, use_graph=True) as ens:
        ens.steps(4)
        assert ens.graph_used
    for (prm, _), a, b in zip(cs, singles, members):
        for n in ["u", "v", "w", "p"] + (["t"] if prm.isscalar else []):
            assert np.array_equal(a.get_field(n), b.get_field(n)), (prm.cds, n)
    for ch in singles + members:
        ch.close()


@pytest.mark.parametrize("kw", [{"dts": 1}, {"dp_opt": 2}, {"dp_opt": 3, "fc": 0.5, "ug": 1.0, "vg": 0.2}])
def test_z6_other_flags_stay_divergence_free(kw):
    """AB2, the bulk-velocity drive and the Coriolis drive with the 6th-order
    operators (no restatement covers these flags): finite fields and a
    divergence at round-off after 5 steps."""
    prm, st = _case("closed", 65, **kw)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.project()
    ch.init()
    ch.steps(5)
    u = ch.get_field("u")
    assert np.isfinite(u).all()
    _, dmax = ch.screen_div_error()
    scale = np.abs(u).max() / (prm.lz / prm.nz)
    assert dmax < 1e-12 * scale, dmax
    ch.close()

"""Mixed precision variant (hc_params.precision = 1): fields, spectra,
right-hand-side stencils and FFTs in FP32 (FD and spectral paths); the column solves of the
low-wavenumber blocks (with the kx = ky = 0 mean mode) and the mean pressure
gradient / bulk velocity update in FP64.  With the 6th/4th-order z operators
every column is solved in FP64 arithmetic from FP32 spectra, with the
pressure correction kept in FP64.  The reference is FP64 only, so the
variant is held against the FP64 path of this library (itself equal to the
oracle to 1e-10), with the tolerances below; needs an MI355X.

Tolerances are relative L2 norms after the stated number of steps from the
same divergence-free FP64 state; v, w and p are small fluctuations, so their
relative errors are larger than u's.  Divergence of the FP32 fields is
bounded by the rounding of the stored velocities (|u| eps32 / dx)."""
import numpy as np
import pytest

from hercules_amd import Channel, Ensemble, cases
from hercules_amd.channel import HerculesError
from tests.helpers import poisson_residual, rel_l2, small_case

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


def _prepared(prm, st):
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.project()
    out = {k: ch.get_field(k) for k in st}
    ch.close()
    return out


def _run(prm, st, nsteps, precision):
    ch = Channel(prm.replace(precision=precision))
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    ch.steps(nsteps)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    out = {k: ch.get_field(k) for k in names}
    out["_div"] = ch.screen_div_error()
    out["_tau"] = ch.wall_shear()
    out["_ubar"] = ch.calculate_ubar()
    ch.close()
    return out


# relative L2 tolerances: u and t (large mean), v and w (fluctuations), p
TOL = {"u": 1e-5, "t": 1e-5, "v": 1e-4, "w": 1e-4, "p": 1e-3}
CASES = [
    # (kind, grid and physics, steps)
    ("closed", dict(nx=64, ny=64, nz=65), 40),
    ("closed", dict(nx=32, ny=32, nz=33, cds=1, dts=1), 40),
    ("closed", dict(nx=64, ny=32, nz=65, dp_opt=2), 40),
    ("closed", dict(nx=32, ny=64, nz=65, dp_opt=3, fc=0.5, ug=1.0, vg=0.2), 40),
    ("open_t", dict(nx=64, ny=64, nz=65), 40),
    ("open_t", dict(nx=64, ny=64, nz=65, tbc_top=1, dtdz_top=0.5), 40),
    ("open_t", dict(nx=128, ny=64, nz=129, is_ri_var=1, ri_str=0.0, ri_end=0.2, imax=40), 40),
    # nx = 512 takes the one-pass right-hand-side kernel
    ("closed", dict(nx=512, ny=16, nz=33, dt=2.5e-4), 10),
    # spectral path (cds = 3): spectra in FP32, low-wavenumber columns in FP64
    ("closed", dict(nx=64, ny=64, nz=65, cds=3, issk=1), 40),
    ("closed", dict(nx=32, ny=32, nz=33, cds=3, issk=0, dts=1), 40),
    ("closed", dict(nx=64, ny=32, nz=65, cds=3, issk=1, dp_opt=2, u_mrf=0.3), 40),
    ("open_t", dict(nx=64, ny=64, nz=65, cds=3, issk=1, tbc_top=1, dtdz_top=0.5), 40),
    ("open_t", dict(nx=128, ny=64, nz=129, cds=3, issk=1, is_ri_var=1, ri_str=0.0, ri_end=0.2, imax=40), 40),
    # 6th/4th-order z operators: FP32 storage, every column solved in FP64
    # arithmetic, the pressure correction held in FP64
    ("closed", dict(nx=64, ny=64, nz=65, zorder=6), 20),
    ("closed", dict(nx=32, ny=32, nz=65, cds=1, zorder=4), 20),
    ("open_t", dict(nx=64, ny=64, nz=65, zorder=6, tbc_top=1, dtdz_top=0.5), 20),
    ("closed", dict(nx=64, ny=64, nz=65, cds=3, issk=1, zorder=6), 20),
    ("open_t", dict(nx=64, ny=32, nz=129, cds=3, issk=1, zorder=6, tbc_top=1, dtdz_top=1.0), 20),
    ("open_t", dict(nx=64, ny=64, nz=65, cds=3, issk=0, zorder=4), 20),
]


@pytest.mark.parametrize("kind,kw,nsteps", CASES)
def test_mixed_step_tracks_fp64(kind, kw, nsteps):
    prm, st = small_case(kind, **kw)
    st = _prepared(prm, st)
    d = _run(prm, st, nsteps, 0)
    m = _run(prm, st, nsteps, 1)
    errs = {k: rel_l2(m[k], d[k]) for k in d if not k.startswith("_")}
    print(prm.nx, prm.ny, prm.nz, {k: f"{v:.2e}" for k, v in errs.items()})
    for k, e in errs.items():
        assert e < TOL[k], (k, errs)
    # plane-averaged wall shear and bulk velocity: FP64 reductions of FP32 fields
    tau_err = abs(m["_tau"][0] - d["_tau"][0]) / abs(d["_tau"][0])
    assert tau_err < 1e-5, (m["_tau"], d["_tau"])
    assert abs(m["_ubar"] - d["_ubar"]) <= 1e-5 * abs(d["_ubar"])
    # divergence of the stored FP32 velocities: their rounding over a grid cell
    umax = max(np.abs(d["u"]).max(), 1.0)
    h = min(prm.lx / prm.nx, prm.ly / prm.ny)
    assert m["_div"][1] < 40 * EPS32 * umax / h, (m["_div"], d["_div"])


def test_mixed_fields_are_fp32_roundings():
    prm, st = small_case("open_t", nx=32, ny=32, nz=33)
    ch = Channel(prm.replace(precision=1))
    for k, v in st.items():
        ch.set_field(k, v)
    for k in ("u", "t"):
        got = ch.get_field(k)
        assert np.array_equal(got, st[k].astype(np.float32).astype(np.float64))
    g = np.random.default_rng(3).standard_normal(ch.ghost_shape)
    ch.set_field("p", g, layout=1)
    back = ch.get_field("p")
    assert np.array_equal(back, g[2:-2, 2:-2, 2:-2].astype(np.float32).astype(np.float64))
    ch.close()


@pytest.mark.parametrize("grid", [(64, 64, 65, 2), (256, 256, 257, 2), (64, 64, 129, 6)])
def test_mixed_poisson_solve(grid):
    nx, ny, nz, zorder = grid
    prm = cases.closed_channel(nx=nx, ny=ny, nz=nz, zorder=zorder)
    rhs = cases.gaussian_poisson_rhs(prm, seed=2024)
    out = {}
    for pr in (0, 1):
        ch = Channel(prm.replace(precision=pr))
        out[pr] = ch.poisson_solver_fft("p", 1.0, rhs)
        ch.close()
    err = rel_l2(out[1], out[0])
    print(grid, f"rel err {err:.2e}")
    assert err < 2e-5
    if zorder == 2:  # the residual helper applies the 2nd-order z operator
        res64 = poisson_residual(prm, 1.0, out[0], rhs)
        res32 = poisson_residual(prm, 1.0, out[1], rhs)
        print(f"   residual fp64 {res64:.2e} mixed {res32:.2e}")
        assert res32 < 30 * EPS32


@pytest.mark.parametrize("zorder", [2, 6])
def test_mixed_projection(zorder):
    prm, st = small_case("closed", nx=64, ny=64, nz=65, zorder=zorder)
    out = {}
    for pr in (0, 1):
        ch = Channel(prm.replace(precision=pr))
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        out[pr] = ({k: ch.get_field(k) for k in ("u", "v", "w")}, ch.screen_div_error())
        ch.close()
    for k in ("u", "v", "w"):
        assert rel_l2(out[1][0][k], out[0][0][k]) < 1e-5, k
    umax = np.abs(out[0][0]["u"]).max()
    assert out[1][1][1] < 40 * EPS32 * umax / (prm.lx / prm.nx)


@pytest.mark.parametrize("kw", [dict(isdamp=1, nzdamp=8, cdamp=0.1), dict(cds=3, issk=1, isdamp=1, nzdamp=8, cdamp=0.1)])
def test_mixed_rejects_unsupported(kw):
    prm, _ = small_case("closed", nx=32, ny=32, nz=33, **kw)
    with pytest.raises(HerculesError, match="status 2"):
        Channel(prm.replace(precision=1))


def test_mixed_ensemble_members_equal_single_runs():
    specs = [("closed", dict(nx=32, ny=32, nz=33)), ("open_t", dict(nx=64, ny=32, nz=65)),
             ("closed", dict(nx=16, ny=16, nz=129, dp_opt=3, fc=0.5, ug=1.0, vg=0.2)),
             ("open_t", dict(nx=32, ny=32, nz=33, cds=3, issk=1))]
    cs = [small_case(kind, **kw) for kind, kw in specs]

    def make(prm, st):
        ch = Channel(prm.replace(precision=1))
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
            assert np.array_equal(a.get_field(n), b.get_field(n)), (prm.nx, n)
    for ch in singles + members:
        ch.close()

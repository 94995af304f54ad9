"""Pins of the CPU oracle (oracle/hc_oracle.c) against mathematics that
does not come from the oracle itself: numpy.fft, scipy banded solves,
operator symbols evaluated by applying the physical stencils to Fourier
modes, closed-form states and discrete conservation.  The reference has no
tests and no golden vectors of its own, and its build needs MPI, FFTW's
generated headers and 2DECOMP (not buildable here), so these checks are what
pins the restatement."""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from hercules_amd import cases, grid
from oracle.oracle import Oracle, fft1d
from tests.helpers import divergence, make_oracle, oracle_kwargs, poisson_residual, rel_l2

RNG = np.random.default_rng(1234)


def _oracle(prm, state=None, init=True):
    o = Oracle(**oracle_kwargs(prm))
    for k, v in (state or {}).items():
        o.set(k, v)
    if init:
        o.init()
    return o


# ---------------------------------------------------------------- FFT (A11)
@pytest.mark.parametrize("n", [8, 12, 16, 24, 48, 64, 96, 192, 256, 384])
def test_fft_matches_numpy(n):
    x = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    assert np.abs(fft1d(x, -1) - np.fft.fft(x)).max() < 1e-13 * np.abs(x).sum()
    assert np.abs(fft1d(x, +1) - np.fft.ifft(x) * n).max() < 1e-13 * np.abs(x).sum()


# ---------------------------------------------------------------- grid (A1)
@pytest.mark.parametrize("ochannel", [0, 1])
@pytest.mark.parametrize("nz,zs", [(17, 1.0), (65, 2.0), (64, 1.1)])
def test_mesh_invariants(ochannel, nz, zs):
    prm = cases.closed_channel(nx=16, ny=16, nz=nz, zstretch=zs, ochannel=ochannel,
                               lz=2.0 if ochannel == 0 else 1.0)
    o = _oracle(prm, init=False)
    o.init()
    m = o.mesh()
    dz = m["dz"][2:-2]
    assert np.all(dz > 0)
    assert abs(dz.sum() - prm.lz) < 1e-13
    if ochannel == 0:
        assert np.allclose(dz, dz[::-1], rtol=0, atol=1e-15)  # symmetric stretching
        assert np.argmin(dz) in (0, nz - 1)
    else:
        assert np.argmin(dz) == 0 and np.all(np.diff(dz) > 0)  # refined at the bottom only
    # ghost copies and derived metrics (src/module_parameters.f90:473-492)
    full = m["dz"]
    assert full[1] == full[2] and full[0] == full[2] and full[-2] == full[-3] and full[-1] == full[-3]
    k = np.arange(1, nz + 3)
    assert np.allclose(m["dz_t"][k], 0.5 * (full[k] + full[k + 1]), rtol=1e-15)
    assert np.allclose(m["l_t"][k], full[k] / (full[k] + full[k + 1]), rtol=1e-15)
    # the package's own grid (used by the host code) agrees exactly
    g = grid.metrics(nz, prm.lz, zs, ochannel)
    assert np.allclose(g["dz"], dz, rtol=0, atol=1e-15 * prm.lz)  # face differences: absolute


# ------------------------------------------------- symbols of the stencils (A8)
def _symbol(stencil, n, l):
    """Eigenvalue of a periodic stencil {offset: coefficient} on mode l."""
    th = 2 * np.pi * l / n
    return sum(c * np.exp(1j * s * th) for s, c in stencil.items())


def _operators(cds, h):
    if cds == 1:
        D = {0: 1 / h, -1: -1 / h}          # divergence at p points
        G = {1: 1 / h, 0: -1 / h}           # gradient at u points
        L = {1: 1 / h ** 2, 0: -2 / h ** 2, -1: 1 / h ** 2}
    else:
        D = {1: -1 / 24 / h, 0: 27 / 24 / h, -1: -27 / 24 / h, -2: 1 / 24 / h}
        G = {2: -1 / 24 / h, 1: 27 / 24 / h, 0: -27 / 24 / h, -1: 1 / 24 / h}
        L = {2: -1 / 12 / h ** 2, 1: 16 / 12 / h ** 2, 0: -30 / 12 / h ** 2,
             -1: 16 / 12 / h ** 2, -2: -1 / 12 / h ** 2}
    return D, G, L


@pytest.mark.parametrize("cds", [1, 2])
def test_wavenumbers_are_stencil_symbols(cds):
    prm = cases.closed_channel(nx=16, ny=12, nz=9, cds=cds)
    o = _oracle(prm)
    dx, dy = prm.lx / prm.nx, prm.ly / prm.ny
    ks = 5  # an interior level (1-based ghosted index)
    aa_p, bb_p, cc_p = o.coeff("p")
    aa_u, bb_u, cc_u = o.coeff("u")
    for j in range(prm.ny):
        for i in range(prm.nx):
            l = i if i < prm.nx // 2 else i - prm.nx
            m = j if j < prm.ny // 2 else j - prm.ny
            Dx, Gx, Lx = _operators(cds, dx)
            Dy, Gy, Ly = _operators(cds, dy)
            # p equation: bb = kx + ky - aa - cc with kx the symbol of D o G
            kx = _symbol(Dx, prm.nx, l) * _symbol(Gx, prm.nx, l)
            ky = _symbol(Dy, prm.ny, m) * _symbol(Gy, prm.ny, m)
            want = kx + ky - aa_p[ks - 1] - cc_p[ks - 1]
            assert abs(bb_p[ks - 1, j, i] - want) < 1e-9 * abs(want)
            # u equation: bb = -(dt nu / 2)(Lx + Ly) - aa - cc
            hx = 0.5 * prm.dt * prm.nu * (_symbol(Lx, prm.nx, l) + _symbol(Ly, prm.ny, m))
            want = -hx - aa_u[ks - 1] - cc_u[ks - 1]
            assert abs(bb_u[ks - 1, j, i] - want) < 1e-12 * (abs(want) + 1)


@pytest.mark.parametrize("wall_rows", [0, 1])
def test_wall_rows_of_first_mode(wall_rows):
    """get_u_eqn_coeff zeroes the 1-D aa(kstr3)/cc(kend3) while visiting the
    first (i,j) = mode (0,0): only that mode gets the -aa / -cc wall terms
    (src/module_poisson_solver.f90:257-272).  wall_rows=1 gives them to all."""
    prm = cases.closed_channel(nx=8, ny=8, nz=9, wall_rows=wall_rows)
    o = _oracle(prm)
    m = o.mesh()
    dz, dzt, dzb = m["dz"], m["dz_t"], m["dz_b"]
    ks, ke = 3, prm.nz + 2
    a = -0.5 * prm.dt * prm.nu / dz[ks - 1] / dzb[ks - 1]
    c = -0.5 * prm.dt * prm.nu / dz[ks - 1] / dzt[ks - 1]
    aa, bb, cc = o.coeff("u")
    assert aa[ks - 1] == 0.0 and cc[ke - 1] == 0.0
    assert abs(bb[ks - 1, 0, 0] - (-2 * a - c)) < 1e-14 * abs(a)
    other = bb[ks - 1, 0, 1]  # mode (1, 0)
    kx = 0.5 * prm.dt * prm.nu * _symbol(_operators(2, prm.lx / prm.nx)[2], prm.nx, 1)
    want = -kx - c + (-2 * a if wall_rows else 0.0)
    assert abs(other - want) < 1e-12 * abs(want)


# ---------------------------------------------------- Thomas + FFT (A10, A11)
@pytest.mark.parametrize("eq", ["u", "v", "w", "t", "p"])
def test_solve_single_mode_matches_banded(eq):
    """A source made of one Fourier mode (l, m) times a random z profile must
    come back as the same mode with the z profile of the tridiagonal system
    (assign_abc scaling, wall rows) solved by scipy.linalg.solve_banded."""
    # zero wall values: the u_mrf / tbot / ttop fix-ups only touch mode (0,0)
    prm = cases.closed_channel(nx=16, ny=16, nz=21, u_mrf=0.0, tbot=0.0, ttop=0.0)
    o = _oracle(prm)
    rk = 2.0 / 15.0
    l, m = 3, 2
    x = np.arange(prm.nx)[None, None, :]
    y = np.arange(prm.ny)[None, :, None]
    prof = RNG.standard_normal(prm.nz)[:, None, None]
    src = prof * np.cos(2 * np.pi * (l * x / prm.nx + m * y / prm.ny))
    out = o.solve(eq, rk, src)
    aa, bb, cc = o.coeff({"u": "u", "v": "u", "w": "w", "t": "t", "p": "p"}[eq])
    n = prm.nz - 1 if eq == "w" else prm.nz
    a = rk * aa[2:2 + n]
    c = rk * cc[2:2 + n]
    b = bb[2:2 + n, m, l].real
    b = rk * b if eq == "p" else 1.0 + rk * b
    ab = np.zeros((3, n))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    prof_out = solve_banded((1, 1), ab, prof[:n, 0, 0])
    want = np.zeros(prm.nz)
    want[:n] = prof_out
    expect = want[:, None, None] * np.cos(2 * np.pi * (l * x / prm.nx + m * y / prm.ny))
    assert rel_l2(out, expect) < 1e-12


@pytest.mark.parametrize("nz,zs", [(17, 1.0), (33, 2.0), (65, 1.0)])
def test_poisson_residual(nz, zs):
    prm = cases.closed_channel(nx=16, ny=16, nz=nz, zstretch=zs)
    o = _oracle(prm)
    src = RNG.standard_normal((nz, 16, 16))
    src -= src.mean(axis=(1, 2), keepdims=True)
    phi = o.solve("p", 8.0 / 15.0, src)
    assert poisson_residual(prm, 8.0 / 15.0, phi, src) < 1e-14
    assert abs(phi[0].mean()) < 1e-13 * np.abs(phi).max()  # gauge: level-1 mean is zero


def test_helmholtz_mean_mode_is_exact_cn():
    """For the (0,0) mode the u system is Crank-Nicolson with the no-slip
    ghost u(kstr3-1) = -u(kstr3): check (1 - rk_c dt nu/2 d2/dz2) u = q on
    the plane means."""
    prm = cases.closed_channel(nx=8, ny=8, nz=17, zstretch=1.0)
    o = _oracle(prm)
    rk = 8.0 / 15.0
    q = np.broadcast_to(RNG.standard_normal(prm.nz)[:, None, None], (prm.nz, 8, 8)).copy()
    u = o.solve("u", rk, q)[:, 0, 0]
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)
    ug = np.concatenate([[-u[0]], u, [-u[-1]]])
    uzz = ((ug[2:] - ug[1:-1]) / m["dz_t"] - (ug[1:-1] - ug[:-2]) / m["dz_b"]) / m["dz"]
    res = u - rk * 0.5 * prm.dt * prm.nu * uzz - q[:, 0, 0]
    assert np.abs(res).max() < 1e-12 * np.abs(q).max()


# ------------------------------------------------------------ RHS (A3-A7)
def _zero_state(prm):
    z = np.zeros((prm.nz, prm.ny, prm.nx))
    return dict(u=z.copy(), v=z.copy(), w=z.copy(), p=z.copy(), t=z.copy())


def test_uniform_flow_right_hand_side():
    prm = cases.closed_channel(nx=16, ny=16, nz=9, zstretch=1.0)
    st = _zero_state(prm)
    U0 = 3.0
    st["u"] += U0
    o = _oracle(prm, st)
    from oracle.oracle import lib
    lib().ho_stage_momentum(o.h, 1)
    ff, qu = o.get("ff"), o.get("qu")
    assert np.abs(ff - 1.0).max() < 1e-13  # only dpx_drive
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)
    ra, rc, dt, nu = 8 / 15, 8 / 15, prm.dt, prm.nu
    uzz = np.zeros(prm.nz)
    uzz[0] = -2 * U0 / m["dz_b"][0] / m["dz"][0]
    uzz[-1] = -2 * U0 / m["dz_t"][-1] / m["dz"][-1]
    want = U0 + dt * ra * 1.0 + rc * 0.5 * dt * nu * uzz
    assert np.abs(qu - want[:, None, None]).max() < 1e-13 * U0


def test_coriolis_and_buoyancy_terms():
    prm = cases.closed_channel(nx=16, ny=16, nz=9, zstretch=1.0, dp_opt=3, fc=0.7, ug=1.3, vg=0.4,
                               isscalar=1, got=2.5, t_ref=0.25, u_mrf=0.0)
    st = _zero_state(prm)
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)
    st["t"] += (0.5 + 0.3 * m["zc"])[:, None, None]  # linear in z
    o = _oracle(prm, st)
    from oracle.oracle import lib
    lib().ho_stage_momentum(o.h, 1)
    assert np.abs(o.get("ff") - (-prm.fc * prm.vg)).max() < 1e-14
    assert np.abs(o.get("gg") - prm.fc * prm.ug).max() < 1e-14
    hh = o.get("hh")[:-1]
    zf = m["z_faces"][1:-1]  # w points
    want = prm.got * (0.5 + 0.3 * zf - prm.t_ref)
    assert np.abs(hh - want[:, None, None]).max() < 1e-12


@pytest.mark.parametrize("n", [16, 32, 64])
def test_convective_flux_is_fourth_order(n):
    """u = sin(2 pi x / lx): the 4th-order interpolation + flux divergence of
    u^2 converges to d(u^2)/dx at 4th order (src/module_cnvdiff.f90:72-75)."""
    errs = []
    for nn in (n, 2 * n):
        prm = cases.closed_channel(nx=nn, ny=16, nz=5, zstretch=1.0)
        st = _zero_state(prm)
        xu = (np.arange(nn) + 1.0) * prm.lx / nn  # u points at x_{i+1/2}, i from 1
        st["u"] += np.sin(2 * np.pi * xu / prm.lx)[None, None, :]
        o = _oracle(prm, st)
        from oracle.oracle import lib
        lib().ho_stage_momentum(o.h, 1)
        ff = o.get("ff")[2, 3]
        exact = 1.0 - (2 * np.pi / prm.lx) * np.sin(4 * np.pi * xu / prm.lx)
        errs.append(np.abs(ff - exact).max())
    order = np.log2(errs[0] / errs[1])
    assert order > 3.8, (errs, order)


def test_mean_momentum_budget_flux_form():
    """With w = 0 on both walls the vertical flux term telescopes: the
    dz-weighted mean of ff over the channel equals dpx_drive (flux form of
    get_momentum_cnvdiff plus the wall rules of update_boundary_uvwhxyz)."""
    prm = cases.closed_channel(nx=16, ny=16, nz=17, zstretch=1.0)
    st = cases.laminar_noise(prm, seed=3)
    o = _oracle(prm, st)
    from oracle.oracle import lib
    lib().ho_stage_momentum(o.h, 1)
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)
    ff = o.get("ff")
    mean = (ff.mean(axis=(1, 2)) * m["dz"]).sum() / prm.lz
    assert abs(mean - 1.0) < 1e-11 * np.abs(ff).max()


# --------------------------------------------------- assembled step (A16)
def test_step_is_divergence_free():
    prm = cases.closed_channel(nx=16, ny=16, nz=17, zstretch=1.0)
    st = cases.laminar_noise(prm, seed=4)
    o = _oracle(prm, st)
    o.steps(2)
    d, terms = divergence(o.get("u"), o.get("v"), o.get("w"), prm)
    assert np.abs(d).max() / terms.max() < 1e-12
    assert np.abs(o.get("w")[-1]).max() == 0.0  # top face stays closed


def test_ubar_and_constant_flow_rate_drive():
    """calculate_ubar and update_dpx_drive (dp_opt=2): dpx after one RK3 step
    is -3 (ubar0 + u_mrf - 1) nu from dpx = 0 (src/module_parameters.f90:594-613)."""
    prm = cases.closed_channel(nx=16, ny=16, nz=9, zstretch=1.0, dp_opt=2, u_mrf=0.1)
    st = cases.laminar_noise(prm, seed=5)
    o = _oracle(prm, st)
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)
    ubar0 = (st["u"].mean(axis=(1, 2)) * m["dz"]).sum() / prm.lz
    assert abs(o.scalar("ubar") - ubar0) < 1e-13 * abs(ubar0)
    o.steps(1)
    want = -3.0 * (ubar0 + prm.u_mrf - 1.0) * prm.nu
    assert abs(o.scalar("dpx_drive") - want) < 1e-13 * abs(want)


def test_rk3_coefficients_are_consistent():
    """Orlandi's RK3 (src/module_parameters.f90:210-212): c = a + b and the
    substep weights of the explicit terms sum to one step."""
    a = np.array([8 / 15, 5 / 12, 3 / 4])
    b = np.array([0, -17 / 60, -5 / 12])
    c = np.array([8 / 15, 2 / 15, 1 / 3])
    assert np.allclose(a + b, c) and abs(c.sum() - 1.0) < 1e-15


def test_first_substep_ignores_old_rhs():
    """rkc2(1) = 0: the state after one step does not depend on ff0 etc.,
    which makes an RK3 restart exact (src/module_navier_stokes.f90:270)."""
    prm = cases.closed_channel(nx=16, ny=16, nz=9, zstretch=1.0)
    st = cases.laminar_noise(prm, seed=6)
    o1 = _oracle(prm, st)
    o1.steps(1)
    o2 = _oracle(prm, st)
    o2.set("ff", np.full(o2.shape, 1e3))
    o2.steps(1)
    for n in ("u", "v", "w", "p"):
        assert np.array_equal(o1.get(n), o2.get(n))


# ------------------------------------------------- Rayleigh damping (A19)
@pytest.mark.parametrize("nzdamp,cdamp", [(4, 0.3), (7, 1.0), (16, 0.05)])
def test_damping_sponge_profile(nzdamp, cdamp):
    """rayleigh_damping (src/module_tools.f90:705-749) keeps every plane mean
    and scales the fluctuation of the top nzdamp+1 levels by 1 - sigma with
    sigma = cdamp sin^2(pi zd / 2), zd = 0 .. 1 from the lowest damped level
    to the top; lower levels and all ghosts are untouched."""
    prm = cases.closed_channel(nx=16, ny=8, nz=17, zstretch=1.0, isdamp=1, nzdamp=nzdamp, cdamp=cdamp)
    st = cases.laminar_noise(prm, seed=21)
    o = _oracle(prm, st)
    before = o.get_ghosted("u")
    from oracle.oracle import lib
    lib().ho_stage_damping(o.h, 0)
    after = o.get_ghosted("u")
    b, a = before[2:-2, 2:-2, 2:-2], after[2:-2, 2:-2, 2:-2]
    k0 = prm.nz - 1 - nzdamp
    for k in range(prm.nz):
        mb, ma = b[k].mean(), a[k].mean()
        assert abs(ma - mb) < 1e-13 * (abs(mb) + 1.0)
        if k < k0:
            assert np.array_equal(a[k], b[k])
            continue
        sigma = cdamp * np.sin(0.5 * np.pi * (k - k0) / nzdamp) ** 2
        fb, fa = b[k] - mb, a[k] - ma
        assert np.abs(fa - (1.0 - sigma) * fb).max() < 1e-13 * (np.abs(b[k]).max() + 1.0)
    if cdamp == 1.0:
        assert np.ptp(a[-1]) < 1e-13 * np.abs(b[-1]).max()  # top level relaxed onto its mean
    # halos and z ghosts keep their pre-damping values
    mask = np.ones(after.shape, bool)
    mask[2:-2, 2:-2, 2:-2] = False
    assert np.array_equal(after[mask], before[mask])


def test_damping_with_zero_strength_is_inert():
    prm0 = cases.closed_channel(nx=16, ny=16, nz=17, zstretch=1.0)
    prm1 = cases.closed_channel(nx=16, ny=16, nz=17, zstretch=1.0, isdamp=1, nzdamp=6, cdamp=0.0)
    st = cases.laminar_noise(prm0, seed=22)
    o0, o1 = _oracle(prm0, st), _oracle(prm1, st)
    o0.steps(2)
    o1.steps(2)
    for n in ("u", "v", "w", "p"):
        assert np.array_equal(o0.get(n), o1.get(n))


def test_damped_step_stays_divergence_free():
    prm = cases.closed_channel(nx=16, ny=16, nz=17, zstretch=1.0, isdamp=1, nzdamp=6, cdamp=0.5)
    st = cases.laminar_noise(prm, seed=23)
    o = _oracle(prm, st)
    o.steps(2)
    d, terms = divergence(o.get("u"), o.get("v"), o.get("w"), prm)
    assert np.abs(d).max() / terms.max() < 1e-12


def test_damping_rejects_levels_outside_the_interior():
    prm = cases.closed_channel(nx=16, ny=16, nz=9, isdamp=1, nzdamp=9)
    with pytest.raises(ValueError):
        Oracle(**oracle_kwargs(prm))


# ---- top temperature gradient (tbc_top = 1): an extension with no reference
# function, pinned by an exact steady state and by mirror symmetry.

@pytest.mark.parametrize("cds", [1, 2, 3])
def test_top_gradient_linear_profile_is_steady(cds):
    """t = tbot + g z with the gradient g imposed at the top and t = tbot at
    the bottom is an exact discrete steady state of the 2nd-order z operator
    on the stretched grid (the ghosts extend the line), while the mean drive
    accelerates u(z) and buoyancy is balanced hydrostatically."""
    g, tb = 0.7, 0.2
    prm = cases.stratified_open_channel(nx=16, ny=16, nz=33, cds=cds, issk=1, tbc_top=1, dtdz_top=g, tbot=tb,
                                        ttop=123.0, wall_rows=1, got=0.1)
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
    shape = (prm.nz, prm.ny, prm.nx)
    t0 = np.broadcast_to((tb + g * m["zc"])[:, None, None], shape).copy()
    z = np.zeros(shape)
    o = make_oracle(prm, dict(u=z, v=z, w=z, p=z, t=t0))
    o.steps(5)
    assert np.abs(o.get("t") - t0).max() < 1e-12 * np.abs(t0).max()
    assert np.abs(o.get("w")).max() < 1e-12
    assert np.abs(o.get("u")).max() > 1e-4  # the drive did act


@pytest.mark.parametrize("cds", [2, 3])
def test_top_zero_gradient_is_the_mirror_of_a_closed_channel(cds):
    """An open channel of height h (free-slip top, dT/dz = 0 at the top) is the
    lower half of a closed channel of height 2h whose state is mirror-symmetric
    about the centre plane (u, v, p, t even, w odd) with t = tbot at both
    walls: the two grids coincide there (get_mesh_param), so after any number
    of steps the open channel equals the closed channel's lower half."""
    m_lev = 16
    common = dict(nx=16, ny=16, cds=cds, issk=1, nu=1.0 / 150.0, alpha=1.0 / 150.0 / 0.71, got=0.0, isscalar=1,
                  tbot=0.3, wall_rows=1, dt=5e-4, dp_opt=1)
    full = cases.closed_channel(nz=2 * m_lev, lz=2.0, ttop=0.3, **common)
    half = cases.stratified_open_channel(nz=m_lev, lz=1.0, tbc_top=1, dtdz_top=0.0, ttop=-5.0, **common)
    rng = np.random.default_rng(4)
    sh = (m_lev, 16, 16)
    zc = grid.metrics(m_lev, 1.0, half.zstretch, 1)["zc"][:, None, None]
    u_l = 10.0 * zc * (2.0 - zc) + 0.3 * rng.standard_normal(sh)
    v_l = 0.3 * rng.standard_normal(sh)
    t_l = 0.3 + 0.5 * zc + 0.1 * rng.standard_normal(sh)
    w_l = 0.1 * rng.standard_normal(sh)
    w_l[-1] = 0.0  # the centre face of the odd w
    even = lambda a: np.concatenate([a, a[::-1]], 0)
    odd_w = np.concatenate([w_l, -w_l[-2::-1], np.zeros((1,) + sh[1:])], 0)
    z = np.zeros((2 * m_lev,) + sh[1:])
    of = make_oracle(full, dict(u=even(u_l), v=even(v_l), w=odd_w, p=z, t=even(t_l)))
    oh = make_oracle(half, dict(u=u_l, v=v_l, w=w_l, p=z[:m_lev], t=t_l))
    # the odd w of the full channel: w(k) of the faces above the centre is
    # -w(2m-2-k); its top face (the wall) is 0
    assert np.array_equal(odd_w[m_lev - 1], np.zeros(sh[1:]))
    of.steps(6)
    oh.steps(6)
    for n in ("u", "v", "w", "p", "t"):
        a, b = oh.get(n), of.get(n)[:m_lev]
        assert rel_l2(a, b) < 1e-10, (n, rel_l2(a, b))


def test_top_gradient_rejects_unknown_flag():
    prm = cases.stratified_open_channel(nx=8, ny=8, nz=9, tbc_top=2)
    with pytest.raises(ValueError):
        Oracle(**oracle_kwargs(prm))

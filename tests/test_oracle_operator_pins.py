"""Operator-level pins of the CPU oracle (oracle/hc_oracle.c) against
mathematics fixed independently of it:

* the explicit right-hand sides of the u, v, w and T equations
  (get_momentum_cnvdiff, src/module_cnvdiff.f90:35-374, and
  get_temperature_cnvdiff, :381-473) converge to the continuous convection,
  diffusion, pressure-gradient and buoyancy terms of smooth fields at the
  order of their stencils: 4th order (cds = 2) or 2nd order (cds = 1) in x
  and y, 2nd order in z on the stretched grid;
* the flux form makes the dz-weighted volume means of the v and T right-hand
  sides vanish when w = 0 on both walls;
* calculate_new_p (src/module_navier_stokes.f90:739-808) and calculate_uvw
  (:659-734) act on one Fourier mode through the symbols of their stencils;
* the Crank-Nicolson rows of the w and u systems (get_w_eqn_coeff,
  get_u_eqn_coeff, src/module_poisson_solver.f90:181-374, solved by
  poisson_solver_fft) are the explicit z operator of the right-hand side
  with the wall ghosts (w = 0 on both wall faces; u no-slip), for the
  (0,0) mode, and for every mode with wall_rows = 1; with the reference's
  rows the other modes see a zero-gradient ghost instead;
* the temperature equation of a substep is advanced with the velocities
  after the projection (src/module_navier_stokes.f90:319-341).

The continuous values come from exact trigonometric algebra (sums of
Fourier modes, multiplied and differentiated exactly), not from the oracle.
"""
import numpy as np
import pytest

from hercules_amd import cases, grid
from oracle.oracle import lib
from tests.helpers import make_oracle, rel_l2


# ----------------------------------------------------------------------------
# exact trigonometric fields: real functions held as sums of complex modes
class Modes:
    """f(x, y, z) = sum c exp(i (a x + b y + g z)); real fields keep every
    mode together with its conjugate."""

    def __init__(self, d=None):
        self.d = dict(d or {})

    @staticmethod
    def const(c):
        return Modes({(0.0, 0.0, 0.0): complex(c)})

    @staticmethod
    def re(c, a=0.0, b=0.0, g=0.0):
        """Re(c exp(i (a x + b y + g z)))"""
        c = complex(c)
        m = Modes({(a, b, g): 0.5 * c})
        return m + Modes({(-a, -b, -g): 0.5 * c.conjugate()})

    def __add__(self, o):
        o = o if isinstance(o, Modes) else Modes.const(o)
        d = dict(self.d)
        for k, c in o.d.items():
            d[k] = d.get(k, 0.0) + c
        return Modes(d)

    __radd__ = __add__

    def __sub__(self, o):
        return self + (o * -1.0)

    def __mul__(self, o):
        if not isinstance(o, Modes):
            return Modes({k: c * o for k, c in self.d.items()})
        d = {}
        for (a1, b1, g1), c1 in self.d.items():
            for (a2, b2, g2), c2 in o.d.items():
                k = (a1 + a2, b1 + b2, g1 + g2)
                d[k] = d.get(k, 0.0) + c1 * c2
        return Modes(d)

    __rmul__ = __mul__

    def __neg__(self):
        return self * -1.0

    def dx(self):
        return Modes({k: c * 1j * k[0] for k, c in self.d.items()})

    def dy(self):
        return Modes({k: c * 1j * k[1] for k, c in self.d.items()})

    def dz(self):
        return Modes({k: c * 1j * k[2] for k, c in self.d.items()})

    def __call__(self, x, y, z):
        out = 0.0
        for (a, b, g), c in self.d.items():
            out = out + c * np.exp(1j * (a * x + b * y + g * z))
        return np.real(out)


def _stage(o, name, t_its=1):
    getattr(lib(), name)(o.h, t_its)


# ----------------------------------------------------------------------------
# horizontal terms: z-independent fields, interior levels
def _xy_case(cds, n):
    prm = cases.closed_channel(nx=n, ny=n, nz=9, cds=cds, isscalar=1, got=0.35, t_ref=0.1,
                               u_mrf=0.2, tbot=0.0, ttop=0.0)
    kx, ky = 2 * np.pi / prm.lx, 2 * np.pi / prm.ly
    U = 1.3 + 0.5 * Modes.re(-1j * np.exp(0.3j), a=kx) * Modes.re(1.0, b=ky)
    V = 0.7 * Modes.re(np.exp(-0.2j), a=kx) * Modes.re(-1j * np.exp(0.1j), b=2 * ky)
    W = 0.4 * Modes.re(-1j, a=kx, b=ky)
    T = 0.6 + 0.2 * Modes.re(1.0, a=2 * kx, b=-ky)
    P = 0.3 * Modes.re(-1j, a=kx) * Modes.re(-1j * np.exp(0.5j), b=ky)
    return prm, dict(U=U, V=V, W=W, T=T, P=P)


def _points(prm):
    """x of p (and w, T) points and of u points, y of p and of v points: u
    sits half a cell in +x of p, v half a cell in +y
    (src/module_variables.f90:34-36)."""
    dx, dy = prm.lx / prm.nx, prm.ly / prm.ny
    xp = (np.arange(prm.nx) * dx)[None, None, :]
    yp = (np.arange(prm.ny) * dy)[None, :, None]
    return xp, xp + 0.5 * dx, yp, yp + 0.5 * dy


def _xy_errors(cds, n):
    prm, F = _xy_case(cds, n)
    U, V, W, T, P = F["U"], F["V"], F["W"], F["T"], F["P"]
    xp, xu, yp, yv = _points(prm)
    shape = (prm.nz, prm.ny, prm.nx)
    z0 = np.zeros((prm.nz, 1, 1))
    ev = lambda f, x, y: np.broadcast_to(f(x, y, z0), shape).copy()
    st = dict(u=ev(U, xu, yp), v=ev(V, xp, yv), w=ev(W, xp, yp), t=ev(T, xp, yp), p=ev(P, xp, yp))
    o = make_oracle(prm, st)
    _stage(o, "ho_stage_momentum", 1)
    ff, gg, hh = o.get("ff"), o.get("gg"), o.get("hh")
    qu, qv, qw = o.get("qu"), o.get("qv"), o.get("qw")
    _stage(o, "ho_stage_temperature", 1)
    ss, qt = o.get("ss"), o.get("qu")
    um, nu, al, dt, got = prm.u_mrf, prm.nu, prm.alpha, prm.dt, prm.got
    ra = rc = 8.0 / 15.0
    lap = lambda f: f.dx().dx() + f.dy().dy()
    want = {
        # -d(uu)/dx - d(uv)/dy + dpx - u_mrf du/dx at u points (dp_opt = 1)
        "f": (ev(-(U * U).dx() - (U * V).dy() - um * U.dx() + 1.0, xu, yp), ff),
        "g": (ev(-(U * V).dx() - (V * V).dy() - um * V.dx(), xp, yv), gg),
        # w points: horizontal fluxes and the buoyancy of the face value of t
        "h": (ev(-(U * W).dx() - (V * W).dy() - um * W.dx() + got * (T - prm.t_ref), xp, yp), hh),
        "s": (ev(-(U * T).dx() - (V * T).dy() - um * T.dx(), xp, yp), ss),
        # linear parts (q - f - dt ra rhs) / (rc dt): nu/2 lap - grad p
        "Lu": (ev(0.5 * nu * lap(U) - P.dx(), xu, yp), (qu - st["u"] - dt * ra * ff) / (rc * dt)),
        "Lv": (ev(0.5 * nu * lap(V) - P.dy(), xp, yv), (qv - st["v"] - dt * ra * gg) / (rc * dt)),
        "Lw": (ev(0.5 * nu * lap(W), xp, yp), (qw - st["w"] - dt * ra * hh) / (rc * dt)),
        "Lt": (ev(0.5 * al * lap(T), xp, yp), (qt - st["t"] - dt * ra * ss) / (rc * dt)),
    }
    inner = slice(2, prm.nz - 2)  # stencils that reach no wall ghost
    return {k: np.abs(a[inner] - b[inner]).max() / np.abs(a[inner]).max() for k, (a, b) in want.items()}


@pytest.mark.parametrize("cds,order", [(2, 4.0), (1, 2.0)])
def test_horizontal_terms_converge_at_stencil_order(cds, order):
    e1, e2 = _xy_errors(cds, 64), _xy_errors(cds, 128)
    for k in e1:
        obs = np.log2(e1[k] / e2[k])
        assert obs > order - 0.25, (k, e1[k], e2[k], obs)
        assert e2[k] < (3e-4 if cds == 2 else 2e-2), (k, e2[k])


# ----------------------------------------------------------------------------
# vertical terms: horizontally uniform fields, interior levels, stretched grid
def _z_errors(nz, ochannel):
    lz = 2.0 if ochannel == 0 else 1.0
    prm = cases.closed_channel(nx=8, ny=8, nz=nz, zstretch=1.0, ochannel=ochannel, lz=lz, isscalar=1, got=0.35,
                               t_ref=0.1, tbot=0.0, ttop=0.0)
    U = 1.0 + 0.5 * Modes.re(-1j * np.exp(0.2j), g=1.3)
    V = 0.3 * Modes.re(1.0, g=0.9)
    W = 0.4 * Modes.re(-1j * np.exp(0.4j), g=0.7)
    T = 0.5 + 0.2 * Modes.re(1.0, g=1.7)
    P = 0.25 * Modes.re(1.0, g=1.1)
    m = grid.metrics(nz, lz, 1.0, ochannel)
    zc = m["zc"][:, None, None]
    zw = m["z_faces"][1:, None, None]  # w(k) sits on the upper face of cell k
    shape = (nz, 8, 8)
    ev = lambda f, z: np.broadcast_to(f(0.0, 0.0, z), shape).copy()
    st = dict(u=ev(U, zc), v=ev(V, zc), w=ev(W, zw), t=ev(T, zc), p=ev(P, zc))
    st["w"][-1] = 0.0  # the top wall face
    o = make_oracle(prm, st)
    _stage(o, "ho_stage_momentum", 1)
    ff, gg, hh = o.get("ff"), o.get("gg"), o.get("hh")
    qu, qv, qw = o.get("qu"), o.get("qv"), o.get("qw")
    _stage(o, "ho_stage_temperature", 1)
    ss, qt = o.get("ss"), o.get("qu")
    nu, al, dt, got = prm.nu, prm.alpha, prm.dt, prm.got
    ra = rc = 8.0 / 15.0
    d2 = lambda f: f.dz().dz()
    want = {
        "f": (ev(-(U * W).dz() + 1.0, zc), ff),
        "g": (ev(-(V * W).dz(), zc), gg),
        "h": (ev(-(W * W).dz() + got * (T - prm.t_ref), zw), hh),
        "s": (ev(-(T * W).dz(), zc), ss),
        "Lu": (ev(0.5 * nu * d2(U), zc), (qu - st["u"] - dt * ra * ff) / (rc * dt)),
        "Lv": (ev(0.5 * nu * d2(V), zc), (qv - st["v"] - dt * ra * gg) / (rc * dt)),
        "Lw": (ev(0.5 * nu * d2(W) - P.dz(), zw), (qw - st["w"] - dt * ra * hh) / (rc * dt)),
        "Lt": (ev(0.5 * al * d2(T), zc), (qt - st["t"] - dt * ra * ss) / (rc * dt)),
    }
    inner = slice(2, nz - 3)  # no wall ghost and not the top face of w
    return {k: np.abs(a[inner] - b[inner]).max() / np.abs(a[inner]).max() for k, (a, b) in want.items()}


@pytest.mark.parametrize("ochannel", [0, 1])
def test_vertical_terms_converge_at_second_order(ochannel):
    e1, e2 = _z_errors(65, ochannel), _z_errors(129, ochannel)
    for k in e1:
        obs = np.log2(e1[k] / e2[k])
        assert obs > 1.75, (k, e1[k], e2[k], obs)
        assert e2[k] < 5e-3, (k, e2[k])


# ----------------------------------------------------------------------------
@pytest.mark.parametrize("cds", [1, 2])
def test_flux_form_budgets_of_v_and_t(cds):
    """w = 0 on both wall faces: the vertical fluxes telescope and the
    horizontal ones are periodic, so the dz-weighted volume means of gg and
    ss vanish (no Coriolis, dp_opt = 1)."""
    prm = cases.closed_channel(nx=16, ny=16, nz=17, zstretch=1.0, cds=cds, isscalar=1, got=0.3, u_mrf=0.15,
                               tbot=0.2, ttop=0.7)
    st = cases.laminar_noise(prm, seed=31)
    st["t"] = 0.4 + np.random.default_rng(32).standard_normal(st["u"].shape)
    o = make_oracle(prm, st)
    _stage(o, "ho_stage_momentum", 1)
    gg = o.get("gg")
    _stage(o, "ho_stage_temperature", 1)
    ss = o.get("ss")
    dz = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)["dz"]
    for f in (gg, ss):
        mean = (f.mean(axis=(1, 2)) * dz).sum() / prm.lz
        assert abs(mean) < 1e-12 * np.abs(f).max(), mean


# ----------------------------------------------------------------------------
def _symbol(stencil, n, l):
    th = 2 * np.pi * l / n
    return sum(c * np.exp(1j * s * th) for s, c in stencil.items())


def _lap_stencil(cds, h):
    if cds == 1:
        return {1: 1 / h ** 2, 0: -2 / h ** 2, -1: 1 / h ** 2}
    return {2: -1 / 12 / h ** 2, 1: 16 / 12 / h ** 2, 0: -30 / 12 / h ** 2, -1: 16 / 12 / h ** 2, -2: -1 / 12 / h ** 2}


def _grad_stencil(cds, h):
    """p points -> the next u (v) point"""
    if cds == 1:
        return {1: 1 / h, 0: -1 / h}
    return {2: -1 / 24 / h, 1: 27 / 24 / h, 0: -27 / 24 / h, -1: 1 / 24 / h}


def _mode(prm, l, m, prof):
    i = np.arange(prm.nx)[None, None, :]
    j = np.arange(prm.ny)[None, :, None]
    return prof[:, None, None] * np.exp(2j * np.pi * (l * i / prm.nx + m * j / prm.ny))


@pytest.mark.parametrize("cds", [1, 2])
@pytest.mark.parametrize("t_its", [1, 2, 3])
def test_new_p_acts_on_a_fourier_mode_through_its_symbols(cds, t_its):
    """calculate_new_p: p += phi - rk_c dt nu/2 (Lx + Ly + Lzz) phi, Lzz the
    2nd-order z operator with the Neumann ghosts of update_boundary_p."""
    prm = cases.closed_channel(nx=16, ny=12, nz=17, zstretch=1.0, cds=cds)
    rng = np.random.default_rng(40 + t_its)
    l, m = 3, 2
    prof = rng.standard_normal(prm.nz) + 1j * rng.standard_normal(prm.nz)
    phi_c = _mode(prm, l, m, prof)
    p0 = rng.standard_normal((prm.nz, prm.ny, prm.nx))
    z = np.zeros_like(p0)
    o = make_oracle(prm, dict(u=z, v=z, w=z, p=p0, t=z))
    o.set("phi", phi_c.real)
    _stage(o, "ho_stage_new_p", t_its)
    dx, dy = prm.lx / prm.nx, prm.ly / prm.ny
    sig = _symbol(_lap_stencil(cds, dx), prm.nx, l) + _symbol(_lap_stencil(cds, dy), prm.ny, m)
    me = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)
    g = np.concatenate([prof[:1], prof, prof[-1:]])  # Neumann ghosts
    lzz = ((g[2:] - g[1:-1]) / me["dz_t"] - (g[1:-1] - g[:-2]) / me["dz_b"]) / me["dz"]
    rc = [8 / 15, 2 / 15, 1 / 3][t_its - 1]
    inc = prof - rc * 0.5 * prm.dt * prm.nu * (sig * prof + lzz)
    want = p0 + _mode(prm, l, m, inc).real
    assert rel_l2(o.get("p") - p0, want - p0) < 1e-13


@pytest.mark.parametrize("cds", [1, 2])
def test_projection_acts_on_a_fourier_mode_through_its_symbols(cds):
    """calculate_uvw: u -= rk_c dt Gx phi, v -= rk_c dt Gy phi, w(k) -= rk_c dt
    (phi(k+1) - phi(k)) / dz_t(k) below the top face, which stays 0."""
    prm = cases.closed_channel(nx=16, ny=12, nz=17, zstretch=1.0, cds=cds)
    rng = np.random.default_rng(50)
    l, m = 5, 3
    prof = rng.standard_normal(prm.nz) + 1j * rng.standard_normal(prm.nz)
    sh = (prm.nz, prm.ny, prm.nx)
    u0, v0, w0 = (rng.standard_normal(sh) for _ in range(3))
    w0[-1] = 0.0
    o = make_oracle(prm, dict(u=u0, v=v0, w=w0, p=np.zeros(sh), t=np.zeros(sh)))
    o.set("phi", _mode(prm, l, m, prof).real)
    _stage(o, "ho_stage_project", 2)
    c = 2.0 / 15.0 * prm.dt
    dx, dy = prm.lx / prm.nx, prm.ly / prm.ny
    gx = _symbol(_grad_stencil(cds, dx), prm.nx, l)
    gy = _symbol(_grad_stencil(cds, dy), prm.ny, m)
    dzt = grid.metrics(prm.nz, prm.lz, prm.zstretch, 0)["dz_t"]
    gz = np.zeros(prm.nz, complex)
    gz[:-1] = (prof[1:] - prof[:-1]) / dzt[:-1]
    # the increments are ~1e-4 of the fields: round-off of the difference
    assert rel_l2(o.get("u") - u0, -c * _mode(prm, l, m, gx * prof).real) < 1e-11
    assert rel_l2(o.get("v") - v0, -c * _mode(prm, l, m, gy * prof).real) < 1e-11
    assert rel_l2(o.get("w") - w0, -c * _mode(prm, l, m, gz).real) < 1e-11


# ----------------------------------------------------------------------------
def _cn_residual(prm, eq, rk, l, m, sol, src, ghost):
    """max |sol - rk dt nu/2 (sigma_L sol + Lzz sol) - src| over the rows of
    one (l, m) profile, Lzz the explicit z operator of the right-hand side
    (src/module_cnvdiff.f90:121-123 for u, :313-339 for w) with the given
    wall ghosts; relative to |src|."""
    me = grid.metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
    dz, dzt, dzb = me["dz"], me["dz_t"], me["dz_b"]
    dx, dy = prm.lx / prm.nx, prm.ly / prm.ny
    sig = _symbol(_lap_stencil(prm.cds, dx), prm.nx, l) + _symbol(_lap_stencil(prm.cds, dy), prm.ny, m)
    if eq == "w":  # faces 0..nz-2; the ghosts are the wall faces
        s = sol[:-1]
        lo, hi = ghost(s)
        g = np.concatenate([[lo], s, [hi]])
        # (w(k+1)-w(k))/dz(k+1) - (w(k)-w(k-1))/dz(k), over dz_t(k)
        lzz = ((g[2:] - g[1:-1]) / dz[1:] - (g[1:-1] - g[:-2]) / dz[:-1]) / dzt[:-1]
        r = s - rk * 0.5 * prm.dt * prm.nu * (sig * s + lzz) - src[:-1]
    else:
        lo, hi = ghost(sol)
        g = np.concatenate([[lo], sol, [hi]])
        lzz = ((g[2:] - g[1:-1]) / dzt - (g[1:-1] - g[:-2]) / dzb) / dz
        r = sol - rk * 0.5 * prm.dt * prm.nu * (sig * sol + lzz) - src
    return np.abs(r).max() / np.abs(src).max()


def _solve_mode(prm, eq, rk, l, m, prof):
    o = make_oracle(prm, {})
    src = _mode(prm, l, m, prof)
    out_re = o.solve(eq, rk, src.real)
    out_im = o.solve(eq, rk, src.imag)
    i = np.arange(prm.nx)[None, None, :]
    j = np.arange(prm.ny)[None, :, None]
    ph = np.exp(-2j * np.pi * (l * i / prm.nx + m * j / prm.ny))
    got = ((out_re + 1j * out_im) * ph).mean(axis=(1, 2))  # the profile of mode (l, m)
    return got


@pytest.mark.parametrize("wall_rows", [0, 1])
@pytest.mark.parametrize("lm", [(0, 0), (2, 1)])
def test_w_rows_are_crank_nicolson_of_the_explicit_operator(wall_rows, lm):
    prm = cases.closed_channel(nx=16, ny=16, nz=21, zstretch=1.0, wall_rows=wall_rows, dt=5e-2)
    rng = np.random.default_rng(60)
    prof = rng.standard_normal(prm.nz) + 1j * rng.standard_normal(prm.nz)
    rk = 2.0 / 15.0
    sol = _solve_mode(prm, "w", rk, *lm, prof)
    assert abs(sol[-1]) == 0.0  # the top face
    wall = lambda s: (0.0, 0.0)  # w = 0 on the wall faces (src/module_boundary.f90:107-109)
    quirk = lambda s: (s[0], s[-1])  # rows of the modes visited after the first
    consistent = wall_rows == 1 or lm == (0, 0)
    assert _cn_residual(prm, "w", rk, *lm, sol, prof, wall if consistent else quirk) < 1e-13
    if not consistent:
        assert _cn_residual(prm, "w", rk, *lm, sol, prof, wall) > 1e-6


@pytest.mark.parametrize("wall_rows", [0, 1])
@pytest.mark.parametrize("ochannel", [0, 1])
def test_u_rows_are_crank_nicolson_of_the_explicit_operator(wall_rows, ochannel):
    lz = 2.0 if ochannel == 0 else 1.0
    prm = cases.closed_channel(nx=16, ny=16, nz=21, zstretch=1.0, wall_rows=wall_rows, dt=5e-2,
                               ochannel=ochannel, lz=lz, u_mrf=0.0)
    rng = np.random.default_rng(61)
    prof = rng.standard_normal(prm.nz) + 1j * rng.standard_normal(prm.nz)
    rk = 1.0 / 3.0
    l, m = 3, 2
    sol = _solve_mode(prm, "u", rk, l, m, prof)
    # no-slip bottom, no-slip (closed) or free-slip (open) top
    top = (lambda s: -s[-1]) if ochannel == 0 else (lambda s: s[-1])
    wall = lambda s: (-s[0], top(s))
    quirk = lambda s: (s[0], s[-1])
    assert _cn_residual(prm, "u", rk, l, m, sol, prof, wall if wall_rows == 1 else quirk) < 1e-13


# ----------------------------------------------------------------------------
def _substep_state(prm, seed):
    st = cases.laminar_noise(prm, seed=seed)
    st["t"] = 0.3 + 0.2 * np.random.default_rng(seed + 1).standard_normal(st["u"].shape)
    return st


@pytest.mark.parametrize("cds", [1, 2])
def test_temperature_is_advected_by_the_projected_velocity(cds):
    """One AB2 step is one substep: its stored T right-hand side must equal
    the temperature operator applied to (u, v, w after the projection, t
    before the update) (src/module_navier_stokes.f90:319-341)."""
    prm = cases.closed_channel(nx=16, ny=16, nz=17, zstretch=1.0, cds=cds, dts=1, isscalar=1, got=0.2)
    st = _substep_state(prm, 70)
    o = make_oracle(prm, st)
    o.steps(1)
    after = {k: o.get(k) for k in ("u", "v", "w")}
    ss = o.get("ss")
    ref = make_oracle(prm, dict(u=after["u"], v=after["v"], w=after["w"], p=st["p"], t=st["t"]))
    _stage(ref, "ho_stage_temperature", 1)
    assert np.array_equal(ss, ref.get("ss"))
    pre = make_oracle(prm, st)  # the start-of-substep velocity gives another answer
    _stage(pre, "ho_stage_temperature", 1)
    assert rel_l2(pre.get("ss"), ss) > 1e-3


def test_temperature_of_a_driven_flow_from_rest_closed_form():
    """From rest under dp/dx = 1, one AB2 step gives u = 1.5 dt in the
    interior (the projection leaves a horizontally uniform u alone), and the
    T right-hand side of a single x mode is -u (D o I) t with D the 4th-order
    flux difference and I the 4th-order interpolation."""
    prm = cases.closed_channel(nx=16, ny=8, nz=17, zstretch=1.0, dts=1, isscalar=1, got=0.0, tbot=0.0, ttop=0.0)
    sh = (prm.nz, prm.ny, prm.nx)
    l, amp = 2, 0.8 * np.exp(0.3j)
    t0 = np.broadcast_to(_mode(prm, l, 0, np.full(prm.nz, amp)).real, sh).copy()
    z = np.zeros(sh)
    o = make_oracle(prm, dict(u=z, v=z, w=z, p=z, t=t0))
    o.steps(1)
    inner = slice(5, prm.nz - 5)
    U = 1.5 * prm.dt
    assert np.abs(o.get("u")[inner] - U).max() < 1e-13 * U
    dx = prm.lx / prm.nx
    I = _symbol({2: -1 / 16, 1: 9 / 16, 0: 9 / 16, -1: -1 / 16}, prm.nx, l)  # p -> u points
    D = _symbol({1: -1 / 24 / dx, 0: 27 / 24 / dx, -1: -27 / 24 / dx, -2: 1 / 24 / dx}, prm.nx, l)  # u -> p
    want = -U * _mode(prm, l, 0, np.full(prm.nz, amp * I * D)).real
    assert rel_l2(o.get("ss")[inner], np.broadcast_to(want, sh)[inner]) < 1e-12

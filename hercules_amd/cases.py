"""Run configurations and seeded synthetic initial states.

The reference reads its initial state from ``results/init_*.dat`` and can
only add unseeded multiplicative noise (src/module_tools.f90:321-357), so
the seeded initial conditions of the benchmark configurations are built
here: a mean profile plus uniform or Gaussian noise whose amplitude is
limited near the walls so that the explicit wall-normal CFL number of the
noise stays below ``cfl_noise`` (the stretched grids have dz_min ~ 1e-4).
"""
from __future__ import annotations

import numpy as np

from .grid import metrics
from .params import Params


# The grid stretching "factor 2.0" of the benchmark configurations is read
# as the argument of tanh(2*zstretch*...) in get_mesh_param
# (src/module_parameters.f90:464-465), i.e. zstretch = 1.0.
ZSTRETCH = 1.0


def closed_channel(nx=64, ny=64, nz=65, **kw) -> Params:
    """Neutral closed channel, Re_tau = 180, constant mean pressure gradient,
    4th-order horizontal FD (cds=2), RK3, dt = 1e-3, tanh(2 zeta) stretching."""
    base = dict(lx=4.0 * np.pi, ly=2.0 * np.pi, lz=2.0, nx=nx, ny=ny, nz=nz, zstretch=ZSTRETCH,
                ochannel=0, dp_opt=1, cds=2, dts=2, nu=1.0 / 180.0, alpha=1.0 / 180.0 / 0.71,
                dt=1e-3, imax=100, isscalar=0, restart=1, isnoise=0, isdamp=0, u_mrf=0.0)
    base.update(kw)
    return Params(**base)


def stratified_open_channel(nx=256, ny=256, nz=257, **kw) -> Params:
    """Stably stratified open channel (free-slip top, no-slip bottom),
    Re_tau = 395, got = 0.1, Pr = 0.71, 4pi x 2pi x 1."""
    nu = 1.0 / 395.0
    base = dict(lx=4.0 * np.pi, ly=2.0 * np.pi, lz=1.0, nx=nx, ny=ny, nz=nz, zstretch=ZSTRETCH,
                ochannel=1, dp_opt=1, cds=2, dts=2, nu=nu, alpha=nu / 0.71, got=0.1,
                dt=1e-3, imax=500, isscalar=1, tbot=0.0, ttop=1.0, t_ref=0.0, restart=1)
    base.update(kw)
    return Params(**base)


def _taper(dz, amp, dt, cfl_noise):
    return amp * np.minimum(1.0, cfl_noise * dz / (dt * max(amp, 1e-300)))


def laminar_noise(prm: Params, seed=0, rel_amp=0.1, cfl_noise=0.2):
    """Laminar Poiseuille u = z(2-z)/(2 nu) (bulk U_b = 1/(3 nu) for lz = 2)
    plus seeded uniform noise of amplitude rel_amp*U_b on u, v, w."""
    m = metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
    zc, dz = m["zc"], m["dz"]
    shape = (prm.nz, prm.ny, prm.nx)
    ub = prm.lz ** 2 / (12.0 * prm.nu)
    rng = np.random.default_rng(seed)
    amp = _taper(dz, rel_amp * ub, prm.dt, cfl_noise)[:, None, None]
    u = (zc * (prm.lz - zc) / (2.0 * prm.nu))[:, None, None] + amp * rng.uniform(-1, 1, shape)
    v = amp * rng.uniform(-1, 1, shape)
    w = amp * rng.uniform(-1, 1, shape)
    w[-1] = 0.0
    z = np.zeros(shape)
    return dict(u=u - prm.u_mrf, v=v, w=w, p=z.copy(), t=z.copy())


def loglaw_noise(prm: Params, seed=0, rel_amp=0.1, t_amp=0.01, cfl_noise=0.2):
    """Reichardt mean profile in wall units (u_tau = 1, nu) plus seeded
    Gaussian noise: rel_amp of the local mean on the velocity, t_amp of the
    wall temperature difference on T (linear conduction profile)."""
    m = metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
    zc, dz = m["zc"], m["dz"]
    shape = (prm.nz, prm.ny, prm.nx)
    # distance to the nearest no-slip wall
    d = zc if prm.ochannel == 1 else np.minimum(zc, prm.lz - zc)
    yp = d / prm.nu
    up = np.log(1.0 + 0.4 * yp) / 0.41 + 7.8 * (1.0 - np.exp(-yp / 11.0) - yp / 11.0 * np.exp(-yp / 3.0))
    rng = np.random.default_rng(seed)
    amp = _taper(dz, rel_amp * up.max(), prm.dt, cfl_noise)[:, None, None]
    u = up[:, None, None] + amp * rng.standard_normal(shape)
    v = amp * rng.standard_normal(shape)
    w = amp * rng.standard_normal(shape)
    w[-1] = 0.0
    t0 = prm.tbot + (prm.ttop - prm.tbot) * zc / prm.lz
    t = t0[:, None, None] + t_amp * abs(prm.ttop - prm.tbot) * rng.standard_normal(shape)
    return dict(u=u - prm.u_mrf, v=v, w=w, p=np.zeros(shape), t=t)


def gaussian_poisson_rhs(prm: Params, seed=0):
    """Seeded unit-variance Gaussian field with the plane mean removed on
    every level (the compatibility condition of the Neumann problem)."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((prm.nz, prm.ny, prm.nx))
    return f - f.mean(axis=(1, 2), keepdims=True)


def wall_shear(u, prm: Params):
    """Plane-averaged wall shear stress nu * du/dz at the no-slip walls from
    the ghost rule u(kstr3-1) = -u(kstr3) - 2 u_mrf (src/module_boundary.f90:52-53)
    and dz_b(kstr3) = dz(kstr3): bottom 2 nu <u_1 + u_mrf>/dz_1, top likewise
    (closed channel only)."""
    m = metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
    dz = m["dz"]
    bot = 2.0 * prm.nu * (u[0].mean() + prm.u_mrf) / dz[0]
    top = 2.0 * prm.nu * (u[-1].mean() + prm.u_mrf) / dz[-1] if prm.ochannel == 0 else 0.0
    return bot, top


def ensemble_24(seed=2024):
    """The 24-member ensemble: eight each at 128x128x129, 192x192x129 and
    256x128x193; neutral closed channels, stratified open channels and
    Ekman-type open channels (Coriolis drive with a top sponge), Re_tau
    seeded uniform in [150, 400] and Ri uniform in [0, 0.2].  Returns
    (params, state) pairs."""
    rng = np.random.default_rng(seed)
    shapes = [(128, 128, 129)] * 8 + [(192, 192, 129)] * 8 + [(256, 128, 193)] * 8
    out = []
    for q, (nx, ny, nz) in enumerate(shapes):
        re_tau = rng.uniform(150.0, 400.0)
        ri = rng.uniform(0.0, 0.2)
        kind = q % 3
        common = dict(nx=nx, ny=ny, nz=nz, nu=1.0 / re_tau, alpha=1.0 / re_tau / 0.71, wall_rows=1,
                      dt=5e-4)
        if kind == 0:
            prm = closed_channel(**common)
        elif kind == 1:
            prm = stratified_open_channel(got=ri, **common)
        else:
            prm = stratified_open_channel(got=ri, dp_opt=3, fc=0.5, ug=1.0, vg=0.0, isdamp=1,
                                          nzdamp=nz // 8, cdamp=0.01, **common)
        st = loglaw_noise(prm, seed=seed + q)
        if not prm.isscalar:
            st.pop("t")
        out.append((prm, st))
    return out

"""FP64 restatement of the finite-difference RK3 step in torch operations —
test infrastructure only, for small grids.

It exists to be differentiated: torch.autograd through this forward gives
a reference gradient for the library's hand-written adjoint.  It follows the
reference's formulas (src/module_cnvdiff.f90:35-374 for the right-hand side,
src/module_poisson_solver.f90 and src/module_navier_stokes.f90:247-355,
523-808 for the solves, projection and pressure update) in their physical
form: explicit stencils with torch.roll and wall ghosts, implicit z operators
as dense per-mode matrices after a 2-D FFT.  Scope: cds = 1 or 2, no
damping, dp_opt = 1, u_mrf = 0, wall_rows = 1 (every mode gets the wall
rows), RK3; temperature (get_temperature_cnvdiff, src/module_cnvdiff.f90:
381-473, with the Dirichlet or constant-gradient top) without the Ri ramp.

With prm.zorder = 6 every z derivative, z interpolation and implicit z
operator is the dense 6th-order matrix of tests/zops6.py (the library's
zorder = 6 extension; the reference is 2nd order in z).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from hercules_amd import grid
from tests.zops6 import zops

RK3 = [(8.0 / 15.0, 0.0, 8.0 / 15.0), (5.0 / 12.0, -17.0 / 60.0, 2.0 / 15.0), (3.0 / 4.0, -5.0 / 12.0, 1.0 / 3.0)]


class TorchChannel:
    def __init__(self, prm):
        assert prm.cds in (1, 2) and prm.isdamp == 0 and prm.dp_opt == 1 and prm.is_ri_var == 0
        assert prm.u_mrf == 0.0 and prm.wall_rows == 1 and prm.dts == 2
        self.p = prm
        m = grid.metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
        t = lambda a: torch.tensor(a, dtype=torch.float64)
        self.dz, self.dzt, self.dzb, self.lt = t(m["dz"]), t(m["dz_t"]), t(m["dz_b"]), t(m["l_t"])
        self.dx, self.dy = prm.lx / prm.nx, prm.ly / prm.ny
        self.z6 = getattr(prm, "zorder", 2) in (4, 6)
        if self.z6:
            self.Z = {k: t(v) for k, v in zops(prm.nz, prm.lz, prm.zstretch, prm.ochannel, prm.zorder).items()}
        self._symbols()

    # ---- horizontal stencils (axis 2 = x, axis 1 = y) ---------------------
    @staticmethod
    def sh(f, d, ax):  # f(i + d) along axis ax
        return torch.roll(f, -d, dims=ax)

    def im(self, f, ax):  # face between i-1 and i
        if self.p.cds == 1:
            return 0.5 * (f + self.sh(f, -1, ax))
        return (-self.sh(f, 1, ax) + 9 * f + 9 * self.sh(f, -1, ax) - self.sh(f, -2, ax)) / 16

    def ip(self, f, ax):  # face between i and i+1
        if self.p.cds == 1:
            return 0.5 * (self.sh(f, 1, ax) + f)
        return (-self.sh(f, 2, ax) + 9 * self.sh(f, 1, ax) + 9 * f - self.sh(f, -1, ax)) / 16

    def d_s(self, g, ax, h):  # derivative of a face quantity at offsets +1, 0 (+2, -1)
        if self.p.cds == 1:
            return (self.sh(g, 1, ax) - g) / h
        return (-self.sh(g, 2, ax) + 27 * self.sh(g, 1, ax) - 27 * g + self.sh(g, -1, ax)) / (24 * h)

    def d_p(self, g, ax, h):  # derivative of a face quantity at offsets 0, -1 (+1, -2)
        if self.p.cds == 1:
            return (g - self.sh(g, -1, ax)) / h
        return (-self.sh(g, 1, ax) + 27 * g - 27 * self.sh(g, -1, ax) + self.sh(g, -2, ax)) / (24 * h)

    def lap_h(self, f):
        out = 0
        for ax, h in ((2, self.dx), (1, self.dy)):
            if self.p.cds == 1:
                out = out + (self.sh(f, 1, ax) - 2 * f + self.sh(f, -1, ax)) / h ** 2
            else:
                out = out + (-self.sh(f, 2, ax) + 16 * self.sh(f, 1, ax) - 30 * f + 16 * self.sh(f, -1, ax)
                             - self.sh(f, -2, ax)) / (12 * h ** 2)
        return out

    # ---- z with wall ghosts -------------------------------------------------
    def up(self, f, top):  # f(k+1) with the top ghost
        return torch.cat([f[1:], top[None]], 0)

    def dn(self, f, bot):  # f(k-1) with the bottom ghost
        return torch.cat([bot[None], f[:-1]], 0)

    def col(self, a):
        return a[:, None, None]

    # ---- symbols and z matrices for the implicit solves ------------------
    def _symbols(self):
        p = self.p
        nx, ny = p.nx, p.ny

        def sym(coefs, n):  # sum_s c_s e^{i s theta} for each wavenumber index
            m = np.fft.fftfreq(n, 1.0 / n)
            th = 2 * math.pi * m / n
            return sum(c * np.exp(1j * s * th) for s, c in coefs)

        if p.cds == 1:
            lap = [(1, 1.0), (0, -2.0), (-1, 1.0)]
            dv = [(0, 1.0), (-1, -1.0)]        # divergence at p points from u faces
            gr = [(1, 1.0), (0, -1.0)]         # gradient at u faces from p
        else:
            lap = [(2, -1 / 12), (1, 16 / 12), (0, -30 / 12), (-1, 16 / 12), (-2, -1 / 12)]
            dv = [(1, -1 / 24), (0, 27 / 24), (-1, -27 / 24), (-2, 1 / 24)]
            gr = [(2, -1 / 24), (1, 27 / 24), (0, -27 / 24), (-1, 1 / 24)]
        lx = sym(lap, nx) / self.dx ** 2
        ly = sym(lap, ny) / self.dy ** 2
        dx_ = sym(dv, nx) / self.dx
        dy_ = sym(dv, ny) / self.dy
        gx_ = sym(gr, nx) / self.dx
        gy_ = sym(gr, ny) / self.dy
        c = lambda a: torch.tensor(a, dtype=torch.complex128)
        self.lap_xy = c(ly[:, None] + lx[None, :]).real.to(torch.float64)          # (ny, nx)
        self.pois_xy = (c(dy_[:, None] * gy_[:, None]) + c(dx_[None, :] * gx_[None, :])).real
        # Nyquist wavenumbers are removed by the transforms
        keep = np.ones((ny, nx))
        keep[ny // 2, :] = 0.0
        keep[:, nx // 2] = 0.0
        self.keep = torch.tensor(keep, dtype=torch.float64)
        nz = p.nz
        dz, dzt, dzb = self.dz.numpy(), self.dzt.numpy(), self.dzb.numpy()
        # cell-centred z Laplacian without wall fluxes (the Neumann operator
        # of p); u, v add the no-slip ghost -f at the bottom and, closed, at
        # the top (free slip adds nothing)
        L = np.zeros((nz, nz))
        for k in range(nz):
            if k > 0:
                L[k, k - 1] += 1 / (dz[k] * dzb[k])
                L[k, k] -= 1 / (dz[k] * dzb[k])
            if k < nz - 1:
                L[k, k + 1] += 1 / (dz[k] * dzt[k])
                L[k, k] -= 1 / (dz[k] * dzt[k])
        Lu = L.copy()
        Lu[0, 0] -= 2 / (dz[0] * dzb[0])
        if p.ochannel == 0:
            Lu[-1, -1] -= 2 / (dz[-1] * dzt[-1])
        Lp = L
        # w on faces 0..nz-2, w = 0 on the walls
        Lw = np.zeros((nz, nz))
        for k in range(nz - 1):
            if k > 0:
                Lw[k, k - 1] = 1 / (dz[k] * dzt[k])
            if k < nz - 2:
                Lw[k, k + 1] = 1 / (dz[k + 1] * dzt[k])
            Lw[k, k] = -(1 / (dz[k + 1] * dzt[k]) + 1 / (dz[k] * dzt[k]))
        # t: Dirichlet bottom (ghost 2 tbot - t), top Dirichlet or constant gradient
        Lt = L.copy()
        Lt[0, 0] -= 2 / (dz[0] * dzb[0])
        bt = np.zeros(nz)
        bt[0] = 2 * p.tbot / (dz[0] * dzb[0])
        if p.tbc_top == 0:
            Lt[-1, -1] -= 2 / (dz[-1] * dzt[-1])
            bt[-1] = 2 * p.ttop / (dz[-1] * dzt[-1])
        else:
            bt[-1] = p.dtdz_top * dzt[-1] / (dz[-1] * dzt[-1])
        t = lambda a: torch.tensor(a, dtype=torch.float64)
        self.Lu, self.Lp, self.Lw, self.Lt, self.bt = t(Lu), t(Lp), t(Lw), t(Lt), t(bt)
        if self.z6:
            Z = self.Z
            c2u = Z["C2_DN"] if p.ochannel else Z["C2_DD"]
            c2t = Z["C2_DN"] if p.tbc_top else Z["C2_DD"]
            self.Lu = c2u[:, 1:-1]
            Lw6 = torch.zeros((nz, nz), dtype=torch.float64)
            Lw6[:-1, :-1] = Z["F2"][:, 1:-1]
            self.Lw = Lw6
            self.Lp = Z["P"]
            self.Lt = c2t[:, 1:-1]
            self.bt = c2t[:, 0] * p.tbot + c2t[:, -1] * self.t_top_datum()

    def t_top_datum(self):  # the top wall value, or the gradient
        return self.p.dtdz_top if self.p.tbc_top else self.p.ttop

    @staticmethod
    def zmul(M, f):  # z operator on (levels, ny, nx)
        return torch.tensordot(M, f, dims=([1], [0]))

    def fft(self, f):
        return torch.fft.fft2(f) / (self.p.nx * self.p.ny) * self.keep

    def ifft(self, F):
        return torch.fft.ifft2(F * self.keep).real * (self.p.nx * self.p.ny)

    def helm(self, q, Lz, beta, trivial_top=False):
        """(I - beta (Lz + lap_xy)) x = q, per mode (wall_rows = 1 rows)."""
        nz = self.p.nz
        Q = self.fft(q).permute(1, 2, 0)  # (ny, nx, nz)
        eye = torch.eye(nz, dtype=torch.float64)
        M = eye - beta * (Lz[None, None] + self.lap_xy[:, :, None, None] * eye)
        if trivial_top:
            M = M.clone()
            M[:, :, -1, :] = 0.0
            M[:, :, -1, -1] = 1.0
            Q = torch.cat([Q[..., :-1], torch.zeros_like(Q[..., -1:])], -1)
        X = torch.linalg.solve(M.to(torch.complex128), Q.unsqueeze(-1)).squeeze(-1)
        return self.ifft(X.permute(2, 0, 1))

    def poisson(self, s):
        """(Lz + pois_xy) phi = s with phi = 0 on level 0 of the (0,0) mode."""
        nz = self.p.nz
        S = self.fft(s).permute(1, 2, 0)
        eye = torch.eye(nz, dtype=torch.float64)
        M = (self.Lp[None, None] + self.pois_xy[:, :, None, None] * eye).to(torch.complex128)
        M = M.clone()
        M[0, 0, 0, :] = 0.0
        M[0, 0, 0, 0] = 1.0
        S = S.clone()
        S[0, 0, 0] = 0.0
        X = torch.linalg.solve(M, S.unsqueeze(-1)).squeeze(-1)
        return self.ifft(X.permute(2, 0, 1))

    # ---- one RK3 substep ----------------------------------------------------
    def substep(self, s, t_its):
        p = self.p
        ra, rb, rc = RK3[t_its - 1]
        u, v, w, pr = s["u"], s["v"], s["w"], s["p"]
        dx, dy, dt, nu = self.dx, self.dy, p.dt, p.nu
        col = self.col
        nz = p.nz
        zero = torch.zeros_like(u[0])
        top_u = -u[-1] if p.ochannel == 0 else u[-1]
        top_v = -v[-1] if p.ochannel == 0 else v[-1]
        wz = torch.cat([w[:-1], zero[None]], 0)  # w with w(kend3) = 0
        # interpolants (src/module_tools.f90:410-468, wall rules of update_boundary_uvwhxyz)
        uhx, uhy = self.im(u, 2), self.ip(u, 1)
        vhx, vhy = self.ip(v, 2), self.im(v, 1)
        whx, why = self.ip(wz, 2), self.ip(wz, 1)
        lt = col(self.lt)
        uhz = self.up(u, top_u) * lt + u * (1 - lt)
        vhz = self.up(v, top_v) * lt + v * (1 - lt)
        # faces: uhz(kend3) = -u_mrf (closed) or u (open); whx, why = 0 there
        uhz = torch.cat([uhz[:-1], (zero if p.ochannel == 0 else u[-1])[None]], 0)
        vhz = torch.cat([vhz[:-1], (zero if p.ochannel == 0 else v[-1])[None]], 0)
        idz = col(1 / self.dz)
        idzt = col(1 / self.dzt)
        tt = s.get("t")
        if p.isscalar:  # buoyancy at the w faces from the old t
            if self.z6:
                tdat = torch.cat([torch.full_like(zero, p.tbot)[None], tt,
                                  torch.full_like(zero, self.t_top_datum())[None]], 0)
                ti = self.Z["I_DN"] if p.tbc_top else self.Z["I_DD"]
                thz_i = self.zmul(ti, tdat)
            else:
                thz_i = ((1 - lt) * tt + lt * self.up(tt, tt[-1]))[:-1]
            buoy = p.got * (thz_i - p.t_ref)
        else:
            buoy = 0.0
        if self.z6:
            Z = self.Z
            c2u = Z["C2_DN"] if p.ochannel else Z["C2_DD"]
            ciu = Z["I_DN"] if p.ochannel else Z["I_DD"]
            udat = torch.cat([zero[None], u, zero[None]], 0)   # u_mrf = 0: wall value 0 (du/dz = 0 open top)
            vdat = torch.cat([zero[None], v, zero[None]], 0)
            wdat = torch.cat([zero[None], wz[:-1], zero[None]], 0)
            uhz_i, vhz_i = self.zmul(ciu, udat), self.zmul(ciu, vdat)
            whx_i, why_i = whx[:-1], why[:-1]
            fuw = torch.cat([zero[None], uhz_i * whx_i, zero[None]], 0)
            fvw = torch.cat([zero[None], vhz_i * why_i, zero[None]], 0)
            f = -self.d_s(uhx * uhx, 2, dx) - self.d_p(uhy * vhx, 1, dy) - self.zmul(Z["D"], fuw) + 1.0
            g = -self.d_p(vhx * uhy, 2, dx) - self.d_s(vhy * vhy, 1, dy) - self.zmul(Z["D"], fvw)
            whz = self.zmul(Z["IFC"], wdat)
            hi = (-self.d_p(whx_i * uhz_i, 2, dx) - self.d_p(why_i * vhz_i, 1, dy)
                  - self.zmul(Z["G"], whz * whz) + buoy)
            h = torch.cat([hi, zero[None]], 0)
            uzz, vzz = self.zmul(c2u, udat), self.zmul(c2u, vdat)
            wzz = torch.cat([self.zmul(Z["F2"], wdat), zero[None]], 0)
            gzp = torch.cat([self.zmul(Z["G"], pr), zero[None]], 0)
        else:
            fuw = uhz * whx
            fvw = vhz * why
            f = (-self.d_s(uhx * uhx, 2, dx) - self.d_p(uhy * vhx, 1, dy)
                 - (fuw - self.dn(fuw, zero)) * idz + 1.0)  # dp_opt = 1: dpx_drive = 1
            g = (-self.d_p(vhx * uhy, 2, dx) - self.d_s(vhy * vhy, 1, dy) - (fvw - self.dn(fvw, zero)) * idz)
            # w equation on faces 0..nz-2
            wp, wm = self.up(wz, zero), self.dn(wz, zero)
            h = (-self.d_p(whx * uhz, 2, dx) - self.d_p(why * vhz, 1, dy)
                 - ((0.5 * (wz + wp)) ** 2 - (0.5 * (wz + wm)) ** 2) * idzt)
            if p.isscalar:
                h = h + torch.cat([buoy, zero[None]], 0)
            # explicit diffusion and pressure gradients
            uzz = ((self.up(u, top_u) - u) / col(self.dzt) - (u - self.dn(u, -u[0])) / col(self.dzb)) * idz
            vzz = ((self.up(v, top_v) - v) / col(self.dzt) - (v - self.dn(v, -v[0])) / col(self.dzb)) * idz
            dzn = torch.cat([self.dz[1:], self.dz[-1:]])
            wzz = ((wp - wz) / col(dzn) - (wz - wm) * idz) * idzt
            gzp = (self.up(pr, pr[-1]) - pr) * idzt
        gxp = self.d_s(pr, 2, dx)
        gyp = self.d_s(pr, 1, dy)
        f0, g0, h0 = s["f"], s["g"], s["h"]
        cn = rc * 0.5 * dt * nu
        qu = u + dt * (ra * f + rb * f0) + cn * (self.lap_h(u) + uzz) - rc * dt * gxp
        qv = v + dt * (ra * g + rb * g0) + cn * (self.lap_h(v) + vzz) - rc * dt * gyp
        qw = wz + dt * (ra * h + rb * h0) + cn * (self.lap_h(wz) + wzz) - rc * dt * gzp
        # Helmholtz solves
        beta = rc * 0.5 * dt * nu
        us = self.helm(qu, self.Lu, beta)
        vs = self.helm(qv, self.Lu, beta)
        ws = self.helm(qw, self.Lw, beta, trivial_top=True)
        # divergence (get_p_eqn_src), pressure Poisson, projection, new p
        if self.z6:
            dwz = self.zmul(self.Z["D"][:, 1:-1], ws[:-1])
        else:
            dwz = (ws - self.dn(ws, zero)) * idz
        div = (self.d_p(us, 2, dx) + self.d_p(vs, 1, dy) + dwz) / dt
        phi = self.poisson(div / rc)
        cdt = rc * dt
        un = us - cdt * self.d_s(phi, 2, dx)
        vn = vs - cdt * self.d_s(phi, 1, dy)
        if self.z6:
            gz = torch.cat([self.zmul(self.Z["G"], phi), zero[None]], 0)
            lz = self.zmul(self.Z["P"], phi)
        else:
            gz = (self.up(phi, phi[-1]) - phi) * idzt
            lz = ((self.up(phi, phi[-1]) - phi) / col(self.dzt) - (phi - self.dn(phi, phi[0])) / col(self.dzb)) * idz
        wn = torch.cat([(ws - cdt * gz)[:-1], zero[None]], 0)
        pn = pr + phi - cn * (self.lap_h(phi) + lz)
        out = dict(u=un, v=vn, w=wn, p=pn, f=f, g=g, h=torch.cat([h[:-1], zero[None]], 0))
        if p.isscalar:
            out["t"], out["s"] = self.temperature(tt, s["s"], un, vn, wn, ra, rb, rc)
        return out

    def temperature(self, t, s0, u, v, w, ra, rb, rc):
        """get_temperature_cnvdiff with the updated velocities, then the t
        Helmholtz solve with the wall data in the (0,0) mode."""
        p = self.p
        dx, dy, dt = self.dx, self.dy, p.dt
        col = self.col
        zero = torch.zeros_like(t[0])
        tux = self.d_p(self.ip(t, 2) * u, 2, dx)
        tvy = self.d_p(self.ip(t, 1) * v, 1, dy)
        if self.z6:
            tdat = torch.cat([torch.full_like(zero, p.tbot)[None], t,
                              torch.full_like(zero, self.t_top_datum())[None]], 0)
            ti = self.Z["I_DN"] if p.tbc_top else self.Z["I_DD"]
            c2t = self.Z["C2_DN"] if p.tbc_top else self.Z["C2_DD"]
            ft = torch.cat([zero[None], self.zmul(ti, tdat) * w[:-1], zero[None]], 0)
            twz = self.zmul(self.Z["D"], ft)
            tzz = self.zmul(c2t, tdat)
        else:
            lt = col(self.lt)
            gdz = p.dtdz_top * float(self.dzt[-1])
            ttop_g = t[-1] + gdz if p.tbc_top else 2 * p.ttop - t[-1]
            tbot_g = 2 * p.tbot - t[0]
            thz = self.up(t, ttop_g) * lt + t * (1 - lt)
            top_face = t[-1] + 0.5 * gdz if p.tbc_top else torch.full_like(zero, p.ttop)
            thz = torch.cat([thz[:-1], top_face[None]], 0)
            ft = thz * w
            twz = (ft - self.dn(ft, zero)) / col(self.dz)
            tzz = ((self.up(t, ttop_g) - t) / col(self.dzt) - (t - self.dn(t, tbot_g)) / col(self.dzb)) / col(self.dz)
        snew = -tux - tvy - twz
        ct = rc * 0.5 * dt * p.alpha
        qt = t + dt * (ra * snew + rb * s0) + ct * (self.lap_h(t) + tzz) + ct * col(self.bt)
        return self.helm(qt, self.Lt, ct), snew

    def steps(self, st, n):
        s = dict(st)
        z = torch.zeros_like(s["u"])
        s.setdefault("f", z)
        s.setdefault("g", z)
        s.setdefault("h", z)
        if self.p.isscalar:
            s.setdefault("s", z)
        for _ in range(n):
            for t in (1, 2, 3):
                s = self.substep(s, t)
        return s

    def wall_shear(self, s):
        return 2.0 * self.p.nu * s["u"][0].mean() / self.dz[0]

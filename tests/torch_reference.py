"""FP64 restatement of the finite-difference RK3 step in torch operations —
test infrastructure only, for small grids.

It exists to be differentiated: torch.autograd through this forward gives
a reference gradient for the library's hand-written adjoint.  It follows the
reference's formulas (src/module_cnvdiff.f90:35-374 for the right-hand side,
src/module_poisson_solver.f90 and src/module_navier_stokes.f90:247-355,
523-808 for the solves, projection and pressure update) in their physical
form: explicit stencils with torch.roll and wall ghosts, implicit z operators
as dense per-mode matrices after a 2-D FFT.  Scope: cds = 1 or 2, no
temperature, no damping, dp_opt = 1, u_mrf = 0, wall_rows = 1 (every mode
gets the wall rows), RK3.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from hercules_amd import grid

RK3 = [(8.0 / 15.0, 0.0, 8.0 / 15.0), (5.0 / 12.0, -17.0 / 60.0, 2.0 / 15.0), (3.0 / 4.0, -5.0 / 12.0, 1.0 / 3.0)]


class TorchChannel:
    def __init__(self, prm):
        assert prm.cds in (1, 2) and prm.isscalar == 0 and prm.isdamp == 0 and prm.dp_opt == 1
        assert prm.u_mrf == 0.0 and prm.wall_rows == 1 and prm.dts == 2
        self.p = prm
        m = grid.metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
        t = lambda a: torch.tensor(a, dtype=torch.float64)
        self.dz, self.dzt, self.dzb, self.lt = t(m["dz"]), t(m["dz_t"]), t(m["dz_b"]), t(m["l_t"])
        self.dx, self.dy = prm.lx / prm.nx, prm.ly / prm.ny
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
        t = lambda a: torch.tensor(a, dtype=torch.float64)
        self.Lu, self.Lp, self.Lw = t(Lu), t(Lp), t(Lw)

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
        fuw = uhz * whx
        fvw = vhz * why
        f = (-self.d_s(uhx * uhx, 2, dx) - self.d_p(uhy * vhx, 1, dy)
             - (fuw - self.dn(fuw, zero)) * idz + 1.0)  # dp_opt = 1: dpx_drive = 1
        g = (-self.d_p(vhx * uhy, 2, dx) - self.d_s(vhy * vhy, 1, dy) - (fvw - self.dn(fvw, zero)) * idz)
        # w equation on faces 0..nz-2
        wp, wm = self.up(wz, zero), self.dn(wz, zero)
        idzt = col(1 / self.dzt)
        h = (-self.d_p(whx * uhz, 2, dx) - self.d_p(why * vhz, 1, dy)
             - ((0.5 * (wz + wp)) ** 2 - (0.5 * (wz + wm)) ** 2) * idzt)
        # explicit diffusion and pressure gradients
        uzz = ((self.up(u, top_u) - u) / col(self.dzt) - (u - self.dn(u, -u[0])) / col(self.dzb)) * idz
        vzz = ((self.up(v, top_v) - v) / col(self.dzt) - (v - self.dn(v, -v[0])) / col(self.dzb)) * idz
        dzn = torch.cat([self.dz[1:], self.dz[-1:]])
        wzz = ((wp - wz) / col(dzn) - (wz - wm) * idz) * idzt
        gxp = self.d_s(pr, 2, dx)
        gyp = self.d_s(pr, 1, dy)
        gzp = (self.up(pr, pr[-1]) - pr) * idzt
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
        div = (self.d_p(us, 2, dx) + self.d_p(vs, 1, dy) + (ws - self.dn(ws, zero)) * idz) / dt
        phi = self.poisson(div / rc)
        cdt = rc * dt
        un = us - cdt * self.d_s(phi, 2, dx)
        vn = vs - cdt * self.d_s(phi, 1, dy)
        gz = (self.up(phi, phi[-1]) - phi) * idzt
        wn = torch.cat([(ws - cdt * gz)[:-1], zero[None]], 0)
        lz = ((self.up(phi, phi[-1]) - phi) / col(self.dzt) - (phi - self.dn(phi, phi[0])) / col(self.dzb)) * idz
        pn = pr + phi - cn * (self.lap_h(phi) + lz)
        return dict(u=un, v=vn, w=wn, p=pn, f=f, g=g, h=torch.cat([h[:-1], zero[None]], 0))

    def steps(self, st, n):
        s = dict(st)
        z = torch.zeros_like(s["u"])
        s.setdefault("f", z)
        s.setdefault("g", z)
        s.setdefault("h", z)
        for _ in range(n):
            for t in (1, 2, 3):
                s = self.substep(s, t)
        return s

    def wall_shear(self, s):
        return 2.0 * self.p.nu * s["u"][0].mean() / self.dz[0]

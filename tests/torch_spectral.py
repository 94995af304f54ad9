"""FP64 restatement of the spectral RK3 substep (time_advancement_sp, cds = 3)
in torch operations — test infrastructure only, for small grids.

It follows src/module_navier_stokes.f90:371-517 and src/module_spectral.f90
in their spectral form: the nonlinear products formed in physical space
(dev_fft_nonlinear_cmplx, :1075-1145, with the skew-symmetric advective
terms of :142-197), transformed, differentiated by i k; the z terms 2nd
order in physical space (or, with prm.zorder = 6, the dense 6th-order
operators of tests/zops6.py); implicit z solves as dense per-mode matrices;
the divergence, pressure, projection and pressure update of get_div_spec,
calculate_new_p_spec, calculate_uvw_spec (:752-894); the 2/3-rule
truncation (:1222-1268).  Full complex 2-D spectra (normalised by nx ny,
Nyquist zeroed) stand for the half spectra of the library.  Scope: no
damping, u_mrf = 0, wall_rows = 1, RK3, no Ri ramp.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from hercules_amd import grid
from tests.torch_reference import RK3
from tests.zops6 import zops


class TorchSpectral:
    def __init__(self, prm):
        assert prm.cds == 3 and prm.isdamp == 0 and prm.u_mrf == 0.0 and prm.wall_rows == 1
        assert prm.dts == 2 and prm.is_ri_var == 0
        self.p = prm
        nx, ny, nz = prm.nx, prm.ny, prm.nz
        m = grid.metrics(nz, prm.lz, prm.zstretch, prm.ochannel)
        t = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64)
        self.dz, self.dzt, self.dzb, self.lt, self.lb = (t(m[k]) for k in ("dz", "dz_t", "dz_b", "l_t", "l_b"))
        self.z6 = getattr(prm, "zorder", 2) in (4, 6)
        if self.z6:
            self.Z = {k: t(v) for k, v in zops(nz, prm.lz, prm.zstretch, prm.ochannel, prm.zorder).items()}
        mx = np.fft.fftfreq(nx, 1.0 / nx)
        my = np.fft.fftfreq(ny, 1.0 / ny)
        kx = 2 * math.pi / prm.lx * mx
        ky = 2 * math.pi / prm.ly * my
        kx[nx // 2] = 0.0
        ky[ny // 2] = 0.0
        self.kx = t(kx)[None, None, :]
        self.ky = t(ky)[None, :, None]
        self.k2 = (self.kx ** 2 + self.ky ** 2)[0]
        keep = np.ones((ny, nx))
        keep[ny // 2, :] = 0.0
        keep[:, nx // 2] = 0.0
        self.keep = t(keep)
        lc, mc = int(nx / 3.0), int(ny / 3.0)
        kept = (np.abs(mx)[None, :] < lc) & (np.abs(my)[:, None] < mc)
        self.kept = t(kept * keep)
        self.coe = 0.5 if prm.issk == 1 else 1.0
        self._zmats()

    # ---- transforms -------------------------------------------------------
    def F(self, f):
        return torch.fft.fft2(f) / (self.p.nx * self.p.ny) * self.keep

    def Fi(self, c):
        return torch.fft.ifft2(c * self.keep).real * (self.p.nx * self.p.ny)

    @staticmethod
    def col(a):
        return a[:, None, None]

    @staticmethod
    def zmul(M, f):
        return torch.tensordot(M.to(f.dtype), f, dims=([1], [0]))

    # ---- z operators --------------------------------------------------------
    def _zmats(self):
        p, nz = self.p, self.p.nz
        dz, dzt, dzb = self.dz.numpy(), self.dzt.numpy(), self.dzb.numpy()
        t = lambda a: torch.tensor(a, dtype=torch.float64)
        if self.z6:
            Z = self.Z
            c2u = Z["C2_DN"] if p.ochannel else Z["C2_DD"]
            c2t = Z["C2_DN"] if p.tbc_top else Z["C2_DD"]
            Lw = torch.zeros((nz, nz), dtype=torch.float64)
            Lw[:-1, :-1] = Z["F2"][:, 1:-1]
            self.Lu, self.Lw, self.Lp, self.Lt = c2u[:, 1:-1], Lw, Z["P"], c2t[:, 1:-1]
            self.bt = c2t[:, 0] * p.tbot + c2t[:, -1] * (p.dtdz_top if p.tbc_top else p.ttop)
            return
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
        Lw = np.zeros((nz, nz))
        for k in range(nz - 1):
            if k > 0:
                Lw[k, k - 1] = 1 / (dz[k] * dzt[k])
            if k < nz - 2:
                Lw[k, k + 1] = 1 / (dz[k + 1] * dzt[k])
            Lw[k, k] = -(1 / (dz[k + 1] * dzt[k]) + 1 / (dz[k] * dzt[k]))
        Lt = L.copy()
        Lt[0, 0] -= 2 / (dz[0] * dzb[0])
        bt = np.zeros(nz)
        bt[0] = 2 * p.tbot / (dz[0] * dzb[0])
        if p.tbc_top == 0:
            Lt[-1, -1] -= 2 / (dz[-1] * dzt[-1])
            bt[-1] = 2 * p.ttop / (dz[-1] * dzt[-1])
        else:
            bt[-1] = p.dtdz_top / dz[-1]
        self.Lu, self.Lw, self.Lp, self.Lt, self.bt = t(Lu), t(Lw), t(L), t(Lt), t(bt)

    def helm(self, Q, Lz, beta, trivial_top=False):
        """(I - beta (Lz - k^2)) X = Q per mode."""
        nz = self.p.nz
        Qm = Q.permute(1, 2, 0)
        eye = torch.eye(nz, dtype=torch.float64)
        M = eye - beta * (Lz[None, None] - self.k2[:, :, None, None] * eye)
        M = M.to(torch.complex128)
        if trivial_top:
            M = M.clone()
            M[:, :, -1, :] = 0.0
            M[:, :, -1, -1] = 1.0
            Qm = torch.cat([Qm[..., :-1], torch.zeros_like(Qm[..., -1:])], -1)
        return torch.linalg.solve(M, Qm.unsqueeze(-1)).squeeze(-1).permute(2, 0, 1)

    def poisson(self, S):
        """(Lp - k^2) phi = S with phi = 0 on level 0 of the (0,0) mode."""
        nz = self.p.nz
        Sm = S.permute(1, 2, 0).clone()
        eye = torch.eye(nz, dtype=torch.float64)
        M = (self.Lp[None, None] - self.k2[:, :, None, None] * eye).to(torch.complex128).clone()
        M[0, 0, 0, :] = 0.0
        M[0, 0, 0, 0] = 1.0
        Sm[0, 0, 0] = 0.0
        return torch.linalg.solve(M, Sm.unsqueeze(-1)).squeeze(-1).permute(2, 0, 1)

    def dzz(self, c, Lz):  # explicit z second derivative (homogeneous wall data)
        return self.zmul(Lz, c)

    # ---- start-up (src/module_navier_stokes.f90:96-123) ------------------------
    def init(self, st):
        s = {}
        for k in ("u", "v", "w", "p") + (("t",) if self.p.isscalar else ()):
            s["c" + k] = self.F(torch.tensor(st[k])) * self.kept
        z = torch.zeros_like(s["cu"])
        s["cff"], s["cgg"], s["chh"] = z, z.clone(), z.clone()
        if self.p.isscalar:
            s["css"] = z.clone()
        return s

    def phys(self, s, name):
        return self.Fi(s["c" + name])

    # ---- one substep --------------------------------------------------------
    def products(self, u, v, w, t):
        p = self.p
        nz = p.nz
        col = self.col
        zero = torch.zeros_like(u[0])
        coe = self.coe
        D = None
        if p.issk:
            cu, cv, cw = self.F(u), self.F(v), self.F(w)
            ik = lambda c, k: 1j * k * c
            D = [self.Fi(ik(cu, self.kx)), self.Fi(ik(cu, self.ky)), self.Fi(ik(cv, self.kx)),
                 self.Fi(ik(cv, self.ky)), self.Fi(ik(cw, self.kx)), self.Fi(ik(cw, self.ky))]
        if self.z6:
            Z = self.Z
            ci = Z["I_DN"] if p.ochannel else Z["I_DD"]
            ct_ = Z["I_DN_top"] if p.ochannel else Z["I_DD_top"]
            udat = torch.cat([zero[None], u, zero[None]], 0)
            vdat = torch.cat([zero[None], v, zero[None]], 0)
            # faces kw = -1..nz-1
            uf = torch.cat([zero[None], self.zmul(ci, udat), self.zmul(ct_[None], udat)], 0)
            vf = torch.cat([zero[None], self.zmul(ci, vdat), self.zmul(ct_[None], vdat)], 0)
            wf = torch.cat([zero[None], w[:-1], zero[None]], 0)
            a1, a1v = uf[1:], vf[1:]
            ru_z = self.zmul(Z["D"], uf * wf)
            rv_z = self.zmul(Z["D"], vf * wf)
            duz, dvz = self.zmul(Z["D"], uf), self.zmul(Z["D"], vf)
            whz = self.zmul(Z["IFC"], wf)
            wwz = torch.cat([self.zmul(Z["G"], whz * whz), zero[None]], 0)
            dwz = torch.cat([self.zmul(Z["G"], whz), zero[None]], 0)
            wc = whz
            if p.isscalar:
                ti = Z["I_DN"] if p.tbc_top else Z["I_DD"]
                tdat = torch.cat([torch.full_like(zero, p.tbot)[None], t,
                                  torch.full_like(zero, p.dtdz_top if p.tbc_top else p.ttop)[None]], 0)
                thz = torch.cat([self.zmul(ti, tdat), zero[None]], 0)
        else:
            lt, lb = col(self.lt), col(self.lb)
            top_u = -u[-1] if p.ochannel == 0 else u[-1]
            top_v = -v[-1] if p.ochannel == 0 else v[-1]
            up = torch.cat([u[1:], top_u[None]], 0)
            umm = torch.cat([-u[0][None], u[:-1]], 0)
            vp = torch.cat([v[1:], top_v[None]], 0)
            vmm = torch.cat([-v[0][None], v[:-1]], 0)
            wz = torch.cat([w[:-1], zero[None]], 0)
            wp = torch.cat([wz[1:], -wz[-2][None]], 0)
            wm = torch.cat([zero[None], wz[:-1]], 0)
            a1 = up * lt + u * (1 - lt)
            a2 = umm * lb + u * (1 - lb)
            a1v = vp * lt + v * (1 - lt)
            a2v = vmm * lb + v * (1 - lb)
            idz = 1 / col(self.dz)
            ru_z = (a1 * wz - a2 * wm) * idz
            rv_z = (a1v * wz - a2v * wm) * idz
            duz, dvz = (a1 - a2) * idz, (a1v - a2v) * idz
            hp, hm = 0.5 * (wz + wp), 0.5 * (wz + wm)
            wwz = (hp * hp - hm * hm) / col(self.dzt)
            dwz = (hp - hm) / col(self.dzt)
            wc = hm
            w = wz
            if p.isscalar:
                tp = torch.cat([t[1:], (t[-1] + p.dtdz_top * float(self.dzt[-1]) if p.tbc_top
                                        else 2 * p.ttop - t[-1])[None]], 0)
                thz = (1 - lt) * t + lt * tp
        out = [u * u, u * v, v * v, None, None, a1 * w, a1v * w, None]
        tmp2u = p.fc * (v - p.vg) if p.dp_opt == 3 else torch.full_like(u, 1.0)
        tmp2v = p.fc * (p.ug - u) if p.dp_opt == 3 else torch.zeros_like(u)
        ru = tmp2u / coe - ru_z
        rv = tmp2v / coe - rv_z
        rw = -wwz
        if p.isscalar:
            rw = rw + (p.got * (thz - p.t_ref)) / coe
        if p.issk:
            ru = ru - u * D[0] - v * D[1] - duz * wc
            rv = rv - u * D[2] - v * D[3] - dvz * wc
            rw = rw - a1 * D[4] - a1v * D[5] - w * dwz
        if self.z6:  # the top level of the w terms is unused (w = 0 at the wall)
            rw = torch.cat([rw[:-1], zero[None]], 0)
        out[3], out[4], out[7] = ru, rv, rw
        return [self.F(o) for o in out]

    def t_products(self, u, v, w, t):
        p = self.p
        zero = torch.zeros_like(t[0])
        if self.z6:
            Z = self.Z
            ti = Z["I_DN"] if p.tbc_top else Z["I_DD"]
            tdat = torch.cat([torch.full_like(zero, p.tbot)[None], t,
                              torch.full_like(zero, p.dtdz_top if p.tbc_top else p.ttop)[None]], 0)
            ft = torch.cat([zero[None], self.zmul(ti, tdat) * w[:-1], zero[None]], 0)
            twz = self.zmul(Z["D"], ft)
        else:
            col = self.col
            lt, lb = col(self.lt), col(self.lb)
            gdz = p.dtdz_top * float(self.dzt[-1])
            tp = torch.cat([t[1:], (t[-1] + gdz if p.tbc_top else 2 * p.ttop - t[-1])[None]], 0)
            tm = torch.cat([(2 * p.tbot - t[0])[None], t[:-1]], 0)
            wz = torch.cat([w[:-1], zero[None]], 0)
            wm = torch.cat([zero[None], wz[:-1]], 0)
            a1 = tp * lt + t * (1 - lt)
            a2 = tm * lb + t * (1 - lb)
            twz = (a1 * wz - a2 * wm) / col(self.dz)
        return [self.F(u * t), self.F(t * v), self.F(-twz)]

    def substep(self, s, t_its):
        p = self.p
        ra, rb, rc = RK3[t_its - 1]
        dt, nu = p.dt, p.nu
        u, v, w = self.phys(s, "u"), self.phys(s, "v"), self.phys(s, "w")
        t = self.phys(s, "t") if p.isscalar else None
        P = self.products(u, v, w, t)
        coe = self.coe
        dif = rc * 0.5 * dt * nu
        hxy = -dif * self.k2
        cdt = rc * dt
        cp = s["cp"]
        ik = lambda c, k: 1j * k * c
        out = {}
        comps = (("cu", "cff", 0, 1, 3, self.Lu), ("cv", "cgg", 1, 2, 4, self.Lu), ("cw", "chh", 5, 6, 7, self.Lw))
        q = {}
        for e, (cn, fn, ix, iy, ir, Lz) in enumerate(comps):
            c = s[cn]
            fnew = -coe * (ik(P[ix], self.kx) + ik(P[iy], self.ky)) + coe * P[ir]
            qq = c + hxy * c + dif * self.dzz(c, Lz)
            if e == 0:
                qq = qq - cdt * ik(cp, self.kx)
            elif e == 1:
                qq = qq - cdt * ik(cp, self.ky)
            else:
                if self.z6:
                    gz = torch.cat([self.zmul(self.Z["G"], cp), torch.zeros_like(cp[:1])], 0)
                else:
                    gz = torch.cat([(cp[1:] - cp[:-1]) / self.col(self.dzt[:-1]), torch.zeros_like(cp[:1])], 0)
                qq = qq - cdt * gz
            q[cn] = qq + rb * dt * s[fn] + ra * dt * fnew
            out[fn] = fnew
        us = self.helm(q["cu"], self.Lu, dif)
        vs = self.helm(q["cv"], self.Lu, dif)
        ws = self.helm(q["cw"], self.Lw, dif, trivial_top=True)
        if self.z6:
            dwz = self.zmul(self.Z["D"][:, 1:-1], ws[:-1])
        else:
            wm = torch.cat([torch.zeros_like(ws[:1]), ws[:-1]], 0)
            dwz = (ws - wm) / self.col(self.dz)
        div = (ik(us, self.kx) + ik(vs, self.ky) + dwz) / dt
        phi = self.poisson(div / rc)
        lz = self.zmul(self.Lp, phi)
        out["cp"] = (cp + phi - dif * (-self.k2 * phi + lz)) * self.kept
        if self.z6:
            gz = torch.cat([self.zmul(self.Z["G"], phi), torch.zeros_like(phi[:1])], 0)
        else:
            gz = torch.cat([(phi[1:] - phi[:-1]) / self.col(self.dzt[:-1]), torch.zeros_like(phi[:1])], 0)
        out["cu"] = (us - cdt * ik(phi, self.kx)) * self.kept
        out["cv"] = (vs - cdt * ik(phi, self.ky)) * self.kept
        wn = ws - cdt * gz
        out["cw"] = torch.cat([wn[:-1], torch.zeros_like(wn[:1])], 0) * self.kept
        if p.isscalar:
            un, vn, wn = self.Fi(out["cu"]), self.Fi(out["cv"]), self.Fi(out["cw"])
            S = self.t_products(un, vn, wn, t)
            ct = s["ct"]
            dft = rc * 0.5 * dt * p.alpha
            snew = -ik(S[0], self.kx) - ik(S[1], self.ky) + S[2]
            d = ct - dft * self.k2 * ct + dft * self.dzz(ct, self.Lt) + rb * dt * s["css"] + ra * dt * snew
            # wall data of t (the (0,0) mode), explicit and implicit halves
            bt = torch.zeros_like(ct)
            bt[:, 0, 0] = self.bt
            d = d + 2 * dft * bt
            out["ct"] = self.helm(d, self.Lt, dft) * self.kept
            out["css"] = snew
        return out

    def steps(self, st, n):
        s = self.init(st)
        for _ in range(n):
            for t_its in (1, 2, 3):
                s = self.substep(s, t_its)
        res = {k: self.phys(s, k) for k in ("u", "v", "w", "p")}
        if self.p.isscalar:
            res["t"] = self.phys(s, "t")
        return res

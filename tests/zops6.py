"""6th-order wall-normal operators (hc_params.zorder = 6) as dense numpy
matrices — test infrastructure, restating the stencil rule the library
builds its tables by (build_zops in hercules_amd/csrc/hc_api.hip):

* second derivatives from 7 data, first derivatives and interpolation from
  6 data: n consecutive entries of the data set centred on the target,
  shifted inwards where the centred window would leave the data set;
* weights that make the stencil exact for every polynomial of degree < n
  (a Neumann top datum enters as a first-derivative condition).

Data sets (columns): "C" = [bottom wall value, levels 0..nz-1, top wall
value or gradient] (nz + 2), "F" = the faces z_f(0..nz), i.e. w levels
-1..nz-1 with zero wall entries (nz + 1), "P" = levels 0..nz-1 (nz).  Rows:
levels 0..nz-1 for centre targets, interior faces kw = 0..nz-2 (at
z_f(kw + 1)) for face targets.  The reference has no such operators (it is
2nd order in z); these are pinned by polynomial exactness (tests/test_zorder6.py).
"""
from __future__ import annotations

import math

import numpy as np

from hercules_amd import grid


def fd_weights(z0, pts, der, deriv_at=()):
    """Weights w with sum_j w_j data_j = f^(der)(z0) for polynomials of degree
    < len(pts); data_j = f(pts[j]), or f'(pts[j]) for j in deriv_at."""
    n = len(pts)
    pts = np.asarray(pts, dtype=np.float64)
    h = np.abs(pts - z0).max()
    s = (pts - z0) / h
    V = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if j in deriv_at:
                V[i, j] = i * s[j] ** (i - 1) / h if i > 0 else 0.0
            else:
                V[i, j] = s[j] ** i
    rhs = np.zeros(n)
    rhs[der] = math.factorial(der) / h ** der
    return np.linalg.solve(V, rhs)


def window(ndata, center, n):
    lo = int(math.ceil(center - 0.5 * n))
    return max(0, min(lo, ndata - n))


def zops(nz, lz, zstretch, ochannel, order=6):
    """Dense operators of the order-`order` z stencils (order 6: 7- and
    6-point, order 4: 5- and 4-point stencils)."""
    n2, n1 = order + 1, order
    zf = grid.faces(nz, lz, zstretch, ochannel)
    zc = 0.5 * (zf[1:] + zf[:-1])
    cpos = np.concatenate([[zf[0]], zc, [zf[-1]]])
    ops = {}

    def build(name, nrows, ndata, pos, n, der, target, center, ntop=False):
        M = np.zeros((nrows, ndata))
        for r in range(nrows):
            lo = window(ndata, center(r), n)
            idx = list(range(lo, lo + n))
            dat = [j for j, d in enumerate(idx) if ntop and d == ndata - 1]
            M[r, idx] = fd_weights(target(r), pos[idx], der, dat)
        ops[name] = M

    for top, ntop in (("DD", False), ("DN", True)):
        build("C2_" + top, nz, nz + 2, cpos, n2, 2, lambda k: zc[k], lambda k: k + 1.0, ntop)
        build("I_" + top, nz - 1, nz + 2, cpos, n1, 0, lambda k: zf[k + 1], lambda k: k + 1.5, ntop)
    build("F2", nz - 1, nz + 1, zf, n2, 2, lambda k: zf[k + 1], lambda k: k + 1.0)
    build("G", nz - 1, nz, zc, n1, 1, lambda k: zf[k + 1], lambda k: k + 0.5)
    build("D", nz, nz + 1, zf, n1, 1, lambda k: zc[k], lambda k: k + 0.5)
    build("IFC", nz, nz + 1, zf, n1, 0, lambda k: zc[k], lambda k: k + 0.5)
    # interpolation to the top wall face z_f(nz): the wall value itself, or
    # (Neumann top) the extrapolation through the wall gradient
    for top, ntop in (("DD", False), ("DN", True)):
        lo = window(nz + 2, nz + 0.5, n1)
        idx = list(range(lo, lo + n1))
        row = np.zeros(nz + 2)
        row[idx] = fd_weights(zf[-1], cpos[idx], 0, [n1 - 1] if ntop else [])
        ops["I_%s_top" % top] = row
    ops["P"] = ops["D"][:, 1:-1] @ ops["G"]
    ops["zf"], ops["zc"] = zf, zc
    return ops

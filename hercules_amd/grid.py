"""Wall-normal grid of the channel: get_mesh_param (src/module_parameters.f90:432-508).

Returns the interior metrics as numpy arrays indexed by level k = 0..nz-1
(the reference's kstr3..kend3).  The HIP library computes the same tables
on the host side of hc_create.
"""
from __future__ import annotations

import numpy as np


def faces(nz, lz, zstretch, ochannel):
    """z of the nz+1 cell faces (src/module_parameters.f90:445-466)."""
    zmin, zmax = 0.0, lz
    if ochannel == 0:
        zmean, a = 0.5 * (zmin + zmax), (2.0 + nz) * 0.5
    else:
        zmean, a = zmin + zmax, (1.0 + nz) * 1.0
    k = np.arange(1, nz + 2, dtype=np.float64)
    return zmean + (zmin - zmean) / np.tanh(2.0 * zstretch) * np.tanh(2.0 * zstretch * (k - a) / (1.0 - a))


def metrics(nz, lz, zstretch, ochannel):
    """dz, dz_t, dz_b, l_t, l_b and cell-centre z for the interior levels."""
    zf = faces(nz, lz, zstretch, ochannel)
    dz = np.diff(zf)
    dzg = np.concatenate([[dz[0]], dz, [dz[-1]]])  # ghost copies (:473-476)
    dz_t = 0.5 * (dzg[1:-1] + dzg[2:])
    dz_b = 0.5 * (dzg[1:-1] + dzg[:-2])
    l_t = dzg[1:-1] / (dzg[1:-1] + dzg[2:])
    l_b = dzg[1:-1] / (dzg[1:-1] + dzg[:-2])
    zc = 0.5 * (zf[1:] + zf[:-1])
    return dict(z_faces=zf, zc=zc, dz=dz, dz_t=dz_t, dz_b=dz_b, l_t=l_t, l_b=l_b)

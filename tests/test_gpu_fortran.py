"""The Fortran 90 host (hercules_amd/fortran) drives the HIP library through
ISO_C_BINDING: parameters.input and results/init_*.dat in, results/bak_*.dat
and results/mean_*.dat out, compared with the CPU oracle (needs an MI355X)."""
import os
import subprocess

import numpy as np
import pytest

from hercules_amd import Channel, cases
from hercules_amd.build import FEXE
from tests.helpers import make_oracle, rel_l2

pytestmark = pytest.mark.gpu

DECK = """&domain
  p_row = 1, p_col = 1, lx = {lx!r}, ly = {ly!r}, lz = 2.0
  nx = {nx}, ny = {ny}, nz = {nz}, zstretch = {zs!r}, ochannel = 0
/
&modeling
  dp_opt = 1, cds = {cds}, issk = 1, dts = 2, isnoise = 0, noise_mag = 0.1
  isdamp = 0, nzdamp = 15, cdamp = 0.01
/
&constants
  nu = {nu!r}, alpha = {alpha!r}, got = 0.0, fc = 0.0, ug = 0.0, vg = 0.0
  dt = 1.0d-3, imax = {imax}, u_mrf = 0.0, isscalar = 0, t_ref = 0
  tbot = -0.5, ttop = 0.5, is_ri_var = 0, ri_str = 0.0, ri_end = 0.0
/
&io
  restart = 1, istmsr = 0, ibackup = {imax}, iinstfl = 100000, imeanfl = {imax}
  isxy2d = 0, xy2d_id = 10, isxz2d = 0, xz2d_id = 10, isyz2d = 0, yz2d_id = 10, intv_2d = 1000
/
"""


@pytest.mark.parametrize("cds", [2, 3])
def test_fortran_driver_matches_oracle(tmp_path, cds):
    """cds=2: time_advancement_fd; cds=3: time_advancement_sp (the driver's
    step loop dispatches on cds as drive does)."""
    if not os.path.exists(FEXE):
        pytest.fail(f"Fortran host not built: {FEXE}")
    prm = cases.closed_channel(nx=32, ny=32, nz=33, imax=4, cds=cds, issk=1)
    st = cases.laminar_noise(prm, seed=21)
    with Channel(prm) as ch:
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        st = {k: ch.get_field(k) for k in ("u", "v", "w", "p", "t")}
    (tmp_path / "results").mkdir()
    (tmp_path / "parameters.input").write_text(DECK.format(
        lx=prm.lx, ly=prm.ly, nx=prm.nx, ny=prm.ny, nz=prm.nz, zs=prm.zstretch, nu=prm.nu,
        alpha=prm.alpha, imax=prm.imax, cds=cds))
    for k, v in st.items():  # Fortran (i,j,k) order == C order (nz, ny, nx)
        np.ascontiguousarray(v).tofile(tmp_path / "results" / f"init_{k}.dat")
    r = subprocess.run([FEXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    o = make_oracle(prm, st)
    o.mean_enable()
    o.steps(prm.imax)
    # output_mean_fields at step imax: the sums over imax steps / imax
    for k in ("u", "uu", "uw", "vw", "ww"):
        got = np.fromfile(tmp_path / "results" / f"mean_{k}.dat").reshape(prm.nz, prm.ny, prm.nx)
        err = rel_l2(got, o.get_mean(k))
        print("mean", k, f"{err:.2e}")
        assert err < 1e-10, (k, err)
    for k in ("u", "v", "w", "p"):
        got = np.fromfile(tmp_path / "results" / f"bak_{k}.dat").reshape(prm.nz, prm.ny, prm.nx)
        err = rel_l2(got, o.get(k))
        print(f"fortran host {k}: rel L2 {err:.3e}")
        assert err < 1e-10

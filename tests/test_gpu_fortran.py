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
  dp_opt = 1, cds = {cds}, issk = 1, dts = {dts}, isnoise = 0, noise_mag = 0.1
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
        alpha=prm.alpha, imax=prm.imax, cds=cds, dts=2))
    for k, v in st.items():  # Fortran (i,j,k) order == C order (nz, ny, nx)
        np.ascontiguousarray(v).tofile(tmp_path / "results" / f"init_{k}.dat")
    r = subprocess.run([FEXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    # screen_cpu_time once per step; screen_div_error after every substep,
    # with the history line after the last one
    assert r.stdout.count("CFL max =") == prm.imax and r.stdout.count("DIV-ERR FINAL") == 3 * prm.imax
    assert r.stdout.count("DIV-ERR HIST_MAX") == prm.imax
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

    # results/geo.info: z, dz, l_t, l_b of each level, format (4e14.6)
    geo = np.loadtxt(tmp_path / "results" / "geo.info")
    from hercules_amd import grid
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
    for col, ref in enumerate((m["zc"], m["dz"], m["l_t"], m["l_b"])):
        assert np.allclose(geo[:, col], ref, rtol=1e-5, atol=1e-12), col


def _deck(tmp_path, prm, dts):
    (tmp_path / "results").mkdir(exist_ok=True)
    (tmp_path / "parameters.input").write_text(DECK.format(
        lx=prm.lx, ly=prm.ly, nx=prm.nx, ny=prm.ny, nz=prm.nz, zs=prm.zstretch, nu=prm.nu,
        alpha=prm.alpha, imax=prm.imax, cds=prm.cds, dts=dts))


@pytest.mark.parametrize("dts", [2, 1])
def test_fortran_restart_is_bitwise(tmp_path, dts):
    """Resume: 4 steps in one run equal, bit for bit, 2 steps, the backup
    files read back as init files (src/module_tools.f90:260-316), and 2 more
    steps.  With AB2 (dts = 1) the backup carries the right-hand-side history
    ff, gg, hh, ss (src/module_io.f90:340-356)."""
    if not os.path.exists(FEXE):
        pytest.fail(f"Fortran host not built: {FEXE}")
    prm = cases.closed_channel(nx=32, ny=32, nz=33, imax=4, cds=2, issk=1)
    st = cases.laminar_noise(prm, seed=5)
    with Channel(prm) as ch:
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        st = {k: ch.get_field(k) for k in ("u", "v", "w", "p", "t")}
    runs = {}
    for name in ("one", "two"):
        d = tmp_path / name
        d.mkdir()
        _deck(d, prm, dts)
        for k, v in st.items():
            np.ascontiguousarray(v).tofile(d / "results" / f"init_{k}.dat")
        if dts != 2:  # an AB2 restart also reads the right-hand-side history
            for k in ("ff", "gg", "hh", "ss"):
                np.zeros_like(st["u"]).tofile(d / "results" / f"init_{k}.dat")
        runs[name] = d
    r = subprocess.run([FEXE, "4"], cwd=runs["one"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    d = runs["two"]
    r = subprocess.run([FEXE, "2"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    names = ["u", "v", "w", "p", "t"] + (["ff", "gg", "hh", "ss"] if dts != 2 else [])
    for k in names:
        (d / "results" / f"bak_{k}.dat").replace(d / "results" / f"init_{k}.dat")
    r = subprocess.run([FEXE, "2"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for k in names:
        a = np.fromfile(runs["one"] / "results" / f"bak_{k}.dat")
        b = np.fromfile(d / "results" / f"bak_{k}.dat")
        assert np.array_equal(a, b), (k, np.abs(a - b).max())
    assert (d / "results" / "bak_ff.dat").exists() == (dts != 2)


def test_fortran_outputs_series_instants_slices(tmp_path):
    """Time series (istmsr = 1, src/module_io.f90:211-305), instantaneous
    fields every iinstfl steps (:366-403) and x-y / x-z / y-z slices every
    intv_2d steps (:408-542), against the oracle's fields."""
    if not os.path.exists(FEXE):
        pytest.fail(f"Fortran host not built: {FEXE}")
    prm = cases.closed_channel(nx=32, ny=32, nz=33, imax=4, cds=2, issk=1)
    st = cases.laminar_noise(prm, seed=8)
    with Channel(prm) as ch:
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        st = {k: ch.get_field(k) for k in ("u", "v", "w", "p", "t")}
    (tmp_path / "results" / "slices").mkdir(parents=True)
    deck = DECK.format(lx=prm.lx, ly=prm.ly, nx=prm.nx, ny=prm.ny, nz=prm.nz, zs=prm.zstretch, nu=prm.nu,
                       alpha=prm.alpha, imax=prm.imax, cds=2, dts=2)
    deck = deck.replace("restart = 1, istmsr = 0, ibackup = 4, iinstfl = 100000, imeanfl = 4",
                        "restart = 1, istmsr = 1, ibackup = 4, iinstfl = 2, imeanfl = 4")
    deck = deck.replace("isxy2d = 0, xy2d_id = 10, isxz2d = 0, xz2d_id = 10, isyz2d = 0, yz2d_id = 10, intv_2d = 1000",
                        "isxy2d = 1, xy2d_id = 3, 10, isxz2d = 1, xz2d_id = 5, isyz2d = 1, yz2d_id = 7, intv_2d = 2")
    assert "istmsr = 1" in deck and "intv_2d = 2" in deck
    (tmp_path / "parameters.input").write_text(deck)
    for k, v in st.items():
        np.ascontiguousarray(v).tofile(tmp_path / "results" / f"init_{k}.dat")
    r = subprocess.run([FEXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    o = make_oracle(prm, st)
    snaps = []
    for _ in range(prm.imax):
        o.steps(1)
        snaps.append({k: o.get(k) for k in ("u", "v", "w", "p", "t")})
    res = tmp_path / "results"
    # time series: one line per step, time then the (0, 0) column
    ser = np.loadtxt(res / "tmsr_u.dat")
    assert ser.shape == (prm.imax, prm.nz + 1)
    assert np.allclose(ser[:, 0], prm.dt * np.arange(1, prm.imax + 1), atol=1e-8)
    for n in range(prm.imax):
        assert np.allclose(ser[n, 1:], snaps[n]["u"][:, 0, 0], atol=2e-8), n
    # instantaneous fields after steps 2 and 4
    for tag, n in (("001", 1), ("002", 3)):
        got = np.fromfile(res / f"inst_u{tag}.dat").reshape(prm.nz, prm.ny, prm.nx)
        assert rel_l2(got, snaps[n]["u"]) < 1e-10, tag
    # slices after step 4 (output 2)
    u4 = snaps[3]["u"]
    xy = np.fromfile(res / "slices" / "xy2d_u_k0003_t000002").reshape(prm.ny, prm.nx)
    assert rel_l2(xy, u4[2]) < 1e-10
    assert (res / "slices" / "xy2d_w_k0010_t000002").exists()
    xz = np.fromfile(res / "slices" / "xz2d_u_j0005_t000002").reshape(prm.nz, prm.nx)
    assert rel_l2(xz, u4[:, 4, :]) < 1e-10
    yz = np.fromfile(res / "slices" / "yz2d_u_i0007_t000002").reshape(prm.nz, prm.ny)
    assert rel_l2(yz, u4[:, :, 6]) < 1e-10


DECK_T = """&domain
  p_row = 1, p_col = 1, lx = {lx!r}, ly = {ly!r}, lz = {lz!r}
  nx = {nx}, ny = {ny}, nz = {nz}, zstretch = {zs!r}, ochannel = {och}
/
&modeling
  dp_opt = {dp}, cds = {cds}, issk = 1, dts = {dts}, isnoise = 0, noise_mag = 0.1
  isdamp = 0, nzdamp = 15, cdamp = 0.01
/
&constants
  nu = {nu!r}, alpha = {alpha!r}, got = {got!r}, fc = 0.0, ug = 0.0, vg = 0.0
  dt = 1.0d-3, imax = {imax}, u_mrf = {mrf!r}, isscalar = {iss}, t_ref = 0
  tbot = 0.0, ttop = 1.0, is_ri_var = {ri}, ri_str = 0.05, ri_end = 0.15
/
&io
  restart = 1, istmsr = 0, ibackup = {imax}, iinstfl = 100000, imeanfl = 100000
  isxy2d = 0, xy2d_id = 10, isxz2d = 0, xz2d_id = 10, isyz2d = 0, yz2d_id = 10, intv_2d = 1000
/
"""


@pytest.mark.parametrize("kind", ["closed_cds2", "open_t_dp2_ri", "closed_cds1_ab2_mrf"])
def test_host_stage_loop_equals_device_step(tmp_path, kind):
    """A Fortran caller running the reference's substep loop on the host with
    the per-stage subroutines (tests/fortran/stage_loop.f90) gives the fields
    of the fused device step run by the driver on the same deck."""
    from hercules_amd.build import STAGE_EXE
    if not os.path.exists(FEXE) or not os.path.exists(STAGE_EXE):
        pytest.fail("Fortran host programs not built")
    if kind == "closed_cds2":
        prm = cases.closed_channel(nx=32, ny=32, nz=33, imax=3)
        st = cases.laminar_noise(prm, seed=31)
        dts, ri = 2, 0
    elif kind == "open_t_dp2_ri":
        prm = cases.stratified_open_channel(nx=32, ny=32, nz=33, imax=3, dp_opt=2, is_ri_var=1,
                                            ri_str=0.05, ri_end=0.15)
        st = cases.loglaw_noise(prm, seed=32)
        dts, ri = 2, 1
    else:
        prm = cases.closed_channel(nx=16, ny=32, nz=17, imax=3, cds=1, dts=1, u_mrf=0.25)
        st = cases.laminar_noise(prm, seed=33)
        dts, ri = 1, 0
    with Channel(prm) as ch:
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        st = {k: ch.get_field(k) for k in ("u", "v", "w", "p", "t")}
    (tmp_path / "results").mkdir()
    (tmp_path / "parameters.input").write_text(DECK_T.format(
        lx=prm.lx, ly=prm.ly, lz=prm.lz, nx=prm.nx, ny=prm.ny, nz=prm.nz, zs=prm.zstretch, och=prm.ochannel,
        dp=prm.dp_opt, cds=prm.cds, dts=dts, nu=prm.nu, alpha=prm.alpha, got=prm.got, imax=prm.imax,
        mrf=prm.u_mrf, iss=prm.isscalar, ri=ri))
    for k, v in st.items():  # init files hold u + u_mrf
        (np.ascontiguousarray(v) + (prm.u_mrf if k == "u" else 0.0)).tofile(tmp_path / "results" / f"init_{k}.dat")
    if dts == 1:  # AB2 restarts read the right-hand-side history too
        for k in ("ff", "gg", "hh", "ss"):
            np.zeros((prm.nz, prm.ny, prm.nx)).tofile(tmp_path / "results" / f"init_{k}.dat")
    nsub = 3 if dts == 2 else 1
    for exe in (FEXE, STAGE_EXE):
        r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        print(exe, r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
        assert r.stdout.count("DIV-ERR FINAL") == nsub * prm.imax
        assert r.stdout.count("HIST_MAX") == prm.imax
        if ri:
            assert r.stdout.count("Ri_b =") == prm.imax
    names = ("u", "v", "w", "p", "t") if prm.isscalar else ("u", "v", "w", "p")
    for k in names:
        a = np.fromfile(tmp_path / "results" / f"bak_{k}.dat")
        b = np.fromfile(tmp_path / "results" / f"stage_{k}.dat")
        err = rel_l2(b, a)
        print(kind, k, f"{err:.2e}")
        assert err < 1e-11, (k, err)


@pytest.mark.parametrize("kind", ["closed_sp_skew", "open_t_sp_dp2_damped", "closed_sp_adv_ab2"])
def test_host_spectral_stage_loop_equals_device_step(tmp_path, kind):
    """A Fortran caller running the reference's spectral substep loop and
    start-up on the host with the spectral per-stage subroutines and complex
    work arrays (tests/fortran/stage_loop_sp.f90) gives the fields of the
    fused device step time_advancement_sp run by the driver on the same deck."""
    from hercules_amd.build import STAGE_SP_EXE
    if not os.path.exists(FEXE) or not os.path.exists(STAGE_SP_EXE):
        pytest.fail("Fortran host programs not built")
    issk, damp = 1, "isdamp = 0, nzdamp = 15, cdamp = 0.01"
    if kind == "closed_sp_skew":
        prm = cases.closed_channel(nx=32, ny=32, nz=33, imax=3, cds=3)
        st = cases.laminar_noise(prm, seed=34)
        dts = 2
    elif kind == "open_t_sp_dp2_damped":
        prm = cases.stratified_open_channel(nx=32, ny=32, nz=33, imax=3, cds=3, dp_opt=2, isdamp=1, nzdamp=6,
                                            cdamp=0.2)
        st = cases.loglaw_noise(prm, seed=35)
        dts, damp = 2, "isdamp = 1, nzdamp = 6, cdamp = 0.2"
    else:
        prm = cases.closed_channel(nx=16, ny=32, nz=17, imax=3, cds=3, dts=1)
        st = cases.laminar_noise(prm, seed=36)
        dts, issk = 1, 0
    with Channel(prm) as ch:
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        st = {k: ch.get_field(k) for k in ("u", "v", "w", "p", "t")}
    (tmp_path / "results").mkdir()
    deck = DECK_T.replace("issk = 1", f"issk = {issk}").replace("isdamp = 0, nzdamp = 15, cdamp = 0.01", damp)
    (tmp_path / "parameters.input").write_text(deck.format(
        lx=prm.lx, ly=prm.ly, lz=prm.lz, nx=prm.nx, ny=prm.ny, nz=prm.nz, zs=prm.zstretch, och=prm.ochannel,
        dp=prm.dp_opt, cds=prm.cds, dts=dts, nu=prm.nu, alpha=prm.alpha, got=prm.got, imax=prm.imax,
        mrf=prm.u_mrf, iss=prm.isscalar, ri=0))
    for k, v in st.items():
        (np.ascontiguousarray(v) + (prm.u_mrf if k == "u" else 0.0)).tofile(tmp_path / "results" / f"init_{k}.dat")
    if dts == 1:
        for k in ("ff", "gg", "hh", "ss"):
            np.zeros((prm.nz, prm.ny, prm.nx)).tofile(tmp_path / "results" / f"init_{k}.dat")
    for exe in (FEXE, STAGE_SP_EXE):
        r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        print(exe, r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
    names = ("u", "v", "w", "p", "t") if prm.isscalar else ("u", "v", "w", "p")
    for k in names:
        a = np.fromfile(tmp_path / "results" / f"bak_{k}.dat")
        b = np.fromfile(tmp_path / "results" / f"stage_{k}.dat")
        err = rel_l2(b, a)
        print(kind, k, f"{err:.2e}")
        assert err < 1e-11, (k, err)

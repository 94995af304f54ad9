"""Per-stage drop-ins of the spectral substep (get_momentum_cnvdiff_spec,
get_temperature_cnvdiff_spec, poisson_solver_fft_spec, get_div_spec,
calculate_uvw_spec, calculate_new_p_spec) through the C ABI, against the
oracle's own stages on the same state.  The reference's complex work arrays
cross the ABI whole; every mode but the x and y Nyquist modes (which the
library does not carry) is compared.  The damped case feeds spectra that
rayleigh_damping_spec changed while the physical fields kept their values,
the mixture the reference's stages see after :406-414."""
import numpy as np
import pytest

from hercules_amd import Channel, cases
from hercules_amd.channel import LAYOUT_GHOSTED, LAYOUT_INTERIOR, HerculesError
from tests.helpers import make_oracle

pytestmark = pytest.mark.gpu

RK = [(8 / 15, 0.0, 8 / 15), (5 / 12, -17 / 60, 2 / 15), (3 / 4, -5 / 12, 1 / 3)]

CASES = {
    "closed_skew": lambda: (cases.closed_channel(nx=32, ny=32, nz=33, cds=3, issk=1), "laminar"),
    "closed_adv_dp2_mrf": lambda: (cases.closed_channel(nx=16, ny=32, nz=33, cds=3, issk=0, dp_opt=2, u_mrf=0.3),
                                   "laminar"),
    "open_t_skew": lambda: (cases.stratified_open_channel(nx=32, ny=32, nz=33, cds=3, issk=1), "loglaw"),
    "open_t_coriolis_damped": lambda: (cases.stratified_open_channel(nx=16, ny=32, nz=33, cds=3, issk=1, dp_opt=3,
                                                                     fc=0.5, ug=1.0, vg=0.2, isdamp=1, nzdamp=6,
                                                                     cdamp=0.2), "loglaw"),
}


def body(a):
    """every mode but the x Nyquist column and the y Nyquist row"""
    nz, ny, nx = a.shape
    keep = np.ones(a.shape, dtype=bool)
    keep[:, :, nx // 2] = False
    keep[:, ny // 2, :] = False
    return a[keep]


def crel(a, b):
    return float(np.linalg.norm(body(a) - body(b)) / max(np.linalg.norm(body(b)), 1e-300))


def ghost(a):
    g = np.full((a.shape[0] + 4, a.shape[1] + 4, a.shape[2] + 4), 7.25)
    g[2:-2, 2:-2, 2:-2] = a
    return g


def setup(name, t_its):
    """oracle and library one step into the same run; the oracle's state
    (fields, spectra, right-hand-side history) is the stages' input"""
    prm, kind = CASES[name]()
    st = (cases.laminar_noise if kind == "laminar" else cases.loglaw_noise)(prm, seed=41)
    if not prm.isscalar:
        st.pop("t", None)
    o = make_oracle(prm, st)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    o.steps(1)
    ch.steps(1)  # the same dpx_drive and ubar on the device
    from oracle.oracle import lib
    if prm.isdamp == 1:
        for fid in range(5):
            lib().ho_stage_damping_spec(o.h, fid)
    phys = {k: o.get(k) for k in ("u", "v", "w", "p", "t")}
    spec = {k: o.spectrum(k) for k in ("u", "v", "w", "p", "t", "ff", "gg", "hh", "ss")}
    return prm, o, ch, phys, spec, lib()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("t_its", [1, 2])
def test_spectral_rhs_stages(name, t_its):
    prm, o, ch, phys, spec, L = setup(name, t_its)
    rk = RK[t_its - 1] if prm.dts == 2 else (1.5, -0.5, 1.0)
    layout = LAYOUT_GHOSTED if t_its == 2 else LAYOUT_INTERIOR
    wrap = ghost if layout == LAYOUT_GHOSTED else (lambda a: a)
    t = wrap(phys["t"]) if prm.isscalar else None
    got = ch.get_momentum_cnvdiff_spec(rk, *(wrap(phys[k]) for k in ("u", "v", "w", "p")), t,
                                       spec["u"], spec["v"], spec["w"], spec["p"],
                                       spec["ff"], spec["gg"], spec["hh"], layout)
    assert L.ho_stage_momentum_spec(o.h, t_its) == 0
    ref = [o.spectrum(k) for k in ("ff", "gg", "hh", "qu", "qv", "qw")]
    for nm, a, b in zip(("c_ff", "c_gg", "c_hh", "c_qu", "c_qv", "c_qw"), got, ref):
        if nm in ("c_hh", "c_qw"):  # the top face of w is a wall: no equation there
            a, b = a[:-1], b[:-1]
        err = crel(a, b)
        print(name, t_its, nm, f"{err:.2e}")
        assert err < 1e-11, (nm, err)
    if prm.isscalar:
        ss, qt = ch.get_temperature_cnvdiff_spec(rk, *(wrap(phys[k]) for k in ("u", "v", "w", "t")),
                                                 spec["t"], spec["ss"], layout)
        assert L.ho_stage_temperature_spec(o.h, t_its) == 0
        for nm, a, b in (("c_ss", ss, o.spectrum("ss")), ("c_qt", qt, o.spectrum("qu"))):
            err = crel(a, b)
            print(name, t_its, nm, f"{err:.2e}")
            assert err < 1e-11, (nm, err)
    ch.close()


@pytest.mark.parametrize("name", ["closed_skew", "open_t_skew"])
def test_spectral_solve_div_project_new_p_stages(name):
    prm, o, ch, phys, spec, L = setup(name, 2)
    rk_c = RK[1][2]
    # poisson_solver_fft_spec of every equation on the same source
    src = np.fft.fft2(np.random.default_rng(5).standard_normal((prm.nz, prm.ny, prm.nx))) / (prm.nx * prm.ny)
    for eq in ("u", "v", "w", "p", "t"):
        if eq == "t" and not prm.isscalar:
            continue
        o.set_spectrum("qu", src)
        assert L.ho_stage_poisson_spec(o.h, {"u": 0, "v": 1, "w": 2, "p": 3, "t": 4}[eq], rk_c) == 0
        ref = o.spectrum("phi")
        got = ch.poisson_solver_fft_spec(eq, rk_c, src)
        if eq == "w":
            ref, got = ref[:-1], got[:-1]
        err = crel(got, ref)
        print(name, "poisson", eq, f"{err:.2e}")
        assert err < (1e-9 if eq == "p" else 1e-12), (eq, err)
    # calculate_uvw_spec and calculate_new_p_spec with a seeded phi and its spectrum
    phi = np.random.default_rng(6).standard_normal((prm.nz, prm.ny, prm.nx))
    c_phi = np.fft.fft2(phi) / (prm.nx * prm.ny)
    o.set("phi", phi)
    o.set_spectrum("phi", c_phi)
    got_uvw = ch.calculate_uvw_spec(rk_c, ghost(phi), c_phi, spec["u"], spec["v"], spec["w"], LAYOUT_GHOSTED)
    got_p = ch.calculate_new_p_spec(rk_c, phi, c_phi, spec["p"])
    assert L.ho_stage_uvw_spec(o.h, 2) == 0
    assert L.ho_stage_new_p_spec(o.h, 2) == 0
    for nm, a, b in (("c_u", got_uvw[0], o.spectrum("u")), ("c_v", got_uvw[1], o.spectrum("v")),
                     ("c_w", got_uvw[2][:-1], o.spectrum("w")[:-1]), ("c_p", got_p, o.spectrum("p"))):
        err = crel(a, b)
        print(name, nm, f"{err:.2e}")
        assert err < 1e-11, (nm, err)
    # get_div_spec(c_u, c_v, w) of seeded fields (the stepped state is
    # divergence-free to round-off, which would compare round-off)
    rng = np.random.default_rng(7)
    sh = (prm.nz, prm.ny, prm.nx)
    cu, cv = (np.fft.fft2(rng.standard_normal(sh)) / (prm.nx * prm.ny) for _ in range(2))
    w = rng.standard_normal(sh)
    w[-1] = 0.0
    o.set_spectrum("u", cu)
    o.set_spectrum("v", cv)
    o.set("w", w)
    got = ch.get_div_spec(cu, cv, w)
    assert L.ho_stage_div_spec(o.h) == 0
    err = crel(got, o.spectrum("qu"))
    print(name, "div", f"{err:.2e}")
    assert err < 1e-11, err
    ch.close()


def test_spectral_stages_need_the_spectral_fp64_path():
    prm = cases.closed_channel(nx=16, ny=32, nz=17)
    ch = Channel(prm)
    z = np.zeros((prm.nz, prm.ny, prm.nx), dtype=np.complex128)
    with pytest.raises(HerculesError, match="status 1"):
        ch.poisson_solver_fft_spec("u", 1.0, z)
    ch.close()
    ch = Channel(prm.replace(cds=3, precision=1))
    with pytest.raises(HerculesError, match="status 2"):
        ch.poisson_solver_fft_spec("u", 1.0, z)
    ch.close()

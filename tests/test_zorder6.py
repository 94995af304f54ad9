"""High-order z operators (hc_params.zorder = 6 and 4) on the CPU: the dense
restatement in tests/zops6.py pinned by what the mathematics fixes (the
reference has no such operators; it is 2nd order in z).

* every stencil is exact for polynomials of degree < its width (order + 1
  for second derivatives, order for first derivatives and interpolation),
  with the wall data entering as values or, for a Neumann top, as a gradient;
* the stencils are centred in the interior and one-sided at the walls;
* the pressure operator D G annihilates constants and its range is the
  range of D (so every divergence has a solution);
* on smooth fields the errors fall at the operator's order with nz;
* the torch restatement of the step with these operators (tests/
  torch_reference.py) leaves a divergence-free velocity after projection.
"""
import numpy as np
import pytest

from hercules_amd import cases
from tests.zops6 import zops

GRIDS = [(33, 2.0, 1.0, 0), (65, 2.0, 1.0, 0), (33, 1.0, 1.0, 1), (129, 1.0, 1.4, 1)]


def _poly(z, deg):
    return z ** deg


def _dpoly(z, deg, der):
    c = 1.0
    for q in range(der):
        c *= deg - q
    return c * z ** (deg - der) if deg >= der else 0.0 * z


@pytest.mark.parametrize("order", [6, 4])
@pytest.mark.parametrize("grid", GRIDS)
def test_stencils_exact_for_polynomials(grid, order):
    nz, lz, zs, och = grid
    Z = zops(nz, lz, zs, och, order)
    zf, zc = Z["zf"], Z["zc"]
    # shift so the monomials are not degenerate about z = 0
    s = lambda z: (z - 0.37 * lz) / lz
    for deg in range(order + 1):
        f = lambda z: _poly(s(z), deg)
        d1 = lambda z: _dpoly(s(z), deg, 1) / lz
        d2 = lambda z: _dpoly(s(z), deg, 2) / lz ** 2
        cdat = np.concatenate([[f(zf[0])], f(zc), [f(zf[-1])]])
        cdat_n = np.concatenate([[f(zf[0])], f(zc), [d1(zf[-1])]])
        fdat = f(zf)

        def check(M, data, exact):
            # round-off of the stencil sum: |weights| . |data| times a few eps
            tol = 1e-13 * (np.abs(M) @ np.abs(data)) + 1e-13
            assert np.all(np.abs(M @ data - exact) <= tol), (deg, np.abs(M @ data - exact).max())

        check(Z["C2_DD"], cdat, d2(zc))
        check(Z["C2_DN"], cdat_n, d2(zc))
        check(Z["F2"], fdat, d2(zf[1:-1]))
        if deg < order:
            check(Z["G"], f(zc), d1(zf[1:-1]))
            check(Z["D"], fdat, d1(zc))
            check(Z["I_DD"], cdat, f(zf[1:-1]))
            check(Z["I_DN"], cdat_n, f(zf[1:-1]))
            check(Z["IFC"], fdat, f(zc))
            check(Z["I_DD_top"][None], cdat, f(zf[-1:]))
            check(Z["I_DN_top"][None], cdat_n, f(zf[-1:]))
    # a polynomial of degree order + 1 is not reproduced: the stencils are not wider
    f7 = lambda z: _poly(s(z), order + 1)
    cdat = np.concatenate([[f7(zf[0])], f7(zc), [f7(zf[-1])]])
    err7 = np.abs(Z["C2_DD"] @ cdat - _dpoly(s(zc), order + 1, 2) / lz ** 2)
    tol = 1e-13 * (np.abs(Z["C2_DD"]) @ np.abs(cdat)) + 1e-13
    assert (err7 > 100 * tol).any()


def test_stencils_centred_inside_one_sided_at_walls():
    nz = 33
    Z = zops(nz, 2.0, 1.0, 0)
    support = lambda row: np.nonzero(row)[0]
    # centre k: data k-2..k+4 of [wall, levels, wall] (level k at index k+1)
    for k in (3, 10, nz - 4):
        assert list(support(Z["C2_DD"][k])) == list(range(k - 2, k + 5))
    assert list(support(Z["C2_DD"][0])) == list(range(0, 7))          # bottom wall value + 6 levels
    assert list(support(Z["C2_DD"][nz - 1])) == list(range(nz - 5, nz + 2))
    # face kw: levels kw-2..kw+3
    for kw in (2, 12, nz - 4):
        assert list(support(Z["G"][kw])) == list(range(kw - 2, kw + 4))
    assert list(support(Z["G"][0])) == list(range(0, 6))
    # centre k: faces k-2..k+3
    assert list(support(Z["D"][10])) == list(range(8, 14))
    # D G: at most 16 diagonals (the library's table width)
    for k in range(nz):
        sup = support(np.abs(Z["P"][k]) > 1e-300)
        assert sup[-1] - sup[0] + 1 <= 16


@pytest.mark.parametrize("order", [6, 4])
@pytest.mark.parametrize("grid", GRIDS)
def test_pressure_operator_nullspace_and_range(grid, order):
    nz, lz, zs, och = grid
    Z = zops(nz, lz, zs, och, order)
    P, D = Z["P"], Z["D"][:, 1:-1]
    assert np.abs(P @ np.ones(nz)).max() < 1e-8 * np.abs(P).max()
    # rank nz - 1, and every D w (w on the interior faces) is in the range of P
    assert np.linalg.matrix_rank(P) == nz - 1
    w = np.random.default_rng(1).standard_normal(nz - 1)
    rhs = D @ w
    phi, res, *_ = np.linalg.lstsq(P, rhs, rcond=None)
    assert np.linalg.norm(P @ phi - rhs) < 1e-10 * np.linalg.norm(rhs)


@pytest.mark.parametrize("order", [6, 4])
def test_convergence_order_on_smooth_fields(order):
    """Errors of d2/dz2, d/dz (centre and face) and interpolation of
    cos(pi z) on the stretched closed-channel grid, nz = 33 .. 257: the
    observed order is at least order - 0.2 (on the tanh grid the one-sided
    wall rows keep the interior order: 6.0 and 4.0 are observed)."""
    errs = {k: [] for k in ("C2_DD", "D", "G", "I_DD")}
    hs = []
    for nz in (33, 65, 129, 257):
        lz = 2.0
        Z = zops(nz, lz, 1.0, 0, order)
        zf, zc = Z["zf"], Z["zc"]
        f = lambda z: np.cos(np.pi * z)
        cd = np.concatenate([[f(zf[0])], f(zc), [f(zf[-1])]])
        errs["C2_DD"].append(np.abs(Z["C2_DD"] @ cd + np.pi ** 2 * f(zc)).max())
        errs["D"].append(np.abs(Z["D"] @ f(zf) + np.pi * np.sin(np.pi * zc)).max())
        errs["G"].append(np.abs(Z["G"] @ f(zc) + np.pi * np.sin(np.pi * zf[1:-1])).max())
        errs["I_DD"].append(np.abs(Z["I_DD"] @ cd - f(zf[1:-1])).max())
        hs.append(1.0 / nz)
    for k, e in errs.items():
        got = np.log(e[-2] / e[-1]) / np.log(hs[-2] / hs[-1])
        print(k, ["%.2e" % x for x in e], f"order {got:.2f}")
        assert got > order - 0.2, (k, e)


@pytest.mark.parametrize("order", [6, 4])
def test_torch_z6_projection_is_divergence_free(order):
    torch = pytest.importorskip("torch")
    from tests.torch_reference import TorchChannel
    prm = cases.closed_channel(nx=16, ny=16, nz=33, wall_rows=1, zorder=order)
    st = cases.laminar_noise(prm, seed=3)
    tc = TorchChannel(prm)
    out = tc.steps({k: torch.tensor(v) for k, v in st.items()}, 1)
    Z = tc.Z
    u, v, w = out["u"], out["v"], out["w"]
    div = (tc.d_p(u, 2, tc.dx) + tc.d_p(v, 1, tc.dy) + tc.zmul(Z["D"][:, 1:-1], w[:-1])).numpy()
    scale = np.abs(u.numpy()).max() / float(tc.dz.min())
    print("max |div|", np.abs(div).max(), "scale", scale)
    assert np.abs(div).max() < 1e-12 * scale

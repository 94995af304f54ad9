"""The ensemble and adjoint configurations at their stated sizes.

Ensemble: 24 channels, eight each at 128x128x129, 192x192x129 and
256x128x193 (neutral, stratified and Ekman-type members with seeded Re_tau
and Ri), advanced together from one captured HIP graph; every member equals,
bit for bit, the same member stepped alone, in FP64 and in the mixed variant
(members with the top sponge stay FP64).

Adjoint: 128x128x129 neutral channel at Re_tau = 180, the plane-averaged
bottom wall shear after K = 400 RK3 steps, checkpoint every 20 steps; its
gradient with respect to the initial velocity against central differences
(eps = 1e-6 relative) along 8 seeded divergence-free directions."""
import numpy as np
import pytest

from hercules_amd import Channel, Ensemble, cases

pytestmark = pytest.mark.gpu


def _channel(prm, st):
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    return ch


@pytest.mark.parametrize("mixed", [False, True])
def test_ensemble_of_24_members_equals_single_runs(mixed):
    members_spec = cases.ensemble_24(seed=2024)
    nsteps = 2
    out_single = []
    for prm, st in members_spec:
        pr = 1 if (mixed and prm.isdamp == 0) else 0
        ch = _channel(prm.replace(precision=pr), st)
        ch.steps(nsteps)
        names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
        out_single.append({n: ch.get_field(n) for n in names})
        ch.close()
    members = []
    for prm, st in members_spec:
        pr = 1 if (mixed and prm.isdamp == 0) else 0
        members.append(_channel(prm.replace(precision=pr), st))
    with Ensemble(members, use_graph=True) as ens:
        ens.steps(nsteps)
        assert ens.graph_used
    cells = 0
    for q, ((prm, _), ref, ch) in enumerate(zip(members_spec, out_single, members)):
        cells += prm.nx * prm.ny * prm.nz
        for n, x in ref.items():
            y = ch.get_field(n)
            assert np.isfinite(x).all(), (q, n)
            assert np.array_equal(x, y), (q, prm.nx, prm.ny, prm.nz, n, np.abs(x - y).max())
        ch.close()
    assert cells == 105_545_728


def test_adjoint_gradient_at_400_steps_against_8_central_differences():
    prm = cases.closed_channel(nx=128, ny=128, nz=129, wall_rows=1)
    st = cases.loglaw_noise(prm, seed=2024, rel_amp=0.05)
    st.pop("t")
    K, ck = 400, 20
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.project()
    x0 = {k: ch.get_field(k) for k in ("u", "v", "w", "p")}
    ch.init()
    loss = ch.wall_shear_gradient(K, ck)
    g = {k: ch.get_adjoint(k) for k in ("u", "v", "w")}
    ch.steps(K)  # the forward run gives the same loss
    assert ch.wall_shear()[0] == pytest.approx(loss, rel=1e-14)
    norm0 = np.sqrt(sum(np.vdot(x0[k], x0[k]) for k in g))
    rng = np.random.default_rng(123)
    for q in range(8):
        for k in ("u", "v", "w"):
            a = rng.standard_normal(x0[k].shape)
            if k == "w":
                a[-1] = 0.0
            ch.set_field(k, a)
        ch.project()  # a divergence-free direction
        d = {k: ch.get_field(k) for k in ("u", "v", "w")}
        eps = 1e-6 * norm0 / np.sqrt(sum(np.vdot(d[k], d[k]) for k in d))
        js = []
        for sgn in (1.0, -1.0):
            for k, v in x0.items():
                ch.set_field(k, v + sgn * eps * d[k] if k in d else v)
            ch.init()
            ch.steps(K)
            js.append(ch.wall_shear()[0])
        fd = (js[0] - js[1]) / (2 * eps)
        ad = sum(float(np.vdot(g[k], d[k])) for k in d)
        err = abs(fd - ad) / abs(ad)
        print(f"direction {q}: fd {fd:.10e} adjoint {ad:.10e} rel {err:.2e}")
        assert err <= 1e-6, (q, fd, ad, err)
    ch.close()

"""Discrete adjoint of the finite-difference step (needs an MI355X).

The substep map (u, v, w, p, f_old) -> (u', v', w', p', f) is quadratic in
the velocities and linear in the rest, so a central difference of the
forward substep is exact up to rounding: <S(x + e d) - S(x - e d), l> / 2e
must equal <d, S'(x)^T l> from hc_adjoint_substep_fd for random d, l.  The
gradient of the wall-shear loss over several RK3 steps (checkpointed,
recomputed) is held against central differences along seeded
divergence-free directions, and checkpoint spacing must not change it."""
import numpy as np
import pytest

# torch before the library: the process keeps the first HIP runtime it loads
torch = pytest.importorskip("torch")

from hercules_amd import Channel, cases  # noqa: E402
from hercules_amd.channel import HerculesError  # noqa: E402

pytestmark = pytest.mark.gpu

OUT = ("u", "v", "w", "p", "ff", "gg", "hh")


def _params(kind, **kw):
    if kind == "closed":
        return cases.closed_channel(**kw)
    base = dict(isscalar=0, got=0.0, nu=1.0 / 180.0)
    base.update(kw)
    return cases.stratified_open_channel(**base)


def _state(prm, seed):
    st = cases.laminar_noise(prm, seed=seed) if prm.ochannel == 0 else cases.loglaw_noise(prm, seed=seed)
    st.pop("t", None)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.project()
    ch.init()
    ch.steps(1)  # a pressure field and stored right-hand sides
    out = {k: ch.get_field(k) for k in OUT}
    ch.close()
    return out


def _set(ch, st):
    for k, v in st.items():
        ch.set_field(k, v)


CASES = [
    ("closed", dict(nx=32, ny=32, nz=33)),
    ("closed", dict(nx=32, ny=32, nz=33, cds=1)),
    ("closed", dict(nx=16, ny=32, nz=65, u_mrf=0.4, wall_rows=1)),
    ("closed", dict(nx=32, ny=32, nz=33, dp_opt=3, fc=0.5, ug=1.0, vg=0.2)),
    ("open", dict(nx=32, ny=32, nz=33)),
    ("closed", dict(nx=32, ny=32, nz=33, dts=1)),
]


@pytest.mark.parametrize("kind,kw", CASES)
def test_substep_adjoint_is_the_transposed_jacobian(kind, kw):
    prm = _params(kind, **kw)
    x = _state(prm, seed=3)
    rng = np.random.default_rng(11)
    shape = (prm.nz, prm.ny, prm.nx)
    nsub = 3 if prm.dts == 2 else 1
    ch = Channel(prm)
    for t_its in range(1, nsub + 1):
        d = {k: rng.standard_normal(shape) for k in OUT}
        lam = {k: rng.standard_normal(shape) for k in OUT}
        d["w"][-1] = 0.0
        lam["w"][-1] = 0.0
        eps = 1e-4
        outs = []
        for sgn in (1.0, -1.0):
            _set(ch, {k: x[k] + sgn * eps * d[k] for k in OUT})
            ch.substep(t_its)
            outs.append({k: ch.get_field(k) for k in OUT})
        fd = sum(float(np.vdot(outs[0][k] - outs[1][k], lam[k])) for k in OUT) / (2 * eps)
        _set(ch, x)
        for k in OUT:
            ch.set_adjoint(k, lam[k])
        ch.adjoint_substep(t_its)
        ad = sum(float(np.vdot(d[k], ch.get_adjoint(k))) for k in OUT)
        print(f"{kind} {kw} t_its={t_its}: fd {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= 1e-9 * abs(ad)
    ch.close()


def _divfree_direction(prm, seed):
    rng = np.random.default_rng(seed)
    shape = (prm.nz, prm.ny, prm.nx)
    ch = Channel(prm)
    for k in ("u", "v", "w"):
        a = rng.standard_normal(shape)
        if k == "w":
            a[-1] = 0.0
        ch.set_field(k, a)
    ch.project()
    d = {k: ch.get_field(k) for k in ("u", "v", "w")}
    ch.close()
    return d


def _loss(prm, st, nsteps):
    ch = Channel(prm)
    _set(ch, st)
    ch.init()
    ch.steps(nsteps)
    tau = ch.wall_shear()[0]
    ch.close()
    return tau


@pytest.mark.parametrize("kind,kw", [("closed", dict(nx=32, ny=32, nz=33)),
                                     ("closed", dict(nx=32, ny=32, nz=65, cds=1, wall_rows=1)),
                                     ("open", dict(nx=32, ny=32, nz=33, dp_opt=3, fc=0.5, ug=1.0))])
def test_wall_shear_gradient_matches_central_differences(kind, kw):
    prm = _params(kind, **kw)
    x = _state(prm, seed=5)
    x = {k: x[k] for k in ("u", "v", "w", "p")}
    nsteps = 6
    ch = Channel(prm)
    _set(ch, x)
    ch.init()
    loss = ch.wall_shear_gradient(nsteps, ckpt_every=4)
    g = {k: ch.get_adjoint(k) for k in ("u", "v", "w", "p")}
    back = {k: ch.get_field(k) for k in ("u", "v", "w", "p")}
    ch.close()
    for k in x:  # the initial state is restored
        assert np.array_equal(back[k], x[k]), k
    assert abs(loss - _loss(prm, x, nsteps)) <= 1e-12 * abs(loss)
    for seed in range(3):
        d = _divfree_direction(prm, 100 + seed)
        d["p"] = np.random.default_rng(200 + seed).standard_normal(x["p"].shape)
        scale = np.sqrt(sum(np.vdot(x[k], x[k]) for k in ("u", "v", "w")) /
                        sum(np.vdot(d[k], d[k]) for k in ("u", "v", "w")))
        eps = 1e-6 * scale
        jp = _loss(prm, {k: x[k] + eps * d[k] for k in x}, nsteps)
        jm = _loss(prm, {k: x[k] - eps * d[k] for k in x}, nsteps)
        fd = (jp - jm) / (2 * eps)
        ad = sum(float(np.vdot(g[k], d[k])) for k in x)
        print(f"{kind} {kw} direction {seed}: fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_checkpoint_spacing_does_not_change_the_gradient():
    prm = _params("closed", nx=32, ny=32, nz=33)
    x = _state(prm, seed=8)
    x = {k: x[k] for k in ("u", "v", "w", "p")}
    grads = []
    for ck in (1, 3, 7):
        ch = Channel(prm)
        _set(ch, x)
        ch.init()
        ch.wall_shear_gradient(7, ckpt_every=ck)
        grads.append({k: ch.get_adjoint(k) for k in x})
        ch.close()
    for gq in grads[1:]:
        for k in x:
            assert np.array_equal(gq[k], grads[0][k]), k


def test_torch_autograd_binding_returns_the_adjoint_gradient():
    torch = pytest.importorskip("torch")
    from hercules_amd.autograd import wall_shear_loss
    prm = _params("closed", nx=32, ny=32, nz=33)
    x = _state(prm, seed=9)
    ch = Channel(prm)
    _set(ch, {k: x[k] for k in ("u", "v", "w", "p")})
    ch.init()
    u0, v0, w0 = (torch.tensor(x[k], requires_grad=True) for k in ("u", "v", "w"))
    loss = wall_shear_loss(ch, u0, v0, w0, nsteps=3, ckpt_every=2)
    (2.0 * loss).backward()
    ref = Channel(prm)
    _set(ref, {k: x[k] for k in ("u", "v", "w", "p")})
    ref.init()
    l2 = ref.wall_shear_gradient(3, ckpt_every=2)
    assert abs(float(loss) - l2) <= 1e-14 * abs(l2)
    for t, k in ((u0, "u"), (v0, "v"), (w0, "w")):
        assert np.array_equal(t.grad.numpy(), 2.0 * ref.get_adjoint(k)), k
    ch.close()
    ref.close()


def test_torch_binding_keeps_device_tensors_on_the_device():
    """The same gradient from tensors on the GPU (device-to-device transfers)
    as from host tensors, bit for bit; the gradients come back on the GPU."""
    from hercules_amd.autograd import wall_shear_loss
    prm = _params("closed", nx=32, ny=32, nz=33)
    x = _state(prm, seed=10)
    out = {}
    for dev in ("cpu", "cuda"):
        ch = Channel(prm)
        _set(ch, {k: x[k] for k in ("u", "v", "w", "p")})
        ch.init()
        ts = [torch.tensor(x[k], device=dev, requires_grad=True) for k in ("u", "v", "w")]
        loss = wall_shear_loss(ch, *ts, nsteps=3, ckpt_every=2)
        loss.backward()
        assert all(t.grad.device.type == dev for t in ts)
        out[dev] = (float(loss), [t.grad.cpu().numpy() for t in ts])
        ch.close()
    assert out["cpu"][0] == out["cuda"][0]
    for a, b in zip(out["cpu"][1], out["cuda"][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kw", [dict(isscalar=1), dict(cds=3, issk=1), dict(precision=1), dict(dp_opt=2)])
def test_adjoint_rejects_unsupported(kw):
    prm = cases.closed_channel(nx=32, ny=32, nz=33, **kw)
    ch = Channel(prm)
    with pytest.raises(HerculesError, match="status 2"):
        ch.adjoint_substep(1)
    ch.close()


@pytest.mark.parametrize("kind,kw", [("closed", {}), ("closed", {"cds": 1}), ("open", {})])
def test_gradient_matches_torch_autograd_reference(kind, kw):
    """The hand-written adjoint against torch.autograd through the FP64
    torch-ops restatement of the step (tests/torch_reference.py, itself
    equal to the oracle to round-off)."""
    torch = pytest.importorskip("torch")
    from tests.torch_reference import TorchChannel
    if kind == "closed":
        prm = cases.closed_channel(nx=16, ny=16, nz=17, wall_rows=1, **kw)
    else:
        prm = _params("open", nx=16, ny=16, nz=17, wall_rows=1, **kw)
    x = _state(prm, seed=12)
    x = {k: x[k] for k in ("u", "v", "w", "p")}
    nsteps = 3
    tc = TorchChannel(prm)
    xt = {k: torch.tensor(v, requires_grad=True) for k, v in x.items()}
    jt = tc.wall_shear(tc.steps(xt, nsteps))
    jt.backward()
    ch = Channel(prm)
    _set(ch, x)
    ch.init()
    loss = ch.wall_shear_gradient(nsteps, ckpt_every=2)
    print(kind, kw, f"loss gpu {loss:.15e} torch {float(jt):.15e}")
    assert abs(loss - float(jt)) <= 1e-12 * abs(loss)
    for k in ("u", "v", "w", "p"):
        g = ch.get_adjoint(k)
        ref = xt[k].grad.numpy()
        err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        print(f"   d loss / d {k}: rel L2 {err:.2e}")
        assert err < 1e-9, (k, err)
    ch.close()

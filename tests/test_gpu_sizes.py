"""Transform lengths 48, 96 and 384 (2^a 3^b, beside the 16..512 powers of
two and 192): the x and y transforms, the right-hand-side tiles of those
widths (16 rows at nx = 48, 8 at 96, the one-pass kernel at 384) and the
column passes, against the oracle (whose DFT is pinned against numpy at
these lengths, tests/test_oracle_pins.py); needs an MI355X."""
import numpy as np
import pytest

from hercules_amd import Channel, cases
from tests.helpers import make_oracle, rel_l2, small_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(48, 96, 17), (96, 48, 33), (384, 16, 17), (16, 384, 17)])
@pytest.mark.parametrize("eq", ["u", "p"])
def test_poisson_solver_fft_new_lengths(shape, eq):
    nx, ny, nz = shape
    prm = cases.closed_channel(nx=nx, ny=ny, nz=nz)
    src = np.random.default_rng(4).standard_normal((nz, ny, nx))
    if eq == "p":
        src -= src.mean(axis=(1, 2), keepdims=True)
    ref = make_oracle(prm, {}).solve(eq, 8.0 / 15.0, src)
    with Channel(prm) as ch:
        got = ch.poisson_solver_fft(eq, 8.0 / 15.0, src)
    err = rel_l2(got, ref)
    print(shape, eq, f"{err:.2e}")
    assert err < (1e-9 if eq == "p" else 1e-13), err


@pytest.mark.parametrize("kind,shape,kw", [
    ("closed", (48, 48, 33), {}),
    ("closed", (96, 96, 33), {}),
    ("closed", (384, 32, 17), {}),
    ("closed", (48, 96, 17), {"cds": 1}),
    ("open_t", (96, 48, 33), {}),
    ("closed", (96, 48, 33), {"cds": 3, "issk": 1}),
    ("open_t", (48, 96, 17), {"cds": 3, "issk": 1}),
    ("closed", (384, 16, 17), {"cds": 3, "issk": 0}),
])
def test_steps_new_lengths_match_oracle(kind, shape, kw):
    nx, ny, nz = shape
    prm, st = small_case(kind, nx=nx, ny=ny, nz=nz, **kw)
    o = make_oracle(prm, st)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    o.steps(2)
    ch.steps(2)
    names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
    for n in names:
        err = rel_l2(ch.get_field(n), o.get(n))
        print(kind, shape, kw, n, f"{err:.2e}")
        assert err < 1e-10, (n, err)
    ch.close()


@pytest.mark.parametrize("shape,kw", [((96, 96, 33), {}), ((48, 96, 33), {"cds": 3, "issk": 1})])
def test_mixed_new_lengths_track_fp64(shape, kw):
    nx, ny, nz = shape
    prm, st = small_case("closed", nx=nx, ny=ny, nz=nz, **kw)
    out = {}
    for pr in (0, 1):
        ch = Channel(prm.replace(precision=pr))
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        ch.init()
        ch.steps(10)
        out[pr] = {n: ch.get_field(n) for n in ("u", "v", "w")}
        ch.close()
    for n in ("u", "v", "w"):
        err = rel_l2(out[1][n], out[0][n])
        print(shape, kw, n, f"{err:.2e}")
        assert err < (1e-5 if n == "u" else 1e-4), (n, err)

"""Screen diagnostics on the GPU: the CFL number of screen_cpu_time
(src/module_io.f90:32-80) against its definition evaluated on the fields,
and the grid of get_mesh_param (geo.info columns); needs an MI355X."""
import numpy as np
import pytest

from hercules_amd import Channel, grid
from tests.helpers import small_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,kw", [("closed", {}), ("open_t", {"u_mrf": 0.4}), ("closed", {"precision": 1})])
def test_cfl_is_its_definition(kind, kw):
    prm, st = small_case(kind, nx=32, ny=32, nz=33, **kw)
    with Channel(prm) as ch:
        for k, v in st.items():
            ch.set_field(k, v)
        ch.init()
        ch.steps(2)
        u, v, w = (ch.get_field(n) for n in "uvw")
        dz = grid.metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)["dz"][:, None, None]
        ref = (prm.dt * (np.abs(u) / (prm.lx / prm.nx) + np.abs(v) / (prm.ly / prm.ny) + np.abs(w) / dz)).max()
        got = ch.screen_cfl()
        print(kind, kw, got, ref)
        assert abs(got - ref) <= 1e-14 * ref


def test_mesh_is_the_grid():
    prm, _ = small_case("open_t", nx=16, ny=16, nz=33)
    with Channel(prm) as ch:
        zc, dz, lt, lb = ch.mesh()
    m = grid.metrics(prm.nz, prm.lz, prm.zstretch, prm.ochannel)
    for a, b in ((zc, m["zc"]), (dz, m["dz"]), (lt, m["l_t"]), (lb, m["l_b"])):
        # std::tanh and numpy tanh may differ in the last bit: dz is a difference of faces
        assert np.allclose(a, b, rtol=1e-11, atol=1e-15)

"""Ensembles of independent channels (hc_ensemble_*): every member's state
after the ensemble steps equals, bit for bit, the same member stepped alone
(needs an MI355X)."""
import numpy as np
import pytest

from hercules_amd import Channel, Ensemble, cases
from tests.helpers import small_case

pytestmark = pytest.mark.gpu


def _members():
    """Uneven grids, both paths, temperature, Coriolis and damping."""
    specs = [
        ("closed", dict(nx=32, ny=32, nz=33)),
        ("open_t", dict(nx=64, ny=32, nz=65)),
        ("closed", dict(nx=16, ny=16, nz=129, dp_opt=3, fc=0.5, ug=1.0, vg=0.2)),
        ("closed", dict(nx=32, ny=32, nz=17, cds=3, issk=1)),
        ("open_t", dict(nx=32, ny=32, nz=33, isdamp=1, nzdamp=8, cdamp=0.2)),
        ("closed", dict(nx=16, ny=32, nz=17, dp_opt=2, dts=1)),
    ]
    return [small_case(kind, **kw) for kind, kw in specs]


def _channel(prm, st):
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    return ch


@pytest.mark.parametrize("use_graph,nsteps", [(True, 5), (False, 3)])
def test_ensemble_members_equal_single_runs(use_graph, nsteps):
    cases_ = _members()
    singles = [_channel(p, s) for p, s in cases_]
    for ch in singles:
        ch.steps(nsteps)
    members = [_channel(p, s) for p, s in cases_]
    with Ensemble(members, use_graph=use_graph) as ens:
        ens.steps(nsteps)
        assert ens.graph_used == use_graph
    for (prm, _), a, b in zip(cases_, singles, members):
        names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
        for n in names:
            x, y = a.get_field(n), b.get_field(n)
            assert np.isfinite(x).all()
            assert np.array_equal(x, y), (prm.nx, prm.ny, prm.nz, prm.cds, n, np.abs(x - y).max())
        assert a.scalar("stps") == b.scalar("stps") == nsteps
    for ch in singles + members:
        ch.close()


def test_ensemble_graph_replays_continue_the_run():
    """Two ensemble calls (graph captured once, replayed) equal one run."""
    cases_ = _members()[:3]
    singles = [_channel(p, s) for p, s in cases_]
    for ch in singles:
        ch.steps(8)
    members = [_channel(p, s) for p, s in cases_]
    with Ensemble(members) as ens:
        ens.steps(4)
        ens.steps(4)
    for (prm, _), a, b in zip(cases_, singles, members):
        for n in ("u", "v", "w", "p"):
            assert np.array_equal(a.get_field(n), b.get_field(n)), (prm.nx, n)
    for ch in singles + members:
        ch.close()

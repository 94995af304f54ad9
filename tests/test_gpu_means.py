"""Mean statistics on the GPU (hc_mean_fields / hc_get_mean_field) against
the oracle's calculate_mean_fields(_spec), itself pinned by the step sums
(tests/test_oracle_means.py); needs an MI355X."""
import numpy as np
import pytest

from hercules_amd import Channel
from hercules_amd.channel import HerculesError, MEANS
from tests.helpers import make_oracle, rel_l2, small_case

pytestmark = pytest.mark.gpu


def _pair(prm, st):
    o = make_oracle(prm, st)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    return o, ch


@pytest.mark.parametrize("kind,kw", [
    ("closed", {}),
    ("closed", {"cds": 1, "u_mrf": 0.3}),
    ("open_t", {}),
    ("open_t", {"cds": 3, "issk": 1}),
    ("closed", {"dts": 1}),
    ("closed", {"isdamp": 1, "nzdamp": 8, "cdamp": 0.3}),
    ("open_t", {"tbc_top": 1, "dtdz_top": 0.5, "cds": 1}),
])
def test_means_match_oracle(kind, kw):
    prm, st = small_case(kind, nx=32, ny=32, nz=33, **kw)
    o, ch = _pair(prm, st)
    o.mean_enable()
    ch.mean_fields(True)
    o.steps(3)
    ch.steps(3)
    assert ch.scalar("mean_steps") == 3
    for name in MEANS:
        ref = o.get_mean(name)
        got = ch.get_mean(name)
        if not np.any(ref):
            assert np.abs(got).max() < 1e-300 + 1e-12 * max(np.abs(o.get_mean("uu")).max(), 1.0), name
            continue
        err = rel_l2(got, ref)
        print(kind, kw, name, f"{err:.2e}")
        assert err < 1e-10, (name, err)
    ch.close()


def test_means_mixed_precision_track_fp64():
    prm, st = small_case("open_t", nx=32, ny=32, nz=33)
    out = {}
    for pr in (0, 1):
        ch = Channel(prm.replace(precision=pr))
        for k, v in st.items():
            ch.set_field(k, v)
        ch.init()
        ch.mean_fields(True)
        ch.steps(4)
        out[pr] = {n: ch.get_mean(n) for n in ("u", "uu", "uw", "tt")}
        ch.close()
    for n in out[0]:
        assert rel_l2(out[1][n], out[0][n]) < 1e-4, n


def test_means_stop_restart_and_errors():
    prm, st = small_case("closed", nx=32, ny=32, nz=33)
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    with pytest.raises(HerculesError):
        ch.get_mean("u")  # nothing accumulated yet
    ch.mean_fields(True)
    ch.steps(2)
    u2 = ch.get_field("u")
    ch.mean_fields(False)
    ch.steps(2)  # not added
    assert ch.scalar("mean_steps") == 2
    ch.mean_fields(True)  # zeroed again
    ch.steps(1)
    assert ch.scalar("mean_steps") == 1
    assert np.allclose(ch.get_mean("u"), ch.get_field("u"), rtol=0, atol=1e-13 * np.abs(u2).max())
    ch.close()


def test_means_of_ensemble_members_equal_single_runs():
    """Statistics accumulated inside the replayed ensemble graph equal those
    of the same channels stepped alone (same kernels, bit for bit), also when
    a member's statistics are switched on after the first capture."""
    from hercules_amd import Ensemble
    specs = [("closed", dict(nx=32, ny=32, nz=33)), ("open_t", dict(nx=32, ny=32, nz=33, cds=3, issk=1))]
    cs = [small_case(kind, **kw) for kind, kw in specs]

    def make(prm, st):
        ch = Channel(prm)
        for k, v in st.items():
            ch.set_field(k, v)
        ch.init()
        return ch

    singles = [make(p, s) for p, s in cs]
    for ch in singles:
        ch.steps(2)
        ch.mean_fields(True)
        ch.steps(4)
    members = [make(p, s) for p, s in cs]
    with Ensemble(members, use_graph=True) as ens:
        ens.steps(2)  # captured without statistics
        for ch in members:
            ch.mean_fields(True)
        ens.steps(4)  # captured again with them
        assert ens.graph_used
    for (prm, _), a, b in zip(cs, singles, members):
        assert a.scalar("mean_steps") == b.scalar("mean_steps") == 4
        for n in MEANS:
            assert np.array_equal(a.get_mean(n), b.get_mean(n)), (prm.cds, n)
    for ch in singles + members:
        ch.close()

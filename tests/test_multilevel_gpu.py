"""multilevel-hip on the GPU against the oracle's tick loop
(crates/graph-compute/src/engines/multilevel.rs:333-468 over coarsen.rs): the
whole descent -- coarsest-level solve, prolongations, refinement at every
level -- bit-identical with fa2-bh-hip and with sgd-stress-hip inside, plus the
settings errors of multilevel.rs:307-331."""
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_settings_errors_and_descriptor():
    e = jc.MultilevelEngine()
    assert e.descriptor["id"] == "multilevel-hip"
    for bad, msg in ((dict(inner="multilevel"), "cannot wrap itself"),
                     (dict(inner="multilevel-hip"), "cannot wrap itself"),
                     (dict(inner="no-such-engine"), "unknown inner engine"),
                     (dict(sweep_schedule="cubic"), "unknown variant"),
                     (dict(max_levels=-1), "invalid type")):
        with pytest.raises(jc.JcError) as err:
            e.set_params(bad)
        assert err.value.code == jc.JC_E_INVALID and msg in str(err.value)
    e.set_params(dict(inner="sgd-stress-hip"))
    e.set_params(None)  # keeps the settings
    with pytest.raises(jc.JcError):
        e.step(1)  # before init
    # inner_params reach the inner engine at init (a bad one fails there)
    e.set_params(dict(inner="fa2-bh-hip", inner_params={"theta": "wide"}))
    off, nb, xyz = workloads.sbm_2000()
    with pytest.raises(jc.JcError) as err:
        e.init(off, nb, xyz)
    assert "theta" in str(err.value)


SET = dict(max_levels=4, target_size=200, sweeps_per_level=5, sweep_schedule="linear", spring_len=2.0,
           seed=77)


def test_fa2_inner_descent_bit_exact():
    off, nb, xyz = workloads.sbm_2000()
    fa2 = O.Fa2Settings.make(theta=1.2)
    e = jc.MultilevelEngine(inner="fa2-bh-hip", inner_params={"theta": 1.2}, **SET)
    e.init(off, nb, xyz)
    st = e.state()
    assert st["level"] == st["n_levels"] - 1 >= 2
    seen = set()
    total = 0
    for k in (7, 13, 1, 19, 40):  # chunks that stop inside levels and at level changes
        got = e.step(k)
        total += k
        want, level, n_levels = O.multilevel_run(off, nb, xyz, total, SET, "fa2-bh-hip", fa2=fa2)
        st = e.state()
        assert (st["level"], st["n_levels"]) == (level, n_levels)
        assert len(got) == 3 * st["n_level_nodes"] == len(want)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (total, np.abs(got - want).max())
        seen.add(level)
    assert 0 in seen and len(seen) >= 3


def test_sgd_inner_descent_bit_exact():
    off, nb, xyz = workloads.sbm_2000()
    sgd = dict(n_pivots=16, eta_max=0.0)
    e = jc.MultilevelEngine(inner="sgd-stress-hip", inner_params=sgd, **SET)
    e.init(off, nb, xyz)
    got = e.step(55)
    want, level, _ = O.multilevel_run(off, nb, xyz, 55, SET, "sgd-stress-hip", sgd=sgd)
    assert e.state()["level"] == level == 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_brute_inner_descent_bit_exact():
    """The default inner engine (fa2-brute-hip, multilevel.rs:176) down the cascade."""
    off, nb, xyz = workloads.sbm_2000()
    e = jc.MultilevelEngine(**SET)
    e.init(off, nb, xyz)
    total = 0
    for k in (11, 30, 40):
        got = e.step(k)
        total += k
        want, level, n_levels = O.multilevel_run(off, nb, xyz, total, SET)
        st = e.state()
        assert (st["level"], st["n_levels"]) == (level, n_levels)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (total, np.abs(got - want).max())
    assert e.state()["level"] == 0


def test_small_graph_is_the_plain_inner_engine():
    """n <= target_size: one level, the wrapper forwards every tick."""
    off, nb = O.ba_graph(300, 3, 4)
    xyz = O.uniform_cube(300, 5.0, 2)
    e = jc.MultilevelEngine()  # defaults: fa2-brute-hip, target_size 500
    e.init(off, nb, xyz)
    assert e.state()["n_levels"] == 1
    got = e.step(25)
    want, _, _, _ = O.fa2_run(off, nb, xyz, 25, O.Fa2Settings.make(), tree_kind=2, exact=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    with pytest.raises(jc.JcError):
        e.get_state()

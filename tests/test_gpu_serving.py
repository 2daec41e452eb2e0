"""Continuous-batching server (lbm3_server_*): every request's result is the
same whether it runs alone or batched with others that arrive at different
times, matches the single-simulation path and the oracle, and admission keeps
to the HBM budget."""
import numpy as np
import pytest

from lbm3_amd import cases
from lbm3_amd import lib
from tests.helpers import rel_max
from tests.oracle_lib import OracleSim

pytestmark = pytest.mark.gpu


def _tgv(n=16):
    return cases.taylor_green(n, u0=0.05, reynolds=400.0)


def _channel():
    return cases.channel(40, 12, 12, force=5e-6, nu=0.05)


def _sphere():
    return cases.sphere(48, 24, 24, diameter=6.0, center=(12.0, 12.0, 12.0), u_inlet=0.05, reynolds=80.0,
                        segments=(16, 32))


CASES = {"tgv": _tgv, "channel": _channel, "sphere": _sphere}


def _cells(make):
    case, dom, tris = make()
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    n = s.info(0).cellCount
    s.close()
    return n


def _drain(srv, expect):
    """Step until every id in `expect` is reported; returns the completion order."""
    order = []
    for _ in range(100000):
        order += srv.step()
        if set(expect) <= set(order):
            return order
    raise AssertionError("server did not finish %s" % sorted(set(expect) - set(order)))


def _alone(make, steps, storage):
    srv = lib.Server(8 << 30, 8)
    case, dom, tris = make()
    rid = srv.submit(case, dom, tris, steps, storage, keep_macros=True)
    _drain(srv, [rid])
    m = srv.macros(rid, _cells(make))
    res = srv.result(rid)
    srv.close()
    return m, res


@pytest.mark.parametrize("name", sorted(CASES))
def test_served_matches_single_simulation(name):
    make = CASES[name]
    steps = 60
    m, res = _alone(make, steps, "f32")
    assert res["status"] == 0 and res["steps"] == steps
    case, dom, tris = make()
    s = lib.Sim(case, "f32")
    fu = s.build(dom, tris)
    s.init()
    s.run(steps)
    ref = s.macros(0)
    ke = s.kinetic_energy()
    s.close()
    assert res["fluid_updates"] == fu * steps
    # the server applies the open boundary as its own pass, the single path folds
    # it into the next step's loads: the same arithmetic, up to contraction
    assert rel_max(m[:, 1:], ref[:, 1:]) < 1e-6
    assert np.max(np.abs(m[:, 0] - ref[:, 0])) < 1e-6
    assert abs(res["kinetic_energy"] / ke - 1.0) < 1e-5


def test_served_taylor_green_matches_oracle():
    """The hot-path parity bar through the server: FP32 rho, u within 1e-5
    relative max-norm of the oracle after 100 steps."""
    case, dom, tris = _tgv(16)
    m, _ = _alone(lambda: (case, dom, tris), 100, "f32")
    o = OracleSim(case, dom, tris, double=False)
    o.init()
    o.run(100)
    mo = o.macros(0)
    assert rel_max(1.0 + m[:, 0], 1.0 + mo[:, 0]) < 1e-5
    assert rel_max(m[:, 1:], mo[:, 1:]) < 1e-5


def test_mix_is_bitwise_identical_to_alone():
    """Requests of every kind and four storage types, submitted while others
    run: each result equals the same request solved alone, bit for bit."""
    plan = [("sphere", "f32", 70, 0), ("tgv", "f32", 50, 0), ("channel", "f16", 40, 5), ("tgv", "f16", 30, 12),
            ("channel", "f32", 45, 12), ("sphere", "f16", 25, 20), ("tgv", "f64", 20, 33), ("sphere", "f16c", 30, 8)]
    alone = {}
    for k, (name, st, steps, _) in enumerate(plan):
        alone[k] = _alone(CASES[name], steps, st)[0]
    srv = lib.Server(8 << 30, 16)
    ids = {}
    order = []
    tick = 0
    while len(order) < len(plan):
        for k, (name, st, steps, at) in enumerate(plan):
            if at == tick:
                case, dom, tris = CASES[name]()
                ids[srv.submit(case, dom, tris, steps, st, keep_macros=True)] = k
        order += srv.step()
        tick += 1
        assert tick < 10000
    assert sorted(ids) == sorted(order)
    for rid, k in ids.items():
        name = plan[k][0]
        m = srv.macros(rid, _cells(CASES[name]))
        assert m.dtype == alone[k].dtype
        assert np.array_equal(m, alone[k]), plan[k]
    srv.close()


def test_admission_keeps_to_the_budget():
    """A budget that holds one request at a time: requests queue, run one after
    another, and all complete; a request larger than the budget is refused."""
    srv = lib.Server(90 << 20, 16)
    ids = []
    for _ in range(3):
        case, dom, tris = _tgv(16)
        ids.append(srv.submit(case, dom, tris, 10, "f32"))
    assert srv.stats()["queued"] == 3
    done = []
    peak = 0
    while len(done) < 3:
        done += srv.step()
        peak = max(peak, srv.stats()["active"])
    assert peak <= 1 and done == ids
    st = srv.stats()
    assert st["active"] == 0 and st["queued"] == 0 and st["bytes"] == 0
    for rid in ids:
        assert srv.result(rid)["status"] == 0
    case, dom, tris = cases.sphere(256, 128, 128)
    with pytest.raises(lib.LBM3Error):
        srv.submit(case, dom, tris, 10, "f32")
    srv.close()


def test_server_rejects_multilevel_and_unfinished_queries():
    srv = lib.Server(1 << 30, 4)
    case, dom, tris = cases.taylor_green(16, levels=2)
    with pytest.raises(lib.LBM3Error):
        srv.submit(case, dom, tris, 10, "f32")
    case, dom, tris = _tgv(16)
    rid = srv.submit(case, dom, tris, 5, "f32")
    with pytest.raises(lib.LBM3Error):
        srv.result(rid)
    _drain(srv, [rid])
    assert srv.result(rid)["status"] == 0
    with pytest.raises(lib.LBM3Error):
        srv.macros(rid, 10)          # macros were not kept
    srv.close()


def test_serving_runners_complete_a_small_mix():
    """The seeded workload through the served and the one-at-a-time runners:
    every request finishes and both count the same fluid updates."""
    from lbm3_amd import serving
    reqs = serving.make_requests(n=6, seed=3, steps_scale=0.02)
    a = serving.run_served(reqs, "f32")
    b = serving.run_sequential(reqs, "f32")
    assert a["requests"] == b["requests"] == 6
    assert a["p99_s"] >= a["p50_s"] > 0 and b["p99_s"] >= b["p50_s"] > 0
    assert abs(a["mlups"] * a["makespan_s"] / (b["mlups"] * b["makespan_s"]) - 1.0) < 1e-9

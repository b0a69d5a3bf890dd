"""Continuous batching on the GPU: step budgets, slots longer than their rods,
device rod copies, and the server's admission / completion loop.

The bar is bitwise: a rod's fused-path arithmetic depends only on its own
data, dt and number of steps, so a request served in a mix of others must
equal the same request served alone, and equal its rods stepped directly.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64, DeviceBlock, ErodError, Server
from elasticapp_amd.workloads import Workload, build_block
from tests.gpu_helpers import STATE, random_workload

pytestmark = pytest.mark.gpu


def _library(n_rods=24, seed=41):
    wl = random_workload(n_rods, 8, 90, seed=seed, steps=0)
    blk = build_block(wl)
    dts = [np.float64(wl.dt * (1.0 + 0.01 * r)) for r in range(n_rods)]   # per-rod dt
    blk.set_rod_dt(np.array(dts))
    return wl, blk


def _direct(wl, rods, steps, precision, dts, steps_per_launch=100):
    """The given library rods stepped `steps` fused steps in a batch of their
    own, in launches of `steps_per_launch` steps (FP32-bulk results depend on
    the launch segmentation at rounding level)."""
    sub = Workload("sub", [wl.n_elems[r] for r in rods], [wl.states[r] for r in rods], wl.dt, steps,
                   forces=[(i,) + tuple(wl.forces[r][1:]) for i, r in enumerate(rods)])
    blk = build_block(sub)
    blk.set_rod_dt(np.array([dts[r] for r in rods]))
    blk.set_fused_config(0, 0, steps_per_launch)
    blk.step_fused(steps, precision)
    return blk


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_step_budgets(precision):
    wl = random_workload(6, 10, 60, seed=43, steps=0)
    a = build_block(wl)
    budgets = [0, 7, 40, 100, 250, None]
    for r, b in enumerate(budgets):
        a.set_step_budget(r, b)
    a.set_fused_config(0, 0, 30)
    a.step_fused(120, precision)           # 4 launches of 30 steps
    assert [a.step_budget(r) for r in range(6)] == [0, 0, 0, 0, 130, None]
    for r, b in enumerate(budgets):
        ref = build_block(Workload("one", [wl.n_elems[r]], [wl.states[r]], wl.dt, 0,
                                   forces=[(0,) + tuple(wl.forces[r][1:])]))
        ref.set_fused_config(0, 0, 30)
        n = 120 if b is None else min(b, 120)
        if n:
            ref.step_fused(n, precision)
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), ref.get_rod(t, 0)), (r, t)
        ref.close()
    # the graph path spends budgets the same way
    g = build_block(wl)
    for r, b in enumerate(budgets):
        g.set_step_budget(r, b)
    g.set_fused_config(0, 0, 30)
    g.step_fused_graph(120, 60, precision)
    assert [g.step_budget(r) for r in range(6)] == [0, 0, 0, 0, 130, None]
    for r in range(6):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), g.get_rod(t, r)), (r, t)


def test_copy_into_longer_slots_and_active_length():
    wl, lib = _library()
    dts = [float(wl.dt * (1.0 + 0.01 * r)) for r in range(len(wl.n_elems))]
    pool = DeviceBlock([128] * 10, damped=True)
    rods = [3, 0, 17, 5]
    pool.copy_rods_from(lib, [9, 2, 4, 7], rods)
    for slot, r in zip([9, 2, 4, 7], rods):
        assert pool.active_length(slot) == wl.n_elems[r]
        for tag in ("Position", "Director", "Mass", "ReferenceCurvature", "BendingTwistRigidityMatrix"):
            assert np.array_equal(pool.get_rod(tag, slot), lib.get_rod(tag, r)), (slot, tag)
    assert pool.active_length(0) == 128
    for slot in (0, 1, 3, 5, 6, 8):
        pool.set_active_length(slot, 0)
    with pytest.raises(ErodError):
        pool.step(wl.dt, 1)                  # the op-by-op path needs full rods
    pool.step_fused(50, FP32_BULK)
    ref = _direct(wl, rods, 50, FP32_BULK, dts)
    for i, slot in enumerate([9, 2, 4, 7]):
        for t in STATE:
            assert np.array_equal(pool.get_rod(t, slot), ref.get_rod(t, i)), (slot, t)


@pytest.mark.parametrize("precision", [FP32_BULK, FP64])
def test_server_mix_equals_alone(precision):
    wl, lib = _library()
    dts = [float(wl.dt * (1.0 + 0.01 * r)) for r in range(len(wl.n_elems))]
    rng = np.random.default_rng(47)
    reqs = []
    for i in range(20):
        rods = list(rng.integers(0, len(wl.n_elems), rng.integers(1, 6)))
        reqs.append((i, rods, int(rng.integers(20, 400))))
    classes = [(32, 12), (64, 16), (128, 16)]
    srv = Server(lib, classes, precision=precision, steps_per_tick=50, hbm_budget_bytes=2e9)
    st = srv.status()
    assert st["pool_bytes"] < 2e9 and st["queued"] == 0
    done, results = set(), {}
    for rid, rods, steps in reqs[:12]:
        srv.submit(rid, rods, steps)
    ticks = 0
    while len(done) < len(reqs):
        if ticks == 3:
            for rid, rods, steps in reqs[12:]:
                srv.submit(rid, rods, steps)
        for rid in srv.tick():
            results[rid] = [{t: srv.read(rid, k, t) for t in STATE} for k in range(len(reqs[rid][1]))]
            srv.release(rid)
            done.add(rid)
        ticks += 1
        assert ticks < 1000
    # every rod advanced exactly its request's steps
    n_ticks, element_steps = srv.stats()
    assert 0 < n_ticks <= ticks
    assert element_steps == sum(sum(wl.n_elems[r] for r in rods) * steps for _, rods, steps in reqs)
    # each request: equal to the same request served alone, and to its rods stepped directly
    for rid, rods, steps in (reqs[0], reqs[7], reqs[15]):
        alone = Server(lib, classes, precision=precision, steps_per_tick=50, hbm_budget_bytes=2e9)
        alone.submit(rid, rods, steps)
        got = []
        while not got:
            got = alone.tick()
        ref = _direct(wl, rods, steps, precision, dts, steps_per_launch=50)
        for k in range(len(rods)):
            for t in STATE:
                a = alone.read(rid, k, t)
                assert np.array_equal(results[rid][k][t], a), (rid, k, t)
                assert np.array_equal(a, ref.get_rod(t, k)), (rid, k, t)
        alone.close()
    srv.close()


def test_server_rejects_what_cannot_fit():
    wl, lib = _library()
    srv = Server(lib, [(64, 2)], precision=FP32_BULK, steps_per_tick=10, hbm_budget_bytes=1e9)
    long_rods = [r for r, n in enumerate(wl.n_elems) if n > 64]
    short = [r for r, n in enumerate(wl.n_elems) if n <= 64]
    if long_rods:
        with pytest.raises(ErodError):
            srv.submit(1, [long_rods[0]], 10)
    with pytest.raises(ErodError):
        srv.submit(2, short[:3], 10)         # three rods, two slots
    with pytest.raises(ErodError):
        Server(lib, [(1024, 100000)], precision=FP32_BULK, hbm_budget_bytes=1e9)   # over budget
    srv.close()


def test_server_steps_in_the_library_environment():
    """Gravity set on the library after the server exists reaches the pool at
    the next tick: a served request equals its rods stepped directly in a
    batch with the same gravity."""
    wl, lib = _library(n_rods=6, seed=43)
    dts = [float(wl.dt * (1.0 + 0.01 * r)) for r in range(len(wl.n_elems))]
    srv = Server(lib, [(128, 8)], precision=FP64, steps_per_tick=50, hbm_budget_bytes=1e9)
    lib.set_gravity([0.0, 0.0, -9.81])
    rods, steps = [0, 2, 5], 120
    srv.submit(0, rods, steps)
    got = []
    while not got:
        got = srv.tick()
    sub = Workload("sub", [wl.n_elems[r] for r in rods], [wl.states[r] for r in rods], wl.dt, steps,
                   forces=[(i,) + tuple(wl.forces[r][1:]) for i, r in enumerate(rods)], gravity=(0.0, 0.0, -9.81))
    ref = build_block(sub)
    ref.set_rod_dt(np.array([dts[r] for r in rods]))
    ref.step_fused(steps, FP64)
    for k in range(len(rods)):
        for t in STATE:
            assert np.array_equal(srv.read(0, k, t), ref.get_rod(t, k)), (k, t)
    srv.close()


def test_pool_view_reads_fields_and_dies_with_the_server():
    """The server's pool reads as a DeviceBlock (a whole field included) while
    the server lives, and refuses calls once it is closed."""
    wl, lib = _library(n_rods=4, seed=44)
    srv = Server(lib, [(128, 4)], precision=FP64, steps_per_tick=20, hbm_budget_bytes=1e9)
    srv.submit(0, [1], 40)
    while not srv.tick():
        pass
    x = srv.pool.get_field("Position")
    assert x.shape == (3, srv.pool.slots(0)) and np.isfinite(x).all()
    srv.close()
    with pytest.raises(ErodError):
        srv.pool.get_field("Position")

"""Admission against the HBM budget (erod_server_create_budgeted), and
admission past a waiting request (both kinds of server).

  * a budgeted server starts with a few slots per class and grows its pool
    while requests wait, as far as its budget allows; the pool never passes
    the budget, results do not depend on when it grew (each request equals
    the same request served alone, bitwise), and a request that could not
    fit the budget even in an otherwise empty pool is refused at submission;
  * the earliest waiting request reserves the classes it needs: a later
    request that fits other classes is admitted past it, and one that needs
    the reserved classes waits behind it.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64, ErodError, Server
from tests.gpu_helpers import STATE
from tests.test_gpu_serve import _library

pytestmark = pytest.mark.gpu


def _run(srv, reqs, submit_at=None, limit=2000):
    """Serve (rid, rods, steps) requests; `submit_at[rid]` = the tick before
    which a request is submitted (default 0).  Returns the results, the tick
    each request finished at, and the largest pool seen."""
    submit_at = submit_at or {}
    results, finished, ticks, pool_max = {}, {}, 0, 0
    pending = list(reqs)
    while len(results) < len(reqs):
        for item in [r for r in pending if submit_at.get(r[0], 0) <= ticks]:
            srv.submit(*item)
            pending.remove(item)
        for rid in srv.tick():
            rods = [r for i, r, _ in reqs if i == rid][0]
            results[rid] = [{t: srv.read(rid, k, t) for t in STATE} for k in range(len(rods))]
            finished[rid] = ticks
            srv.release(rid)
        pool_max = max(pool_max, srv.status()["pool_bytes"])
        ticks += 1
        assert ticks < limit
    return results, finished, pool_max


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_budgeted_pool_grows_and_results_do_not_change(precision):
    wl, lib = _library()
    rng = np.random.default_rng(53)
    reqs = [(i, [int(r) for r in rng.integers(0, 24, int(rng.integers(2, 9)))], int(rng.integers(30, 300)))
            for i in range(40)]
    srv = Server.budgeted(lib, [32, 64, 128], precision=precision, steps_per_tick=50, hbm_budget_bytes=2e9)
    info0 = srv.pool_info()
    assert info0["slots"] == 48 and info0["growths"] == 0
    results, _, pool_max = _run(srv, reqs)
    for rid, res in results.items():
        for k, st in enumerate(res):
            assert all(np.isfinite(a).all() for a in st.values()), (rid, k)
    info = srv.pool_info()
    # 40 requests of 2-8 rods submitted at once: the pool grew to hold them
    assert info["growths"] >= 1 and info["slots"] > 48, info
    assert pool_max <= 2e9
    assert len(srv.pool.n_elems) == info["slots"]      # the pool view follows the rebuilt pool
    srv.close()
    for rid in (0, 17, 39):
        rods, steps = reqs[rid][1], reqs[rid][2]
        alone = Server(lib, [(32, 16), (64, 16), (128, 16)], precision=precision, steps_per_tick=50,
                       hbm_budget_bytes=2e9)
        got, _, _ = _run(alone, [(rid, rods, steps)])
        for k in range(len(rods)):
            for t in STATE:
                assert np.array_equal(results[rid][k][t], got[rid][k][t]), (rid, k, t)
        alone.close()


def test_pool_grows_while_requests_run():
    """Requests arrive over several ticks, so that the pool grows while
    earlier requests are half way through their steps: the rods moved into
    the new pool continue bit for bit (each request equals the same request
    served alone)."""
    wl, lib = _library()
    rng = np.random.default_rng(59)
    reqs = [(i, [int(r) for r in rng.integers(0, 24, int(rng.integers(3, 9)))], int(rng.integers(150, 400)))
            for i in range(36)]
    srv = Server.budgeted(lib, [32, 64, 128], precision=FP32_BULK, steps_per_tick=50, hbm_budget_bytes=2e9)
    results, finished, _ = _run(srv, reqs, submit_at={i: i // 6 for i in range(36)})
    info = srv.pool_info()
    srv.close()
    assert info["growths"] >= 2, info
    for rid in (0, 5, 11, 23, 35):
        rods, steps = reqs[rid][1], reqs[rid][2]
        alone = Server(lib, [(32, 16), (64, 16), (128, 16)], precision=FP32_BULK, steps_per_tick=50,
                       hbm_budget_bytes=2e9)
        got, _, _ = _run(alone, [(rid, rods, steps)])
        for k in range(len(rods)):
            for t in STATE:
                assert np.isfinite(got[rid][k][t]).all()
                assert np.array_equal(results[rid][k][t], got[rid][k][t]), (rid, k, t)
        alone.close()


def test_pool_settings_survive_growth():
    """The tile shape and the re-orthonormalisation schedule set on the pool
    carry over to the pool a growth builds."""
    wl, lib = _library()
    srv = Server.budgeted(lib, [32, 64, 128], precision=FP32_BULK, steps_per_tick=50, hbm_budget_bytes=2e9)
    srv.pool.set_fused_config(256, 1, 50)
    srv.pool.set_reortho_interval(100)
    rng = np.random.default_rng(61)
    for i in range(40):
        srv.submit(i, [int(r) for r in rng.integers(0, 24, int(rng.integers(2, 9)))], int(rng.integers(120, 200)))
    srv.tick()  # admits what fits, growing the pool, and steps it
    assert srv.pool_info()["growths"] >= 1
    assert srv.pool.reortho_interval()[0] == 100
    assert srv.pool.fused_config() == (256, 1, 50)
    done = 0
    for _ in range(200):
        for rid in srv.tick():
            assert np.isfinite(srv.read(rid, 0, "Position")).all()
            srv.release(rid)
            done += 1
        if done == 40:
            break
    assert done == 40
    srv.close()


def test_budget_bounds_the_pool():
    """A budget that holds a little more than the initial pool: the pool
    grows once or twice, never past the budget, and the requests beyond it
    wait for slots to be released."""
    wl, lib = _library()
    small = [r for r, n in enumerate(wl.n_elems) if n <= 32]
    assert len(small) >= 4
    probe = Server.budgeted(lib, [32], precision=FP32_BULK, steps_per_tick=20, hbm_budget_bytes=1e9)
    base = probe.status()["pool_bytes"]      # 16 slots of 32 elements
    probe.close()
    budget = 2.6 * base                       # room for at most 41 slots
    srv = Server.budgeted(lib, [32], precision=FP32_BULK, steps_per_tick=20, hbm_budget_bytes=budget)
    reqs = [(i, [small[i % len(small)]] * 4, 40 + 20 * (i % 3)) for i in range(30)]   # 120 rods
    results, finished, pool_max = _run(srv, reqs)
    assert pool_max <= budget
    info = srv.pool_info()
    # 16 slots doubled to 32, then grown by what the next requests lacked
    assert info["growths"] == 3 and info["slots"] == 40, info
    # 120 rods through at most 40 slots: several waves
    assert max(finished.values()) >= 3 * 2
    with pytest.raises(ErodError, match="budget"):
        srv.submit(99, [small[0]] * 42, 10)     # 42 slots of 32 elements pass the budget
    srv.close()


def test_later_request_admitted_past_a_waiting_one():
    wl, lib = _library()
    short = [r for r, n in enumerate(wl.n_elems) if n <= 32]
    long_ = [r for r, n in enumerate(wl.n_elems) if 32 < n <= 128]
    assert short and long_
    # two classes; the long class has 2 slots
    srv = Server(lib, [(32, 8), (128, 2)], precision=FP32_BULK, steps_per_tick=20, hbm_budget_bytes=1e9)
    reqs = [
        (0, [long_[0], long_[0]], 200),   # takes both long slots for 10 ticks
        (1, [long_[0]], 20),              # waits for a long slot: reserves the long class
        (2, [short[0]], 20),              # short class only: admitted past request 1
        (3, [long_[0]], 20),              # needs the reserved class: waits behind request 1
    ]
    results, finished, _ = _run(srv, reqs)
    assert finished[2] == 0                      # admitted at once, done in its first tick
    assert finished[1] >= finished[0]            # waited for request 0's slots
    assert finished[3] >= finished[1]            # never overtook the earlier request 1
    srv.close()

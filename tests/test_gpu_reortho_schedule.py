"""Re-orthonormalisation of the directors scheduled on the fused paths
(erod_set_reortho_interval): after the launch during which the batch's fused
step count reaches a multiple of the interval, one Newton-Schulz polar step
on the directors of every rod with steps left.

  * direct launches: equal, bitwise, to stepping by the interval and calling
    erod_reorthonormalize between (same Newton-Schulz arithmetic);
  * captured graphs: a chunk that is a multiple of the interval, an interval
    that is a multiple of the chunk, and neither (the graph is captured again
    at each new phase), all bitwise equal to the direct launches, with the
    phase carried across calls;
  * a rod whose step budget is spent keeps its state bit for bit;
  * over 10,000 FP32-bulk steps with a step every 1,000, directors stay
    orthonormal to 1e-13.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64
from elasticapp_amd.workloads import build_block
from tests.gpu_helpers import STATE, random_workload

pytestmark = pytest.mark.gpu


def _as_matrices(d):
    return d.T.reshape(-1, 3, 3)


def _equal(a, b, rods):
    for r in rods:
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_schedule_equals_manual_steps(precision):
    wl = random_workload(12, 10, 300, seed=701, steps=0)
    a, b = build_block(wl), build_block(wl)
    for blk in (a, b):
        blk.set_fused_config(0, 0, 50)
    a.set_reortho_interval(200)
    a.step_fused(600, precision)
    assert a.reortho_interval() == (200, 0)
    for _ in range(3):
        b.step_fused(200, precision)
        b.reorthonormalize()
    _equal(a, b, range(12))
    # the schedule really acted: without it the directors differ
    c = build_block(wl)
    c.set_fused_config(0, 0, 50)
    c.step_fused(600, precision)
    assert any(not np.array_equal(c.get_rod("Director", r), a.get_rod("Director", r)) for r in range(12))
    for blk in (a, b, c):
        blk.close()


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
@pytest.mark.parametrize("every,chunk", [(100, 300), (300, 100), (250, 100), (70, 100)])
def test_graph_schedule_equals_direct(precision, every, chunk):
    wl = random_workload(10, 20, 400, seed=703, steps=0)
    a, b = build_block(wl), build_block(wl)
    for blk in (a, b):
        blk.set_fused_config(0, 0, 50)
        blk.set_reortho_interval(every)
    # two calls, so that the phase carries from one to the next
    a.step_fused(650, precision)
    a.step_fused(550, precision)
    b.step_fused_graph(650, chunk, precision)
    b.step_fused_graph(550, chunk, precision)
    assert a.reortho_interval() == b.reortho_interval()
    _equal(a, b, range(10))
    for blk in (a, b):
        blk.close()


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_schedule_over_rod_pieces(precision):
    """Rods longer than the tile run as pieces (several launches per chunk,
    two state copies): the scheduled step equals stepping by the interval and
    calling erod_reorthonormalize between, bitwise."""
    wl = random_workload(5, 300, 1500, seed=711, steps=0)
    a, b = build_block(wl), build_block(wl)
    for blk in (a, b):
        blk.set_fused_config(256, 1, 40)
        blk.set_fused_pieces(256, 1, 16)
    a.set_reortho_interval(80)
    a.step_fused_graph(240, 80, precision)
    for _ in range(3):
        b.step_fused(80, precision)
        b.reorthonormalize()
    assert a.fused_pieces()[4] > 5
    _equal(a, b, range(5))
    for blk in (a, b):
        blk.close()


def test_spent_budgets_keep_their_state():
    wl = random_workload(6, 30, 200, seed=705, steps=0)
    blk = build_block(wl)
    blk.set_fused_config(0, 0, 40)
    blk.set_reortho_interval(80)
    budgets = [None, 40, 0, 120, 80, None]
    for r, n in enumerate(budgets):
        blk.set_step_budget(r, n)
    # the rods stepped alone, each with its budget and the same schedule
    # while it still has steps left
    blk.step_fused(40, FP32_BULK)
    frozen = {r: {t: blk.get_rod(t, r) for t in STATE} for r in (1, 2)}
    blk.step_fused(160, FP32_BULK)
    for r, st in frozen.items():
        for t in STATE:
            assert np.array_equal(blk.get_rod(t, r), st[t]), (r, t)
    # a budget that runs out in the launch that reaches the schedule's step:
    # the rod has no steps left after it, so it keeps its state as it is
    # without a schedule
    twin = build_block(wl)
    twin.set_fused_config(0, 0, 40)
    twin.set_step_budget(4, 80)
    twin.step_fused(80, FP32_BULK)
    assert np.array_equal(twin.get_rod("Director", 4), blk.get_rod("Director", 4))
    blk.close()
    twin.close()


def test_long_run_stays_orthonormal():
    wl = random_workload(8, 20, 300, seed=707, steps=0)
    blk = build_block(wl)
    blk.set_reortho_interval(1000)
    blk.step_fused_graph(10000, 100, FP32_BULK)
    eye = np.eye(3)
    for r in range(8):
        Q = _as_matrices(blk.get_rod("Director", r))
        assert np.abs(Q @ Q.transpose(0, 2, 1) - eye).max() <= 1e-13, r
    blk.close()


def test_interval_arguments():
    blk = build_block(random_workload(2, 10, 20, seed=709, steps=0))
    blk.set_reortho_interval(0)
    assert blk.reortho_interval() == (0, 0)
    with pytest.raises(ValueError):
        blk.set_reortho_interval(-1)
    blk.close()

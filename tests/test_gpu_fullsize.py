"""Full-size ragged batch (4096 rods, n log-uniform in [16, 1024]) on the GPU:
properties that do not need the CPU oracle at this size.

  * every path stays finite;
  * the fused FP64 path equals the op-by-op FP64 path (itself pinned to the
    oracle and the golden vectors) to <= 1e-10 relative;
  * the FP32-bulk path stays within 1e-6 of FP64 in position;
  * a sample of rods stepped alone reproduces their in-batch state bitwise.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64
from elasticapp_amd.workloads import Workload, build_block, ragged_4096
from tests.gpu_helpers import STATE

pytestmark = pytest.mark.gpu

STEPS = 200


@pytest.fixture(scope="module")
def ragged():
    return ragged_4096(steps=STEPS)


@pytest.fixture(scope="module")
def fused64(ragged):
    blk = build_block(ragged)
    blk.step_fused(STEPS, FP64)
    yield blk
    blk.close()


def test_ragged_layout(ragged):
    n = np.array(ragged.n_elems)
    assert len(n) == 4096 and n.min() >= 16 and n.max() <= 1024
    assert 0.9e6 < n.sum() < 1.1e6


def test_ragged_fused_fp64_equals_op_by_op(ragged, fused64):
    assert not fused64.check_finite().any()
    ops = build_block(ragged)
    ops.step(ragged.dt, STEPS)
    for t in STATE:
        a, b = fused64.get_field(t), ops.get_field(t)
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err <= 1e-10, (t, err)
    ops.close()


def _rod_position_errors(a, b, wl):
    """per rod: max over its nodes of |x_a - x_b|, over the rod's length"""
    xa, xb = a.get_field("Position"), b.get_field("Position")
    err = np.empty(len(wl.n_elems))
    for r, (s0, n) in enumerate(zip(b.starts, wl.n_elems)):
        L = wl.states[r]["ReferenceElementLength"].sum()
        err[r] = np.abs(xa[:, s0:s0 + n + 1] - xb[:, s0:s0 + n + 1]).max() / L
    return err


def test_ragged_fp32_bulk_close_to_fp64(ragged, fused64):
    blk = build_block(ragged)
    blk.step_fused_graph(STEPS, 100, FP32_BULK)
    assert not blk.check_finite().any()
    err = _rod_position_errors(blk, fused64, ragged)
    assert err.max() <= 1e-6, (int(err.argmax()), err.max())
    blk.close()


def test_ragged_rods_alone_equal_in_batch(ragged, fused64):
    for r in (0, 1234, 4095, int(np.argmax(ragged.n_elems))):
        one = Workload("one", [ragged.n_elems[r]], [ragged.states[r]], ragged.dt, STEPS, forces=[(0,) + tuple(ragged.forces[r][1:])])
        blk = build_block(one)
        blk.step_fused(STEPS, FP64)
        for t in STATE:
            assert np.array_equal(blk.get_rod(t, 0), fused64.get_rod(t, r)), (r, t)
        blk.close()


def test_fp32_bulk_error_after_ten_thousand_steps():
    """The accuracy target of the FP32-bulk path over the full horizon: after
    10,000 steps of the whole 4,096-rod ragged batch, each rod at its own
    stable dt, every rod's positions stay within 1e-6 of its length of the
    FP64 trajectory."""
    wl = ragged_4096(steps=10000)
    a, b = build_block(wl), build_block(wl)
    for blk in (a, b):
        blk.set_rod_dt(wl.info["rod_dt"])
    a.step_fused_graph(10000, 100, FP32_BULK)
    b.step_fused_graph(10000, 100, FP64)
    assert not a.check_finite().any() and not b.check_finite().any()
    err = _rod_position_errors(a, b, wl)
    print("fp32-bulk vs fp64 after 10,000 steps, per rod |dx|/L: max %.3e, median %.3e"
          % (err.max(), np.median(err)), flush=True)
    assert err.max() <= 1e-6, (int(err.argmax()), err.max())
    a.close()
    b.close()

"""Rods run as pieces on the fused path (erod_set_fused_pieces): rods longer
than the tile, or than a requested piece size, are cut into pieces that own
a part of the rod and recompute a halo of their neighbours' elements, a
launch over pieces taking at most halo / 2 steps.

  * rods of 1,025 to 4,096 elements (longer than any tile): fused FP64
    equal to the oracle built with the device library's transcendental
    functions, bitwise, and FP32-bulk positions within 1e-6 of the rod length;
  * the cut does not change results: a batch stepped whole and stepped with
    its longer rods in pieces is bitwise the same (both precisions, equal
    steps per launch), for several piece sizes and halos;
  * clocks, step budgets, a clamp at the rod start and a ramped end force on
    cut rods, and a captured graph over pieces, against the uncut run.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64, RAMP_SIN
from elasticapp_amd.workloads import build_block
from tests.gpu_helpers import STATE, oracle_system, random_workload

pytestmark = pytest.mark.gpu


def _equal(a, b, rods, tags=STATE):
    for r in rods:
        for t in tags:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)


def test_long_rods_match_the_oracle():
    wl = random_workload(0, 0, 0, seed=101, steps=120, lengths=[1025, 1500, 2048, 3001, 4096, 64, 700])
    sysm = oracle_system(wl, variant="devmath")
    sysm.step(wl.dt, wl.steps)
    f64 = build_block(wl)
    f64.step_fused(wl.steps, FP64)
    lanes, epl, halo, tiles, pieces = f64.fused_pieces()
    assert pieces >= 5 * 2 and tiles > 0 and halo == 16
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(f64.get_rod(t, r), sysm.rod_field(r, t)), (r, t)
    assert np.array_equal(f64.rod_time(), np.full(len(wl.n_elems), f64.rod_time()[0]))
    f32 = build_block(wl)
    f32.step_fused(wl.steps, FP32_BULK)
    for r in range(len(wl.n_elems)):
        L = float(wl.states[r]["ReferenceElementLength"].sum())
        x, c = f32.get_rod("Position", r), sysm.rod_field(r, "Position")
        assert np.abs(x - c).max() <= 1e-6 * L, r


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
@pytest.mark.parametrize("lanes,halo", [(256, 16), (128, 8), (512, 32)])
def test_cut_does_not_change_results(precision, lanes, halo):
    wl = random_workload(30, 2, 1024, seed=103, steps=0)
    spl = halo // 2
    whole, cut = build_block(wl), build_block(wl)
    whole.set_fused_config(1024, 2, spl)
    cut.set_fused_config(1024, 2, spl)
    cut.set_fused_pieces(lanes, 0, halo)
    for blk in (whole, cut):
        blk.step_fused(5 * spl + 3, precision)
    assert cut.fused_pieces()[4] > 0 and whole.fused_pieces()[4] == 0
    _equal(whole, cut, range(len(wl.n_elems)))
    assert np.array_equal(whole.rod_time(), cut.rod_time())


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_budgets_clocks_and_hooks_on_cut_rods(precision):
    """Step budgets that end inside a launch, a clamp at the start and a
    sin-ramped end force on rods that are cut, per-rod dt: equal to the same
    rods stepped whole with the same steps per launch."""
    wl = random_workload(8, 300, 900, seed=107, steps=0)
    for r in range(3):
        wl.clamps.append((r, wl.states[r]["Position"][:, 0].copy(), wl.states[r]["Director"][:, 0].copy()))
    wl.forces[:] = [(r, np.array([0.3, -0.2, 0.1]) * (r + 1), RAMP_SIN, 0.0, 40 * wl.dt) for r in range(8)]
    dts = wl.dt * (1.0 + 0.05 * np.arange(8))
    budgets = [None, 13, 0, 40, 7, None, 29, 100]
    blocks = []
    for pieces in (False, True):
        b = build_block(wl)
        b.set_rod_dt(dts)
        b.set_fused_config(1024, 2, 8)
        if pieces:
            b.set_fused_pieces(256, 0, 16)
        for r, n in enumerate(budgets):
            b.set_step_budget(r, n)
        b.step_fused(45, precision)
        blocks.append(b)
    whole, cut = blocks
    # the rods are cut in this plan (a spent budget empties the plan: a twin without budgets shows it)
    probe = build_block(wl)
    probe.set_fused_config(1024, 2, 8)
    probe.set_fused_pieces(256, 0, 16)
    probe.step_fused(1, precision)
    assert probe.fused_pieces()[4] >= 8 * 2
    _equal(whole, cut, range(8))
    assert np.array_equal(whole.rod_time(), cut.rod_time())
    assert [cut.step_budget(r) for r in range(8)] == [whole.step_budget(r) for r in range(8)]
    assert [cut.step_budget(r) for r in range(8)] == [None, 0, 0, 0, 0, None, 0, 55]
    # the clocks advanced by each rod's own steps
    t = cut.rod_time()
    for r, n in enumerate(budgets):
        k = 45 if n is None else min(n, 45)
        assert abs(t[r] - k * dts[r]) <= 1e-12 * max(1.0, k * dts[r]), r


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_cut_does_not_change_results_on_the_ragged_batch(precision):
    """The 4,096-rod ragged batch (about 4,000 tiles, run in several waves of
    workgroups): with its rods of more than 256 elements in pieces, 25 steps
    equal the uncut run bitwise (each launch over pieces reads the state the
    previous one wrote and writes the other copy, so no piece reads a
    neighbour's halo while that neighbour writes it)."""
    from elasticapp_amd.workloads import ragged_4096
    wl = ragged_4096()
    dts = np.asarray(wl.info["rod_dt"])
    blocks = []
    for cut in (False, True):
        b = build_block(wl)
        b.set_rod_dt(dts)
        if cut:
            b.set_fused_config(256, 1, 8)
            b.set_fused_split(128, 1)
        else:
            b.set_fused_config(1024, 2, 8)
        b.step_fused(25, precision)
        blocks.append(b)
    whole, cut = blocks
    assert cut.fused_pieces()[4] > 1000 and not cut.check_finite().any()
    x0, x1 = whole.get_field("Position"), cut.get_field("Position")
    q0, q1 = whole.get_field("Director"), cut.get_field("Director")
    assert np.array_equal(x0, x1) and np.array_equal(q0, q1)
    assert np.array_equal(whole.get_field("AngularVelocity"), cut.get_field("AngularVelocity"))
    for b in blocks:
        b.close()


def test_graph_over_pieces_equals_direct():
    wl = random_workload(6, 1100, 2500, seed=109, steps=0)
    a, b = build_block(wl), build_block(wl)
    for blk in (a, b):
        blk.set_fused_config(0, 0, 8)
    a.step_fused(40, FP32_BULK)
    b.step_fused_graph(40, 16, FP32_BULK)
    _equal(a, b, range(6))
    assert np.array_equal(a.rod_time(), b.rod_time())


def test_piece_arguments():
    from elasticapp_amd import ErodError
    blk = build_block(random_workload(3, 10, 40, seed=111, steps=0))
    with pytest.raises(ErodError):
        blk.set_fused_pieces(0, 0, 1)
    blk.set_fused_pieces(64, 0, 40)      # 2 * halo + 2 > 64 lanes
    blk.set_fused_config(64, 1, 0)
    with pytest.raises(ErodError, match="halo"):
        blk.step_fused(1, FP64)
    blk.close()

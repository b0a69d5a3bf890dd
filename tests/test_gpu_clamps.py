"""Clamps at any index, up to two per rod (erod_add_clamp).

A clamp at index i holds node i and element min(i, n - 1): position and
director after the kinematic half step, velocity and angular velocity zeroed
after the dynamic step, as the reference's per-index constrain_primitives /
constrain_rates (elastica/Simulator/Constraints/Constraint.hpp:290-313,
elastica/Simulator/Interface/Constraints/ExecutionPolicy.hpp:591-593).

  * ragged batch with clamps at the first node, the last node, interior
    nodes and pairs of them, tip forces and gravity: fused FP64 equals the
    oracle built with the device library's sin / cos / acos bitwise, the
    op-by-op path is within 1e-12 of it, FP32-bulk positions within 1e-6 of
    the rod length;
  * rods cut into pieces with clamps inside pieces and in their halos equal
    the uncut run bitwise;
  * the clamped node and element hold their pose exactly;
  * a beam clamped at both ends under a uniform load against the shearable
    closed form w(s) = q s^2 (L-s)^2 / (24 EI) + q s (L-s) / (2 a G A),
    a = 27/28, first order in the element length;
  * a checkpoint keeps both clamps of a rod (a resumed run equals the
    uninterrupted one);
  * refused arguments.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64, DeviceBlock, ErodError, straight_rod_state
from elasticapp_amd.workloads import Workload, build_block
from tests.gpu_helpers import STATE, oracle_system, random_workload

pytestmark = pytest.mark.gpu


def _clamped_workload(seed=801, steps=300):
    wl = random_workload(12, 6, 200, seed=seed, steps=steps)
    rng = np.random.default_rng(seed + 1)
    for r in range(12):
        st, n = wl.states[r], wl.n_elems[r]
        kind = r % 6
        idx = {0: [0], 1: [n], 2: [n // 2], 3: [0, n], 4: [n // 3, (2 * n) // 3 + 1], 5: []}[kind]
        for i in idx:
            e = min(i, n - 1)
            # a pose off the rod's own, so that holding it is visible
            x = st["Position"][:, i] + 1e-3 * rng.normal(size=3)
            wl.clamps.append((r, x, st["Director"][:, e].copy(), i))
    wl.gravity = (0.0, -0.5, 0.0)
    return wl


def test_clamps_match_the_oracle():
    wl = _clamped_workload()
    sysm = oracle_system(wl, variant="devmath")
    sysm.step(wl.dt, wl.steps)
    f64 = build_block(wl)
    f64.step_fused(wl.steps, FP64)
    ops = build_block(wl)
    ops.step(wl.dt, wl.steps)
    f32 = build_block(wl)
    f32.step_fused(wl.steps, FP32_BULK)
    for r in range(12):
        L = float(wl.states[r]["ReferenceElementLength"].sum())
        for t in STATE:
            c = sysm.rod_field(r, t)
            assert np.array_equal(f64.get_rod(t, r), c), (r, t)
            o = ops.get_rod(t, r)
            assert np.abs(o - c).max() <= 1e-12 * max(np.abs(c).max(), 1e-300), (r, t)
        x32, c = f32.get_rod("Position", r), sysm.rod_field(r, "Position")
        assert np.abs(x32 - c).max() <= 1e-6 * L, r
    for blk in (f64, ops, f32):
        blk.close()
    sysm.close()


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_clamped_pose_is_held(precision):
    wl = _clamped_workload(seed=803, steps=0)
    blk = build_block(wl)
    blk.step_fused(120, precision)
    for rod, x, Q, i in wl.clamps:
        n = wl.n_elems[rod]
        e = min(i, n - 1)
        assert np.array_equal(blk.get_rod("Position", rod)[:, i], x), (rod, i)
        assert np.array_equal(blk.get_rod("Director", rod)[:, e], Q), (rod, i)
        assert not blk.get_rod("Velocity", rod)[:, i].any() and not blk.get_rod("AngularVelocity", rod)[:, e].any()
    blk.close()


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_clamps_in_pieces_and_halos(precision):
    """Rods of 600-1,000 elements in pieces of 256 lanes with a halo of 16:
    clamps inside pieces, at piece boundaries (in a neighbour's halo) and at
    both ends, against the same rods stepped whole (one 1024-lane tile each)."""
    wl = random_workload(4, 600, 1000, seed=805, steps=0)
    for r in range(4):
        st, n = wl.states[r], wl.n_elems[r]
        for i in ([0, n], [240, 250], [500, n], [7, 590])[r]:
            wl.clamps.append((r, st["Position"][:, i].copy(), st["Director"][:, min(i, n - 1)].copy(), i))
    blocks = []
    for cut in (False, True):
        b = build_block(wl)
        b.set_fused_config(0, 0, 8)
        if cut:
            b.set_fused_pieces(256, 0, 16)
        b.step_fused(40, precision)
        blocks.append(b)
    whole, cut = blocks
    assert cut.fused_pieces()[4] > 4 and whole.fused_pieces()[4] == 0
    for r in range(4):
        for t in STATE:
            assert np.array_equal(whole.get_rod(t, r), cut.get_rod(t, r)), (r, t)
    for b in blocks:
        b.close()


def _beam_error(n, precision):
    L, r, rho, E, G, gx = 3.0, 0.25, 5e3, 1e6, 1e4, -0.01
    dt = 0.01 * L / n
    st = straight_rod_state(n, r, L, [0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0], rho, E, G, 0.5, 0.5)
    clamps = [(0, st["Position"][:, 0].copy(), st["Director"][:, 0].copy(), 0),
              (0, st["Position"][:, n].copy(), st["Director"][:, n - 1].copy(), n)]
    wl = Workload("clamped-clamped", [n], [st], dt, int(200.0 / dt), clamps=clamps, gravity=(gx, 0.0, 0.0))
    blk = build_block(wl)
    blk.step_fused_graph(wl.steps, 100, precision)
    x = blk.get_rod("Position", 0)
    blk.close()
    A = np.pi * r * r
    I = A * A / (4 * np.pi)
    q = rho * A * abs(gx)
    s = x[2]
    want = -(q * s ** 2 * (L - s) ** 2 / (24 * E * I) + q * s * (L - s) / (2 * (27.0 / 28.0) * G * A))
    return np.abs(x[0] - want).max(), np.abs(want).max()


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_beam_clamped_at_both_ends(precision):
    e10, wmax = _beam_error(10, precision)
    e20, _ = _beam_error(20, precision)
    print("clamped-clamped beam: relative error %.3e (n=10), %.3e (n=20)" % (e10 / wmax, e20 / wmax), flush=True)
    assert e20 <= 0.025 * wmax, (e20, wmax)
    assert 1.6 <= e10 / e20 <= 2.4, (e10, e20)


def test_checkpoint_keeps_both_clamps(tmp_path):
    wl = _clamped_workload(seed=807, steps=0)
    a = build_block(wl)
    a.step_fused(50, FP64)
    path = str(tmp_path / "c.ckpt")
    a.save(path)
    b = DeviceBlock.load(path)
    a.step_fused(70, FP64)
    b.step_fused(70, FP64)
    for r in range(12):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)
    a.close()
    b.close()


def test_clamp_arguments():
    wl = random_workload(2, 10, 10, seed=809, steps=0)
    blk = build_block(wl)
    x, Q = np.zeros(3), np.eye(3).ravel()
    with pytest.raises(ErodError, match="index past"):
        blk.add_clamp(0, x, Q, 11)
    blk.add_clamp(0, x, Q, 3)
    blk.add_clamp(0, x, Q, 3)            # the same index: replaced
    blk.add_clamp(0, x, Q, 10)
    with pytest.raises(ErodError, match="at most two"):
        blk.add_clamp(0, x, Q, 5)
    blk.add_clamp(1, x, Q, 9)
    with pytest.raises(ErodError, match="last element"):
        blk.add_clamp(1, x, Q, 10)       # n - 1 and n would both hold element n - 1
    blk.close()

"""Rods resting on a plane under gravity with friction (the 100,000 x 128
contact batch and a 48-rod cut of it) on the GPU.

  * the 48-rod batch against the oracle: both FP64 paths <= 1e-10, FP32-bulk
    positions <= 1e-6 relative.  The contact law is non-smooth (max and min
    switch per node and step), which amplifies the last-bit differences of
    the device acos / sin / cos over the steps; the smooth batches hold the
    op-by-op path to 1e-12.
  * the full batch: every path stays finite, no node sinks deeper into the
    plane than a bound well above the static deflection m g / k, and rods
    stepped alone reproduce their in-batch state bitwise.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64
from elasticapp_amd.workloads import Workload, build_block, rods_on_plane
from tests.gpu_helpers import STATE, oracle_system, rel_state_error

pytestmark = pytest.mark.gpu


def test_contact_cut_matches_oracle():
    wl = rods_on_plane(n_rods=48, n=128, steps=300)
    sysm = oracle_system(wl)
    sysm.step(wl.dt, wl.steps)
    a = build_block(wl)
    a.step(wl.dt, wl.steps)
    b = build_block(wl)
    b.step_fused(wl.steps, FP64)
    c = build_block(wl)
    c.step_fused(wl.steps, FP32_BULK)
    e_ops, e_f64, e_f32 = rel_state_error(a, sysm), rel_state_error(b, sysm), rel_state_error(c, sysm)
    print("op-by-op", e_ops, "fused fp64", e_f64, "fp32-bulk", e_f32)
    assert max(e_ops.values()) <= 1e-10, e_ops
    assert max(e_f64.values()) <= 1e-10, e_f64
    assert e_f32["Position"] <= 1e-6, e_f32
    # the rods came down onto the plane and are held by it
    z = np.stack([a.get_rod("Position", r)[2] for r in range(48)])
    assert z.min() < 0.01 and z.min() > 0.01 - 1e-3


def test_tilted_plane_matches_oracle():
    """The 48-rod cut turned by a rotation, so that it rests on a tilted
    plane with gravity along the turned axis: every path against the oracle
    on the same turned batch, and the turned result against the turned
    result of the level batch."""
    from scipy.spatial.transform import Rotation
    from tests.gpu_helpers import rotated
    base = rods_on_plane(n_rods=48, n=128, steps=300)
    R = Rotation.from_rotvec([0.3, -0.5, 0.2]).as_matrix()
    wl = rotated(base, R)
    sysm = oracle_system(wl)
    sysm.step(wl.dt, wl.steps)
    a = build_block(wl)
    a.step(wl.dt, wl.steps)
    b = build_block(wl)
    b.step_fused(wl.steps, FP64)
    c = build_block(wl)
    c.step_fused(wl.steps, FP32_BULK)
    e_ops, e_f64, e_f32 = rel_state_error(a, sysm), rel_state_error(b, sysm), rel_state_error(c, sysm)
    print("op-by-op", e_ops, "fused fp64", e_f64, "fp32-bulk", e_f32)
    assert max(e_ops.values()) <= 1e-10, e_ops
    assert max(e_f64.values()) <= 1e-10, e_f64
    assert e_f32["Position"] <= 1e-6, e_f32
    level = build_block(base)
    level.step_fused(base.steps, FP64)
    for r in range(0, 48, 7):
        x, y = R @ level.get_rod("Position", r), b.get_rod("Position", r)
        assert np.abs(x - y).max() <= 1e-9, r
    # the rods are held by the tilted plane
    n = R @ np.array([0.0, 0.0, 1.0])
    h = np.concatenate([n @ a.get_rod("Position", r) for r in range(48)])
    assert h.min() < 0.01 and h.min() > 0.01 - 1e-3


STEPS = 200


@pytest.fixture(scope="module")
def plane_batch():
    return rods_on_plane(steps=STEPS)


def test_full_contact_batch(plane_batch):
    blk = build_block(plane_batch)
    blk.step_fused_graph(STEPS, 100, FP32_BULK)
    assert not blk.check_finite().any()
    x = blk.get_field("Position")
    # static deflection m g / k = (rho A L / n) g / k ~ 2.4e-5 m; allow 40x for the dynamics
    r = plane_batch.info["radius"]
    m = plane_batch.info["density"] * np.pi * r * r * plane_batch.info["length"] / 128
    sink = r - x[2, np.concatenate([np.arange(s, s + 129) for s in blk.starts[:2000]])].min()
    assert sink < 40 * m * 9.81 / 1e3, sink
    for rod in (0, 77777, 99999):
        one = Workload("one", [128], [plane_batch.states[rod]], plane_batch.dt, STEPS, gravity=plane_batch.gravity, contact=plane_batch.contact)
        b1 = build_block(one)
        b1.step_fused_graph(STEPS, 100, FP32_BULK)
        for t in STATE:
            assert np.array_equal(b1.get_rod(t, 0), blk.get_rod(t, rod)), (rod, t)
        b1.close()
    blk.close()

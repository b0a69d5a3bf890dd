"""The device paths against closed forms, with no oracle in between.

* Axial stretching of a clamped rod pulled at its tip: steady elongation
  F L / (E A - F) (examples/axial_stretching/convergence.py:62-74).
* The shearable Timoshenko cantilever under a sin-ramped tip load:
  y(s) = s(-F/(aGA) + s(-FL/(2EI) + sF/(6EI))) with a = 27/28
  (examples/timoshenko/helpers/analytical.py:4-31), run to steady state with
  the reference's convergence settings (examples/timoshenko/convergence.py:25-60);
  the discretisation error is first order in the element length.
* Director re-orthonormalisation: from a director perturbed by eps it lands
  on the polar factor (U V^T of the SVD) to O(eps^2), and leaves an
  orthonormal director unchanged to rounding.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64, RAMP_LINEAR, RAMP_SIN, straight_rod_state
from elasticapp_amd.workloads import Workload, build_block
from tests.gpu_helpers import random_workload

pytestmark = pytest.mark.gpu


def _clamped(n, r, L, direction, normal, rho, E, G, damping, F, ramp, interval, steps):
    st = straight_rod_state(n, r, L, [0.0, 0.0, 0.0], direction, normal, rho, E, G, damping, damping)
    clamps = [(0, st["Position"][:, 0].copy(), st["Director"][:, 0].copy())]
    forces = [(0, np.array(F, dtype=np.float64), ramp, 0.0, interval)]
    return Workload("clamped", [n], [st], 0.01 * L / n, steps, clamps=clamps, forces=forces)


def _axial(steps=None):
    n, L = 10, 1.0
    dt = 0.01 * L / n
    return _clamped(n, 0.025, L, [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], 1e3, 1e4, 1e4 / 1.5, 10.0,
                    [1.0, 0.0, 0.0], RAMP_LINEAR, 2.0, steps or int(20.0 / dt))


def _axial_want():
    A = np.pi * 0.025 ** 2
    return 1.0 * 1.0 / (A * 1e4 - 1.0)


@pytest.mark.parametrize("precision,tol", [(FP64, 1e-6), (FP32_BULK, 1e-5)])
def test_axial_stretching_fused(precision, tol):
    """FP64 to the oracle test's 1e-6 of the elongation; FP32-bulk to 1e-5
    (its strain e - 1 ~ 0.054 is formed from FP32 lengths, about 1e-6 relative)."""
    wl = _axial()
    blk = build_block(wl)
    blk.step_fused(wl.steps, precision)
    tip = blk.get_rod("Position", 0)[0, -1] - 1.0
    want = _axial_want()
    assert abs(tip - want) <= tol * want, (tip, want)
    blk.close()


def test_axial_stretching_op_by_op():
    wl = _axial()
    blk = build_block(wl)
    blk.step(wl.dt, wl.steps)
    tip = blk.get_rod("Position", 0)[0, -1] - 1.0
    want = _axial_want()
    assert abs(tip - want) <= 1e-6 * want, (tip, want)
    blk.close()


def _timoshenko_error(n, precision):
    L, r, rho, E, G, F = 3.0, 0.25, 5e3, 1e6, 1e4, -15.0
    dt = 0.01 * L / n
    wl = _clamped(n, r, L, [0.0, 0.0, 1.0], [0.0, 1.0, 0.0], rho, E, G, 0.5, [F, 0.0, 0.0], RAMP_SIN, 100.0,
                  int(200.0 / dt))
    blk = build_block(wl)
    blk.step_fused(wl.steps, precision)
    x = blk.get_rod("Position", 0)
    blk.close()
    s = x[2]
    A = np.pi * r * r
    I = A * A / (4 * np.pi)
    f = abs(F)
    want = s * (-f / ((27.0 / 28.0) * G * A) + s * (-f * L / 2 / (E * I) + s * f / 6 / (E * I)))
    return np.abs(x[0] - want).max(), np.abs(want).max()


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_timoshenko_cantilever_converges(precision):
    e10, ymax = _timoshenko_error(10, precision)
    e20, _ = _timoshenko_error(20, precision)
    assert e20 <= 0.06 * ymax, (e20, ymax)
    assert 1.8 <= e10 / e20 <= 2.2, (e10, e20)


def _as_matrices(d):
    # director lines 3 p + r hold Q(p, r) of every element
    return d.T.reshape(-1, 3, 3)


def test_reorthonormalize_is_the_polar_factor():
    """One Newton-Schulz step X (3I - X^T X) / 2 acts on the SVD X = U S V^T as
    S -> S (3 - S^2) / 2 with U V^T unchanged: checked against numpy's SVD, and
    a second call brings a director perturbed by 1e-6 to orthonormal at rounding.
    (Using X X^T in place of X^T X moves the singular values by 5e-6 here.)"""
    wl = random_workload(5, 8, 40, seed=61, steps=0)
    blk = build_block(wl)
    rng = np.random.default_rng(62)
    eps = 1e-6
    before = []
    for r in range(5):
        d = blk.get_rod("Director", r)
        if r % 2 == 0:
            d = d + eps * rng.normal(size=d.shape)
            blk.set_rod("Director", r, d)
        before.append(d)
    blk.reorthonormalize()
    once = [blk.get_rod("Director", r) for r in range(5)]
    blk.reorthonormalize()
    eye = np.eye(3)
    for r in range(5):
        Q0, Q1, Q2 = (_as_matrices(a) for a in (before[r], once[r], blk.get_rod("Director", r)))
        u0, s0, v0 = np.linalg.svd(Q0)
        u1, s1, v1 = np.linalg.svd(Q1)
        # the map is not monotonic around 1, so compare the sorted values
        assert np.abs(np.sort(s1, 1) - np.sort(s0 * (3.0 - s0 * s0) / 2.0, 1)).max() <= 1e-14, r
        assert np.abs(u1 @ v1 - u0 @ v0).max() <= 1e-13, r
        assert np.abs(Q2 @ Q2.transpose(0, 2, 1) - eye).max() <= 1e-14, r
        if r % 2 == 0:
            assert np.abs(Q0 @ Q0.transpose(0, 2, 1) - eye).max() >= 1e-6, r    # it was perturbed
        else:
            assert np.abs(Q1 - Q0).max() <= 1e-15, r
    blk.close()


def test_periodic_reorthonormalisation_in_a_long_run():
    """FP32-bulk with a re-orthonormalisation every 1000 steps: directors stay
    orthonormal to 1e-13, and positions stay within 1e-7 of the rod length of
    the same run without it (the directors differ at 1e-14; through the FP32
    rounding of the loads that grows to FP32-bulk's own rounding noise, below
    its 1e-6 accuracy bound)."""
    wl = random_workload(6, 20, 60, seed=63, steps=0)
    a = build_block(wl)
    b = build_block(wl)
    for _ in range(5):
        a.step_fused(1000, FP32_BULK)
        b.step_fused(1000, FP32_BULK)
        b.reorthonormalize()
    eye = np.eye(3)
    for r in range(6):
        Q = _as_matrices(b.get_rod("Director", r))
        assert np.abs(Q @ Q.transpose(0, 2, 1) - eye).max() <= 1e-13, r
        xa, xb = a.get_rod("Position", r), b.get_rod("Position", r)
        L = np.linalg.norm(np.diff(xa, axis=1), axis=0).sum()
        assert np.abs(xa - xb).max() <= 1e-7 * L, r
    a.close()
    b.close()

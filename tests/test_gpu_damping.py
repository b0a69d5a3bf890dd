"""Damping rates that vary along a rod (the reference stores ForceDampingRate
per node and TorqueDampingRate per element,
ExplicitDampingAdapterFacade.hpp:185-198): the fused paths read them per
node and element for such rods and keep one rate per rod for the others.
Fused FP64 against the oracle built with the device library's
transcendental functions, bitwise; FP32-bulk positions within 1e-6 of the
rod length; op-by-op bitwise too."""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64
from elasticapp_amd.workloads import build_block
from tests.gpu_helpers import STATE, oracle_system, random_workload

pytestmark = pytest.mark.gpu


def test_rates_along_the_rod():
    wl = random_workload(12, 2, 700, seed=121, steps=150)
    rng = np.random.default_rng(122)
    for r, st in enumerate(wl.states):
        if r % 3 != 2:  # two rods in three get varying rates, the third keeps one rate
            n = wl.n_elems[r]
            st["ForceDampingRate"] = rng.uniform(0.0, 0.8, (1, n + 1))
            st["TorqueDampingRate"] = rng.uniform(0.0, 0.8, (1, n))
    sysm = oracle_system(wl, variant="devmath")
    sysm.step(wl.dt, wl.steps)
    ops = build_block(wl)
    ops.step(wl.dt, wl.steps)
    f64 = build_block(wl)
    f64.step_fused(wl.steps, FP64)
    f32 = build_block(wl)
    f32.step_fused(wl.steps, FP32_BULK)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(ops.get_rod(t, r), sysm.rod_field(r, t)), (r, t)
            assert np.array_equal(f64.get_rod(t, r), sysm.rod_field(r, t)), (r, t)
        L = float(wl.states[r]["ReferenceElementLength"].sum())
        assert np.abs(f32.get_rod("Position", r) - sysm.rod_field(r, "Position")).max() <= 1e-6 * L, r
    # the rates change after the first steps: the next fused call reads the new ones
    st = wl.states[0]
    new = np.linspace(0.1, 0.9, wl.n_elems[0] + 1)[None, :]
    f64.set_rod("ForceDampingRate", 0, new)
    sysm.rod_field(0, "ForceDampingRate")[...] = new
    f64.step_fused(20, FP64)
    sysm.step(wl.dt, 20)
    for t in STATE:
        assert np.array_equal(f64.get_rod(t, 0), sysm.rod_field(0, t)), t

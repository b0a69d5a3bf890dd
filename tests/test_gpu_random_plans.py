"""Random batches, hooks and tile plans against the oracle.

Each case draws a batch whose rod lengths include the tile-shape edges
(2, 3, P*NT - 1, P*NT, P*NT + 1 and rods cut into several pieces), random
hooks (clamps at any index, one or two per rod, end forces with and without
ramps, damping rates along some rods, gravity, sometimes a contact plane,
level or of any normal) and
a random fused plan (main tile shape, second group, piece shape and halo,
steps per launch).  Then:

  * fused FP64 equals the oracle built with the device library's sin / cos /
    acos bit for bit, whatever the plan (rods never interact and the
    arithmetic does not depend on the tile);
  * FP32-bulk under that plan equals FP32-bulk in 1024-lane tiles with the
    same halo and steps per launch bit for bit (at most halo / 2 steps per launch,
    so that cut and whole rods take the same launches), and its positions
    stay within 1e-6 of the rod length of the oracle's.
"""
import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64, RAMP_LINEAR, RAMP_SIN
from elasticapp_amd.workloads import build_block
from tests.gpu_helpers import STATE, oracle_system, random_workload

pytestmark = pytest.mark.gpu

# (lanes, elements per thread) of the kernel's tile shapes
SHAPES = [(64, 1), (128, 1), (256, 1), (512, 1), (512, 2), (1024, 2), (512, 4), (1024, 4)]


def _case(seed):
    rng = np.random.default_rng(seed)
    main = SHAPES[int(rng.integers(0, len(SHAPES)))]
    cap, p = main
    edges = [2, 3, cap - 1, cap, cap + 1, 2 * cap + 7, int(rng.integers(4, 3 * cap))]
    lengths = [int(x) for x in rng.choice(edges, int(rng.integers(6, 12)))]
    lengths += [int(x) for x in rng.integers(2, 300, int(rng.integers(4, 16)))]
    # a shape with P elements per thread runs only rods of at least P elements
    # whole; shorter ones leave the shape infeasible, so lengthen them
    lengths = [max(n, p) for n in lengths]
    steps = int(rng.integers(40, 90))
    wl = random_workload(0, 0, 0, seed=seed + 1, steps=steps, lengths=lengths)
    wl.forces = []
    for r, n in enumerate(wl.n_elems):
        st = wl.states[r]
        kind = int(rng.integers(0, 5))
        if kind == 1:
            idx = [int(rng.integers(0, n + 1))]
        elif kind == 2:
            # two distinct clamps, never n - 1 and n together (they would hold
            # the same last element twice)
            a, b = (int(x) for x in rng.choice(n + 1, 2, replace=False))
            if {a, b} == {n - 1, n}:
                b = 0 if a != 0 else 1
            idx = [a, b] if a != b else [a]
        else:
            idx = []
        for i in idx:
            e = min(i, n - 1)
            wl.clamps.append((r, st["Position"][:, i] + 1e-4 * rng.normal(size=3), st["Director"][:, e].copy(), i))
        if rng.random() < 0.6:
            ramp = int(rng.choice([0, RAMP_LINEAR, RAMP_SIN]))
            wl.forces.append((r, rng.normal(0.0, 0.1, 3), ramp, float(rng.uniform(0, 0.5)) * steps * wl.dt,
                              float(rng.uniform(0.2, 1.0)) * steps * wl.dt))
        if rng.random() < 0.3:
            st["ForceDampingRate"] = rng.uniform(0.0, 0.8, (1, n + 1))
            st["TorqueDampingRate"] = rng.uniform(0.0, 0.8, (1, n))
    if rng.random() < 0.5:
        wl.gravity = tuple(rng.normal(0.0, 0.5, 3))
    if rng.random() < 0.4:
        # a level plane, or one of any normal (the kernels' general law); the
        # lowest nodes just inside it
        nrm = np.array([0.0, 0.0, 1.0])
        if rng.random() < 0.5:
            nrm = rng.normal(size=3)
            nrm /= np.linalg.norm(nrm)
        lowest = min(float((nrm @ st["Position"]).min()) for st in wl.states)
        wl.contact = (1e3, 10.0, 0.3, 10.0, lowest + 1e-5, 0.0) + ((tuple(nrm),) if nrm[2] != 1.0 else ())
    # the plan: a second group of smaller tiles, pieces of a shape and halo
    # that fit, steps per launch
    smaller = [s for s in SHAPES if s[0] < cap and s[1] <= p]
    split = smaller[int(rng.integers(0, len(smaller)))] if smaller and rng.random() < 0.5 else None
    piece_shapes = [s for s in SHAPES if s[1] <= p and s[0] >= 64]
    pieces = piece_shapes[int(rng.integers(0, len(piece_shapes)))]
    halo = int(rng.choice([2, 4, 8, 16, 32]))
    while 2 * halo + 2 > pieces[0]:
        halo //= 2
    halo = max(halo, 2)
    # at most halo / 2 steps per launch, so that pieces take the launches
    # whole rods take (FP32-bulk's merged half steps then fall alike)
    spl = int(rng.integers(1, halo // 2 + 1))
    return wl, (cap, p, spl), split, (pieces[0], pieces[1], halo)


@pytest.mark.parametrize("seed", range(16))
def test_random_plan_matches_the_oracle(seed):
    wl, config, split, pieces = _case(1000 + 17 * seed)
    print("case", seed, "rods", len(wl.n_elems), "lengths", sorted(wl.n_elems)[-6:], "config", config,
          "split", split, "pieces", pieces, "clamps", len(wl.clamps), "steps", wl.steps, flush=True)
    sysm = oracle_system(wl, variant="devmath")
    sysm.step(wl.dt, wl.steps)

    def planned(precision):
        blk = build_block(wl)
        blk.set_fused_config(*config)
        if split:
            blk.set_fused_split(*split)
        blk.set_fused_pieces(*pieces)
        blk.step_fused(wl.steps, precision)
        return blk

    f64 = planned(FP64)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(f64.get_rod(t, r), sysm.rod_field(r, t)), (seed, r, wl.n_elems[r], t)
    f32 = planned(FP32_BULK)
    ref = build_block(wl)
    ref.set_fused_config(1024, 2, config[2])
    ref.set_fused_pieces(0, 0, pieces[2])
    ref.step_fused(wl.steps, FP32_BULK)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(f32.get_rod(t, r), ref.get_rod(t, r)), (seed, r, wl.n_elems[r], t)
        L = float(wl.states[r]["ReferenceElementLength"].sum())
        dx = np.abs(f32.get_rod("Position", r) - sysm.rod_field(r, "Position")).max()
        assert dx <= 1e-6 * L, (seed, r, dx / L)
    for blk in (f64, f32, ref):
        blk.close()

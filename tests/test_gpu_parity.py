"""GPU parity: every path of liberod.so against the reference's golden vectors
and the CPU oracle.

Tolerances
  * op-by-op FP64 path: the reference's golden tolerances (4e-13 .. 4e-12);
    against the oracle <= 1e-12 relative (same operation order, no FMA; only
    the device acos/sin/cos differ from glibc by an ulp).
  * fused FP64 path: <= 1e-10 relative after hundreds of steps (it is built
    without FMA contraction and keeps the reference's order of evaluation, so
    on these batches it equals the op-by-op path bitwise; the bound leaves
    room for the device math library).
  * FP32-bulk path: position error <= 1e-6 relative to the position scale.
  * both fused paths: bitwise independent of the tile shape (lanes per tile,
    elements per thread) and of the other rods in the batch; the FP64 path
    also of the steps per launch.
"""
import numpy as np
import pytest

from elasticapp_amd import FIELDS, FP32_BULK, FP64, DeviceBlock, ErodError, straight_rod_state
from elasticapp_amd.workloads import Workload
from tests.golden_fixtures import as_field, load, rel_err_ok
from tests.gpu_helpers import STATE, device_block, oracle_system, random_workload, rel_state_error

pytestmark = pytest.mark.gpu


# --- golden vectors ---------------------------------------------------------------

def _golden_block(case):
    p = case["init"]
    n = p["n_elem"]
    blk = DeviceBlock([n] * case["n_rods"], damped=case["plugin"] == "CosseratRod")
    st = straight_rod_state(n, p["radius"], p["length"], p["origin"], p["direction"], p["normal"], p["density"],
                            p["youngs_modulus"], p["shear_modulus"], p["force_damping"] or 0.0,
                            p["torque_damping"] or 0.0)
    for r in range(case["n_rods"]):
        for tag, a in st.items():
            blk.set_rod(tag, r, a)
    for a in case["assign"]:
        entry = case["replaced"][str(a["rod"])][a["tag"]]
        blk.set_rod(a["tag"], a["rod"], as_field(entry, a["tag"], a["shrink_diagonal"]))
    return blk


def _golden_cases():
    out = []
    for f in ("fullrod_PerturbedRodStep.json", "fullrod_PerturbedRodStepWithoutDamping.json"):
        for case in load(f)["cases"]:
            for sec in case["sections"]:
                out.append(pytest.param(case, sec, id="%s-%s" % (case["name"].split(".")[-2], sec["name"])))
    return out


@pytest.mark.parametrize("case,section", _golden_cases())
def test_golden_op_by_op(case, section):
    blk = _golden_block(case)
    blk.step(case["dt"], section["steps"])
    bad = []
    for rod, vars_ in section["expected"].items():
        for tag, entry in vars_.items():
            ok = rel_err_ok(blk.get_rod(tag, int(rod)), as_field(entry, tag), section["epsilon"])
            if not ok.all():
                bad.append("rod%s %s %d/%d" % (rod, tag, (~ok).sum(), ok.size))
    assert not bad, bad


@pytest.mark.parametrize("case,section", [p for p in _golden_cases() if p.values[1]["steps"] > 0])
@pytest.mark.parametrize("precision", [FP64])
def test_golden_fused_state(case, section, precision):
    blk = _golden_block(case)
    blk.set_rod_dt(case["dt"])
    blk.step_fused(section["steps"], precision)
    bad = []
    for rod, vars_ in section["expected"].items():
        for tag in STATE:
            ok = rel_err_ok(blk.get_rod(tag, int(rod)), as_field(vars_[tag], tag), max(section["epsilon"], 1e-12))
            if not ok.all():
                bad.append("rod%s %s %d/%d" % (rod, tag, (~ok).sum(), ok.size))
    assert not bad, bad


# --- the oracle on random ragged batches ----------------------------------------------

@pytest.fixture(scope="module")
def ragged():
    wl = random_workload(40, 2, 300, seed=7, steps=300)
    sysm = oracle_system(wl)
    sysm.step(wl.dt, wl.steps)
    return wl, sysm


def test_op_by_op_matches_oracle(ragged):
    wl, sysm = ragged
    blk = device_block(wl)
    blk.step(wl.dt, wl.steps)
    err = rel_state_error(blk, sysm, STATE + ("InternalLoads", "InternalTorques", "Curvature", "ShearStretchStrain"))
    assert max(err.values()) <= 1e-12, err


def test_fused_fp64_matches_oracle(ragged):
    wl, sysm = ragged
    blk = device_block(wl)
    blk.step_fused(wl.steps, FP64)
    err = rel_state_error(blk, sysm)
    assert max(err.values()) <= 1e-10, err
    assert np.allclose(blk.rod_time(), sysm.time, rtol=1e-14)


def test_fused_fp32_bulk_error_vs_oracle(ragged):
    wl, sysm = ragged
    blk = device_block(wl)
    blk.step_fused(wl.steps, FP32_BULK)
    err = rel_state_error(blk, sysm)
    assert err["Position"] <= 1e-6, err
    assert err["Director"] <= 1e-5, err


SHAPES = [(256, 1), (512, 1), (512, 2), (1024, 2), (512, 4), (1024, 4)]


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
@pytest.mark.parametrize("lanes,epl", SHAPES)
def test_fused_is_independent_of_tiling(lanes, epl, precision):
    """Per-rod arithmetic does not depend on the tile, the thread or the
    position within a thread a rod lands in.  (FP32-bulk takes the two half
    steps between consecutive steps of a launch as one, so runs compared
    bitwise use the same steps per launch.)"""
    wl = random_workload(30, 4, 250, seed=11, steps=50)
    ref = device_block(wl)
    ref.set_fused_config(256, 1, 10)
    ref.step_fused(wl.steps, precision)
    blk = device_block(wl)
    blk.set_fused_config(lanes, epl, 10)
    blk.step_fused(wl.steps, precision)
    assert blk.fused_config()[:2] == (lanes, epl)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(blk.get_rod(t, r), ref.get_rod(t, r)), (lanes, epl, r, t)


def _edge_lengths():
    # rods at the tile capacities (1024 and 512 lanes) and one off them, and
    # hundreds of two- and three-element rods that the planner packs into the
    # room the long rods leave
    return [1024, 1023, 513, 512, 511, 257, 256, 255, 2, 3] + [2] * 300 + [3] * 60 + [1024]


def test_edge_lengths_fused_fp64_equals_op_by_op():
    wl = random_workload(0, 0, 0, seed=23, steps=100, lengths=_edge_lengths())
    plain = device_block(wl)
    plain.step(wl.dt, wl.steps)
    fused = device_block(wl)
    fused.set_fused_config(0, 0, 50)
    fused.step_fused(wl.steps, FP64)
    assert fused.fused_config()[0] == 1024
    for r in range(len(wl.n_elems)):
        for t in STATE:
            a, b = fused.get_rod(t, r), plain.get_rod(t, r)
            assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1.0), (r, t)
    f32 = device_block(wl)
    f32.set_fused_config(0, 0, 50)
    f32.step_fused(wl.steps, FP32_BULK)
    for r in range(len(wl.n_elems)):
        x, y = f32.get_rod("Position", r), plain.get_rod("Position", r)
        L = np.linalg.norm(np.diff(y, axis=1), axis=0).sum()
        assert np.abs(x - y).max() <= 1e-6 * L, r


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_edge_lengths_are_independent_of_placement(precision):
    """The same rods in reverse order land in other tiles and lanes (the
    planner keeps the order of equal lengths) and step bitwise the same; a
    shape with more elements per thread than the shortest rod is refused."""
    lengths = _edge_lengths()
    wl = random_workload(0, 0, 0, seed=29, steps=60, lengths=lengths)
    rev = Workload("reversed", wl.n_elems[::-1], wl.states[::-1], wl.dt, wl.steps,
                   forces=[(len(lengths) - 1 - f[0],) + tuple(f[1:]) for f in wl.forces])
    a, b = device_block(wl), device_block(rev)
    for blk in (a, b):
        blk.set_fused_config(1024, 2, 30)
        blk.step_fused(wl.steps, precision)
    n = len(lengths)
    for r in range(n):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, n - 1 - r)), (r, t)
    a.set_fused_config(1024, 4, 30)
    with pytest.raises(ErodError):
        a.step_fused(1, precision)


def test_rejected_arguments_leave_the_batch_as_it_was():
    """Bad arguments raise and change nothing: a ramped force with an empty
    interval (its ramp would be 0/0 at the start time), a rod dt that is not
    positive (the call is all or nothing), a graph chunk of zero steps.  The
    batch then steps bitwise like an untouched one."""
    wl = random_workload(4, 8, 40, seed=71, steps=50)
    a, b = device_block(wl), device_block(wl)
    dts = np.asarray([wl.dt * (1.0 - 0.1 * r) for r in range(4)])
    for blk in (a, b):
        blk.set_rod_dt(dts)
    with pytest.raises(ErodError, match="interval"):
        a.add_end_force(0, [1.0, 0.0, 0.0], ramp=1, start=0.0, interval=0.0)
    with pytest.raises(ErodError, match="interval"):
        a.add_end_force(0, [1.0, 0.0, 0.0], ramp=2, start=0.0, interval=float("nan"))
    bad = dts.copy()
    bad[2] = -1.0
    with pytest.raises(ErodError, match="positive"):
        a.set_rod_dt(bad)
    with pytest.raises(ErodError, match="chunk"):
        a.step_fused_graph(30, 0, FP64)
    for blk in (a, b):
        blk.step_fused(wl.steps, FP64)
    for r in range(4):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)
    assert np.array_equal(a.rod_time(), b.rod_time())
    a.close()
    b.close()


def test_autotune_keeps_state_and_picks_a_shape():
    wl = random_workload(25, 4, 200, seed=31, steps=30)
    a = device_block(wl)
    ms = a.autotune(FP32_BULK, 5)
    assert ms > 0 and a.fused_config()[0] in (256, 512, 1024)
    for r in range(len(wl.n_elems)):
        assert np.array_equal(a.get_rod("Position", r), wl.states[r]["Position"])
    a.step_fused(wl.steps, FP32_BULK)
    b = device_block(wl)
    b.step_fused(wl.steps, FP32_BULK)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)


def test_fp64_fused_is_independent_of_steps_per_launch():
    wl = random_workload(10, 4, 120, seed=19, steps=60)
    a = device_block(wl)
    a.set_fused_config(0, 0, 60)
    a.step_fused(60, FP64)
    b = device_block(wl)
    b.set_fused_config(0, 0, 7)
    b.step_fused(60, FP64)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)


def test_fp32_bulk_launch_segmentation_is_a_rounding_effect():
    """Merging the half steps inside a launch changes x at FP64 rounding
    level; through the FP32 rounding of node differences that can grow to
    FP32-bulk's own rounding noise, still 1000x below its accuracy bound."""
    wl = random_workload(10, 4, 120, seed=19, steps=60)
    a = device_block(wl)
    a.set_fused_config(0, 0, 60)
    a.step_fused(60, FP32_BULK)
    b = device_block(wl)
    b.set_fused_config(0, 0, 1)
    b.step_fused(60, FP32_BULK)
    for r in range(len(wl.n_elems)):
        x, y = a.get_rod("Position", r), b.get_rod("Position", r)
        assert np.abs(x - y).max() <= 1e-9 * max(1.0, np.abs(y).max()), r


@pytest.mark.parametrize("spl", [10, 20])
def test_fused_graph_replay_equals_direct(spl):
    """Chunks of 20 steps: two launches replayed as a captured graph (10 steps
    per launch), or one launch, which goes out as that launch (20)."""
    wl = random_workload(20, 2, 120, seed=13, steps=60)
    a = device_block(wl)
    a.set_fused_config(0, 0, spl)
    a.step_fused(60, FP32_BULK)
    b = device_block(wl)
    b.set_fused_config(0, 0, spl)
    b.step_fused_graph(60, 20, FP32_BULK)
    assert b.last_kernel_ms()[1] == 60 // spl
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)


@pytest.mark.parametrize("precision", [FP32_BULK, FP64])
def test_fused_graph_takes_steps_that_do_not_fill_a_chunk(precision):
    """n not a multiple of the chunk: the whole chunks are replayed and the
    rest launched directly with the chunk's launch lengths, so the run equals
    the direct path over the same steps (also with fewer steps than one
    chunk), and the launch count covers both parts."""
    wl = random_workload(20, 2, 120, seed=17, steps=0)
    for n, chunk in ((75, 20), (5, 20), (20, 20)):
        a, b = device_block(wl), device_block(wl)
        for blk in (a, b):
            blk.set_fused_config(0, 0, 10)
        a.step_fused(n, precision)
        b.step_fused_graph(n, chunk, precision)
        assert b.last_kernel_ms()[1] == -(-n // 10), (n, chunk)
        for r in range(len(wl.n_elems)):
            for t in STATE:
                assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (n, chunk, r, t)
        assert np.array_equal(a.rod_time(), b.rod_time()), (n, chunk)
        a.close()
        b.close()


@pytest.mark.parametrize("precision", [FP32_BULK, FP64])
def test_graph_replays_with_eager_work_between(precision):
    """One captured chunk (two launches of 10 steps) replayed four times, each
    replay followed by eager work on the device (another batch's fused steps,
    the finiteness check with its clear, downloads): every replay equals the
    eager path advanced by the same steps."""
    wl = random_workload(20, 2, 120, seed=19, steps=0)
    a, b = device_block(wl), device_block(wl)
    for blk in (a, b):
        blk.set_fused_config(0, 0, 10)
    other = device_block(random_workload(5, 4, 30, seed=20, steps=0))
    for rep in range(4):
        a.step_fused(20, precision)
        b.step_fused_graph(20, 20, precision)
        other.step_fused(7, FP64)
        assert not b.check_finite().any()
        for r in range(len(wl.n_elems)):
            for t in STATE:
                assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (rep, r, t)
        assert np.array_equal(a.rod_time(), b.rod_time()), rep
    for blk in (a, b, other):
        blk.close()


def test_rod_alone_equals_rod_in_batch():
    wl = random_workload(12, 2, 90, seed=17, steps=40)
    full = device_block(wl)
    full.step_fused(wl.steps, FP64)
    from elasticapp_amd.workloads import Workload
    for r in (0, 5, 11):
        one = Workload("one", [wl.n_elems[r]], [wl.states[r]], wl.dt, wl.steps,
                       forces=[(0,) + tuple(wl.forces[r][1:])])
        blk = device_block(one)
        blk.step_fused(wl.steps, FP64)
        for t in STATE:
            assert np.array_equal(blk.get_rod(t, 0), full.get_rod(t, r)), (r, t)


# --- hooks: clamp, ramped tip force, gravity, plane contact ---------------------------

def _hooked_workload():
    wl = random_workload(6, 8, 60, seed=23, steps=400, forces=False)
    for r in range(3):
        st = wl.states[r]
        wl.clamps.append((r, st["Position"][:, 0].copy(), st["Director"][:, 0].copy()))
        wl.forces.append((r, np.array([0.3, -0.2, 0.1]), r % 3, 0.0, wl.dt * 200))
    wl.gravity = (0.0, 0.0, -9.81)
    zmin = min(st["Position"][2].min() for st in wl.states)
    wl.contact = (1e3, 10.0, 0.3, 100.0, zmin + 0.01, 0.0)
    return wl


def test_hooks_op_by_op_and_fused_match_oracle():
    wl = _hooked_workload()
    sysm = oracle_system(wl)
    sysm.step(wl.dt, wl.steps)
    a = device_block(wl)
    a.step(wl.dt, wl.steps)
    err = rel_state_error(a, sysm)
    assert max(err.values()) <= 1e-12, err
    b = device_block(wl)
    b.step_fused(wl.steps, FP64)
    err = rel_state_error(b, sysm)
    assert max(err.values()) <= 1e-10, err
    # clamped ends stay put
    for r in range(3):
        assert np.array_equal(b.get_rod("Position", r)[:, 0], wl.states[r]["Position"][:, 0])


def _verlet_by_calls(blk, dt, n, t=0.0):
    # the reference stepper's sequence over one block (PositionVerlet.hpp:49-88,
    # SymplecticSteppers.hpp:152-216): kinematic half step, post-kinematic
    # hooks, internal dynamics (external loads zeroed first), pre-dynamic hooks,
    # dynamic step, post-dynamic hooks, kinematic half step
    for _ in range(n):
        blk.kinematic_update(dt)
        t += 0.5 * dt
        blk.apply_post_kinematic(t)
        blk.compute_internal_dynamics(t)
        blk.apply_pre_dynamic(t)
        blk.dynamic_update(dt)
        blk.apply_post_dynamic(t)
        t += 0.5 * dt
        blk.kinematic_update(dt)
    return t


def test_per_block_calls_compose_the_step():
    """The three per-block calls and the hook stages, driven one by one from
    the host as the reference's stepper drives a block, equal erod_step
    bitwise in every field (hooks: clamps, ramped end forces, gravity, plane
    contact with friction)."""
    wl = _hooked_workload()
    a, b = device_block(wl), device_block(wl)
    _verlet_by_calls(a, wl.dt, 60)
    b.step(wl.dt, 60)
    for tag in FIELDS:
        for r in range(len(wl.n_elems)):
            assert np.array_equal(a.get_rod(tag, r), b.get_rod(tag, r)), (tag, r)


@pytest.mark.parametrize("case,section", [p for p in _golden_cases() if p.values[1]["steps"] > 0])
def test_golden_per_block_calls(case, section):
    """The golden steps through the three per-block calls, every variable the
    reference's test compares, at its tolerances."""
    blk = _golden_block(case)
    _verlet_by_calls(blk, case["dt"], section["steps"])
    bad = []
    for rod, vars_ in section["expected"].items():
        for tag, entry in vars_.items():
            ok = rel_err_ok(blk.get_rod(tag, int(rod)), as_field(entry, tag), section["epsilon"])
            if not ok.all():
                bad.append("rod%s %s %d/%d" % (rod, tag, (~ok).sum(), ok.size))
    assert not bad, bad


def test_finite_flags():
    wl = random_workload(5, 4, 20, seed=29, steps=10)
    blk = device_block(wl)
    blk.step_fused(10, FP64)
    assert not blk.check_finite().any()
    v = blk.get_rod("Velocity", 3)
    v[1, 2] = np.nan
    blk.set_rod("Velocity", 3, v)
    flags = blk.check_finite()
    assert flags.tolist() == [0, 0, 0, 1, 0]


@pytest.mark.parametrize("ramp", [1, 2])
def test_device_ramp_matches_the_reference_known_answers(ramp):
    """The end force the device hook applies at time t, F ramp(t), at the
    points Test_Rampup.cpp checks (start -2, interval 4) and between them,
    against the oracle's ramp (the device sin may differ in the last bit)."""
    from oracle import pyoracle
    st = straight_rod_state(6, 0.01, 1.0, [0, 0, 0], [1, 0, 0], [0, 1, 0], 1e3, 1e6, 4e5, 0.1, 0.1)
    blk = DeviceBlock([6])
    for tag, a in st.items():
        blk.set_rod(tag, 0, a)
    s, T = -2.0, 4.0
    blk.add_end_force(0, [1.0, -2.0, 0.5], ramp=ramp, start=s, interval=T)
    for t in [-2.5, s] + [s + f * T for f in np.linspace(0.05, 1.0, 20)] + [s + 1.5 * T]:
        blk.compute_internal_dynamics(t)
        blk.apply_pre_dynamic(t)
        got = blk.get_rod("ExternalLoads", 0)[:, -1]
        want = pyoracle.ramp_value(ramp, t, s, T)
        assert np.abs(got - want * np.array([1.0, -2.0, 0.5])).max() <= 2e-16 * 2.0, (t, got, want)
    blk.close()


def test_batch_length_pins_two_ghost_elements_per_gap():
    """Test_CosseratRod.py:13-22 on the device batch: 4 * 10 + 3 * 2 = 46
    element slots, rods starting 12 slots apart."""
    blk = DeviceBlock([10] * 4)
    assert blk.element_slots == 46 and blk.starts == [0, 12, 24, 36]
    assert blk.get_field("Position").shape == (3, 47) and blk.get_field("ReferenceCurvature").shape == (3, 45)
    blk.close()


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_split_plan_equals_single_plan(precision):
    """Rods that find no room beside longer rods in 1024-lane tiles run in a
    second group of 512-lane tiles, launched after the first: every rod steps
    bitwise as in the single plan (the arithmetic does not depend on the tile).
    With a split shape of more elements per thread than some rods have,
    those rods stay in the main group, and the run is still the same."""
    wl = random_workload(0, 0, 0, seed=37, steps=60, lengths=_edge_lengths())
    a, b = device_block(wl), device_block(wl)
    for blk in (a, b):
        blk.set_fused_config(1024, 2, 30)
    b.set_fused_split(512, 2)
    a.step_fused(wl.steps, precision)
    b.step_fused(wl.steps, precision)
    lanes, epl, tiles = b.fused_split()
    assert (lanes, epl) == (512, 2) and tiles > 0 and a.fused_split() == (0, 0, 0)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)
    assert np.array_equal(a.rod_time(), b.rod_time())
    b.step_fused_graph(60, 30, precision)
    a.step_fused_graph(60, 30, precision)
    for r in range(0, len(wl.n_elems), 7):
        assert np.array_equal(a.get_rod("Position", r), b.get_rod("Position", r)), r
    b.set_fused_split(512, 4)
    a.step_fused(30, precision)
    b.step_fused(30, precision)
    assert b.fused_split()[:2] == (512, 4)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)
    for blk in (a, b):
        blk.close()


def test_autotuned_plan_steps_like_the_default_and_keeps_a_set_second_group():
    """On a batch of short rods no second group can help (every rod fits the
    smallest tiles' room), and the autotuner's plan steps bitwise like the
    default plan; a second group set by the caller is kept as set."""
    wl = random_workload(40, 4, 60, seed=43, steps=40)
    a, b = device_block(wl), device_block(wl)
    a.autotune(FP64, 10)
    assert a.fused_split()[2] == 0 or a.fused_split()[0] < a.fused_config()[0]
    a.step_fused(wl.steps, FP64)
    b.step_fused(wl.steps, FP64)
    for r in range(len(wl.n_elems)):
        for t in STATE:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)
    c = device_block(wl)
    c.set_fused_config(256, 1, 0)
    c.set_fused_split(128, 1)
    c.autotune(FP64, 10)
    assert c.fused_split()[:2] in ((128, 1), (0, 0))
    for blk in (a, b, c):
        blk.close()

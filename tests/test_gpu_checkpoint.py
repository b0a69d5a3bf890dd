"""Checkpoints of device batches: a run resumed from erod_save / erod_load
continues bit for bit (op-by-op and both fused precisions, with hooks, step
budgets and shortened slots), and the numpy reader returns exactly what the
device holds."""
import numpy as np
import pytest

from elasticapp_amd import FIELDS, FP32_BULK, FP64, DeviceBlock, ErodError
from elasticapp_amd import checkpoint
from elasticapp_amd.workloads import build_block
from tests.gpu_helpers import STATE, random_workload

pytestmark = pytest.mark.gpu


def _hooked():
    wl = random_workload(7, 8, 70, seed=53, steps=0, forces=False)
    for r in range(3):
        st = wl.states[r]
        wl.clamps.append((r, st["Position"][:, 0].copy(), st["Director"][:, 0].copy()))
        wl.forces.append((r, np.array([0.2, -0.1, 0.3]), 1 + r % 2, 0.0, wl.dt * 120))
    wl.gravity = (0.0, 0.0, -9.81)
    # a slightly tilted plane just above the lowest node
    n = np.array([0.1, 0.0, 1.0]) / np.sqrt(1.01)
    hmin = min((n @ st["Position"]).min() for st in wl.states)
    wl.contact = (1e3, 10.0, 0.3, 10.0, hmin + 0.01, 0.0, tuple(n))
    return wl


def _same(a, b, tags):
    for r in range(len(a.n_elems)):
        if a.active_length(r) == 0:
            continue
        for t in tags:
            assert np.array_equal(a.get_rod(t, r), b.get_rod(t, r)), (r, t)


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_fused_run_resumes_bitwise(tmp_path, precision):
    wl = _hooked()
    a = build_block(wl)
    a.set_step_budget(1, 150)
    a.set_active_length(4, max(2, wl.n_elems[4] - 5))
    a.step_fused(100, precision)
    path = str(tmp_path / "run.ckpt")
    a.save(path)
    a.step_fused(100, precision)
    b = DeviceBlock.load(path)
    assert b.n_elems == a.n_elems and b.step_budget(1) == 50 and b.active_length(4) == a.active_length(4)
    b.step_fused(100, precision)
    _same(a, b, STATE)
    assert np.array_equal(a.rod_time(), b.rod_time())
    assert b.step_budget(1) == 0 and b.step_budget(0) is None


def test_op_by_op_run_resumes_bitwise(tmp_path):
    wl = _hooked()
    a = build_block(wl)
    a.step(wl.dt, 60)
    path = str(tmp_path / "ops.ckpt")
    a.save(path)
    t_saved = a.time
    a.step(wl.dt, 60)
    b = DeviceBlock.load(path)
    assert b.time == t_saved
    b.step(wl.dt, 60)
    assert b.time == a.time
    _same(a, b, list(FIELDS))


def test_reader_returns_device_contents(tmp_path):
    wl = _hooked()
    a = build_block(wl)
    a.step_fused(40, FP32_BULK)
    path = str(tmp_path / "r.ckpt")
    a.save(path)
    ck = checkpoint.read(path)
    assert list(ck["n_elems"]) == a.n_elems and ck["damped"]
    assert np.array_equal(ck["rod_time"], a.rod_time())
    for tag in FIELDS:
        assert np.array_equal(ck["fields"][tag], a.get_field(tag)), tag
    for r in (0, 3, 6):
        for tag in STATE:
            assert np.array_equal(checkpoint.rod_field(ck, tag, r), a.get_rod(tag, r)), (r, tag)
    assert ck["env"]["gravity"] and ck["env"]["contact"] and ck["env"]["plane"]["k"] == 1e3
    n = np.array([0.1, 0.0, 1.0]) / np.sqrt(1.01)
    assert np.allclose(ck["env"]["plane"]["normal"], n, rtol=0, atol=1e-16) and ck["version"] == 3
    assert ck["rods"][0]["flags"] & 1 and ck["rods"][1]["ramp"] == 2 and ck["rods"][0]["dt"] == wl.dt
    assert ck["rods"][0]["clamp_index"] == (0, 0) and not ck["rods"][0]["flags"] & 8


def test_load_rejects_a_foreign_file(tmp_path):
    p = tmp_path / "bad.ckpt"
    p.write_bytes(b"ERODCKP1" + b"\x00" * 40)
    with pytest.raises((ErodError, ValueError)):
        DeviceBlock.load(str(p))


def test_version_one_file_loads_with_the_level_plane(tmp_path):
    """A version 1 checkpoint (written before the plane normal and the clamp
    indices existed: the same layout without the 104 bytes of clamp indices
    and second clamp after each rod record and without the normal's 24
    bytes) loads and steps as the batch it came from with its plane set
    level."""
    import struct
    wl = _hooked()
    wl.contact = tuple(wl.contact[:6])  # level plane
    a = build_block(wl)
    a.step_fused(50, FP64)
    path = tmp_path / "v3.ckpt"
    a.save(str(path))
    data = bytearray(path.read_bytes())
    n = len(wl.n_elems)
    # header 32 B; n_elems, n_active (u64), steps_left (u32); budgets (u8);
    # rod clocks (f64); 168 + 104 B per rod record; environment 80 B; then the normal
    head = 32 + n * (8 + 8 + 4) + 1 + 8 * n
    recs = b"".join(bytes(data[head + 272 * r:head + 272 * r + 168]) for r in range(n))
    at = head + 272 * n + 80
    assert struct.unpack_from("<3d", data, at) == (0.0, 0.0, 1.0)
    v1 = data[:8] + struct.pack("<I", 1) + data[12:head] + recs + data[head + 272 * n:at] + data[at + 24:]
    p1 = tmp_path / "v1.ckpt"
    p1.write_bytes(bytes(v1))
    assert checkpoint.read(str(p1))["version"] == 1
    b = DeviceBlock.load(str(p1))
    a.step_fused(50, FP64)
    b.step_fused(50, FP64)
    _same(a, b, STATE)

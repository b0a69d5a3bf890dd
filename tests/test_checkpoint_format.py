"""The numpy checkpoint reader against a file packed by hand from the layout
documented in erod_io.hip (no GPU)."""
import ctypes
import os
import struct

import numpy as np
import pytest

import elasticapp_amd as ea
from elasticapp_amd import checkpoint


def _pack(n_elems, version=2):
    ea.load_library()
    n_rods = len(n_elems)
    E = sum(n_elems) + 2 * n_rods - 2
    b = bytearray(b"ERODCKP1")
    b += struct.pack("<IIQQ", version, 1, n_rods, E)
    b += np.array(n_elems, "<u8").tobytes() + np.array(n_elems, "<u8").tobytes()
    b += np.array([7] * n_rods, "<u4").tobytes() + b"\x01"
    b += np.arange(n_rods, dtype="<f8").tobytes()
    for r in range(n_rods):
        flags = 2 | (1 | 8 if version >= 3 and r == 0 else 0)
        b += struct.pack("<3dI3d9d3di2d", 1e-3 * (r + 1), 0.1, 0.2, flags, 0, 0, 0, *([0.0] * 9), 1.0, 2.0, 3.0, 1, 0.0, 5.0)
        if version >= 3:
            b += struct.pack("<2I3d9d", 0, n_elems[r], 0.5, 0.25, r, *np.eye(3).ravel())
    b += struct.pack("<i3di6d", 1, 0.0, 0.0, -9.81, 0, 0, 0, 0, 0, 0, 0)
    if version >= 2:
        b += struct.pack("<3d", 0.6, 0.0, 0.8)
    b += struct.pack("<I", len(ea.FIELDS))
    fields = {}
    for name, (fid, nc, st) in sorted(ea.FIELDS.items(), key=lambda kv: kv[1][0]):
        cnt = E + 1 if st == 0 else (E if st == 1 else E - 1)
        a = np.random.default_rng(fid).normal(size=(nc, cnt))
        fields[name] = a
        b += struct.pack("<IIQ", fid, nc, cnt) + a.astype("<f8").tobytes()
    b += b"ERODEND1"
    return bytes(b), fields


def test_reader_parses_the_documented_layout(tmp_path):
    data, fields = _pack([5, 3, 9])
    p = tmp_path / "x.ckpt"
    p.write_bytes(data)
    assert checkpoint.read_header(str(p)) == (True, [5, 3, 9])
    ck = checkpoint.read(str(p))
    assert list(ck["starts"]) == [0, 7, 12] and ck["budgets"] and list(ck["steps_left"]) == [7, 7, 7]
    assert ck["rods"][2]["dt"] == 3e-3 and list(ck["rods"][1]["force"]) == [1.0, 2.0, 3.0] and ck["rods"][0]["ramp"] == 1
    assert ck["env"]["gravity"] and not ck["env"]["contact"] and ck["env"]["g"][2] == -9.81
    assert list(ck["env"]["plane"]["normal"]) == [0.6, 0.0, 0.8]
    for name, a in fields.items():
        assert np.array_equal(ck["fields"][name], a), name
    x = checkpoint.rod_field(ck, "Position", 1)
    assert np.array_equal(x, fields["Position"][:, 7:11])


def test_reader_reads_the_clamp_indices_of_version_three(tmp_path):
    data, fields = _pack([5, 3], version=3)
    p = tmp_path / "v3.ckpt"
    p.write_bytes(data)
    ck = checkpoint.read(str(p))
    assert ck["version"] == 3 and ck["rods"][0]["flags"] == 11
    assert ck["rods"][0]["clamp_index"] == (0, 5) and ck["rods"][1]["clamp_index"] == (0, 3)
    assert list(ck["rods"][1]["clamp1_x"]) == [0.5, 0.25, 1.0]
    assert np.array_equal(ck["rods"][0]["clamp1_q"], np.eye(3).ravel())
    assert np.array_equal(ck["fields"]["Director"], fields["Director"])


def test_reader_reads_version_one_with_the_horizontal_plane(tmp_path):
    data, fields = _pack([5, 3], version=1)
    p = tmp_path / "v1.ckpt"
    p.write_bytes(data)
    ck = checkpoint.read(str(p))
    assert ck["version"] == 1 and list(ck["env"]["plane"]["normal"]) == [0.0, 0.0, 1.0]
    assert np.array_equal(ck["fields"]["Velocity"], fields["Velocity"])


def test_reader_rejects_truncation(tmp_path):
    data, _ = _pack([4, 4])
    p = tmp_path / "t.ckpt"
    p.write_bytes(data[:-20])
    with pytest.raises(ValueError):
        checkpoint.read(str(p))


def test_load_refuses_a_rod_count_the_file_cannot_hold(tmp_path):
    # a corrupt count must fail on the header, before any allocation or device call
    data, _ = _pack([4, 4])
    bad = data[:16] + struct.pack("<Q", 1 << 31) + data[24:]
    p = tmp_path / "n.ckpt"
    p.write_bytes(bad)
    with pytest.raises(ValueError):
        checkpoint.read_header(str(p))
    L = ea.load_library()
    h = ctypes.c_void_p()
    assert L.erod_load(os.fsencode(str(p)), ctypes.byref(h)) != 0 and not h.value
    assert b"rod count exceeds" in L.erod_last_error()

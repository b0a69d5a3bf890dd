"""Reader of the batch checkpoints written by erod_save (numpy; no GPU).

The file layout is documented in elasticapp_amd/csrc/erod_io.hip.  `read`
returns every array of the batch as host data, so that a checkpoint written
on a GPU can be inspected or compared offline.
"""
import struct

import numpy as np

MAGIC, END = b"ERODCKP1", b"ERODEND1"
_ROD = struct.Struct("<3dI3d9d3di2d")
_CLAMP2 = struct.Struct("<2I3d9d")   # version 3: clamp indices, the second clamp's pose


def _field_names():
    from . import _FIELD_NAMES
    return list(_FIELD_NAMES)


def read_header(path):
    """(damped, n_elems) of a checkpoint, without reading its fields."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError("not an erod checkpoint: %s" % path)
        version, damped, n_rods, _ = struct.unpack("<IIQQ", f.read(24))
        if version not in (1, 2, 3):
            raise ValueError("unsupported checkpoint version %d" % version)
        raw = f.read(8 * n_rods)
        if len(raw) != 8 * n_rods:
            raise ValueError("truncated checkpoint header: %s" % path)
        n_elems = np.frombuffer(raw, dtype="<u8")
    return bool(damped), [int(x) for x in n_elems]


def read(path):
    names = _field_names()
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != MAGIC or data[-8:] != END:
        raise ValueError("not a complete erod checkpoint: %s" % path)
    o = 8
    version, damped, n_rods, slots = struct.unpack_from("<IIQQ", data, o)
    o += 24
    if version not in (1, 2, 3):
        raise ValueError("unsupported checkpoint version %d" % version)

    def arr(dtype, n):
        nonlocal o
        a = np.frombuffer(data, dtype=dtype, count=n, offset=o).copy()
        o += a.nbytes
        return a

    out = {"version": version, "damped": bool(damped), "element_slots": slots}
    out["n_elems"] = arr("<u8", n_rods)
    out["n_active"] = arr("<u8", n_rods)
    out["steps_left"] = arr("<u4", n_rods)
    out["budgets"] = bool(data[o])
    o += 1
    out["rod_time"] = arr("<f8", n_rods)
    rods = []
    for _ in range(n_rods):
        v = _ROD.unpack_from(data, o)
        o += _ROD.size
        rod = {"dt": v[0], "gamma_f": v[1], "gamma_t": v[2], "flags": v[3], "clamp_x": np.array(v[4:7]),
               "clamp_q": np.array(v[7:16]), "force": np.array(v[16:19]), "ramp": v[19],
               "ramp_start": v[20], "ramp_interval": v[21], "clamp_index": (0, 0)}
        # version 3: the clamps' indices and a second clamp (flag bit 3)
        if version >= 3:
            w = _CLAMP2.unpack_from(data, o)
            o += _CLAMP2.size
            rod["clamp_index"] = (w[0], w[1])
            rod["clamp1_x"], rod["clamp1_q"] = np.array(w[2:5]), np.array(w[5:14])
        rods.append(rod)
    out["rods"] = rods
    grav, = struct.unpack_from("<i", data, o)
    o += 4
    g = struct.unpack_from("<3d", data, o)
    o += 24
    contact, = struct.unpack_from("<i", data, o)
    o += 4
    plane = dict(zip(("k", "c", "mu", "ct", "z0", "offset"), struct.unpack_from("<6d", data, o)))
    o += 48
    # version 2 adds the plane's unit normal; version 1 planes are z = z0
    if version >= 2:
        plane["normal"] = np.array(struct.unpack_from("<3d", data, o))
        o += 24
    else:
        plane["normal"] = np.array([0.0, 0.0, 1.0])
    out["env"] = {"gravity": bool(grav), "g": np.array(g), "contact": bool(contact), "plane": plane}
    nf, = struct.unpack_from("<I", data, o)
    o += 4
    fields = {}
    for _ in range(nf):
        fid, nc, cnt = struct.unpack_from("<IIQ", data, o)
        o += 16
        fields[names[fid]] = arr("<f8", nc * cnt).reshape(nc, cnt)
    out["fields"] = fields
    if data[o:o + 8] != END:
        raise ValueError("checkpoint layout mismatch")
    starts = np.concatenate([[0], np.cumsum(out["n_elems"].astype(np.int64) + 2)[:-1]])
    out["starts"] = starts
    return out


def rod_field(ckpt, tag, rod):
    """One rod's slice of a field, as DeviceBlock.get_rod returns it."""
    from . import ELEMENT, FIELDS, NODE, load_library
    load_library()
    _, _, st = FIELDS[tag]
    n = int(ckpt["n_active"][rod])
    cnt = n + 1 if st == NODE else (n if st == ELEMENT else n - 1)
    s = int(ckpt["starts"][rod])
    return ckpt["fields"][tag][:, s:s + cnt]

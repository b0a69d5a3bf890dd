"""elasticapp_amd — an MI355X-native Cosserat-rod block timestep.

Python host layer over the C ABI in include/erod.h (liberod.so, HIP kernels
for gfx950).  `DeviceBlock` mirrors the reference's blocks::Block<CosseratRod>
(elastica/Systems/CosseratRods/Block.hpp): rods live in one ghost-padded,
component-major SoA in HBM; `block.rod(i)[tag]` reads or writes a rod slice the
way blocks::get<tag>(slice) does, and the three stepper calls
(kinematic_update, compute_internal_dynamics, dynamic_update) are the
per-block operations the reference's position-Verlet stepper issues
(elastica/Time/SymplecticSteppers/Algorithms/PositionVerlet.hpp:49-88).

There is no CPU fallback: every call goes to the HIP library and fails loudly
when it is missing or no GPU is visible.
"""
import ctypes
import os

import numpy as np

__all__ = ["DeviceBlock", "Server", "FIELDS", "FP64", "FP32_BULK", "ErodError", "load_library", "library_path",
           "straight_rod_state", "assemble", "RAMP_NONE", "RAMP_LINEAR", "RAMP_SIN"]

_HERE = os.path.dirname(os.path.abspath(__file__))
M_1_PI = 0.318309886183790671537767526745028724  # <cmath> M_1_PI, correctly rounded
FP64, FP32_BULK = 0, 1
RAMP_NONE, RAMP_LINEAR, RAMP_SIN = 0, 1, 2
NODE, ELEMENT, VORONOI = 0, 1, 2

# field name -> (id, ncomp, stagger), in the order of erod_field (include/erod.h)
_FIELD_NAMES = [
    "Position", "Velocity", "Acceleration", "Mass", "InvMass", "InternalLoads", "ExternalLoads",
    "ForceDampingRate",
    "Director", "AngularVelocity", "AngularAcceleration", "ReferenceElementLength", "ElementLength",
    "ElementDilatation", "Tangent", "ElementDimension", "ElementVolume", "ShearStretchRigidityMatrix",
    "MassSecondMomentOfInertia", "InvMassSecondMomentOfInertia", "ShearStretchStrain",
    "ReferenceShearStretchStrain", "InternalStress", "InternalTorques", "ExternalTorques",
    "TorqueDampingRate", "_DummyElementVector", "_DummyElementVector2",
    "ReferenceVoronoiLength", "VoronoiLength", "VoronoiDilatation", "Curvature", "ReferenceCurvature",
    "BendingTwistRigidityMatrix", "InternalCouple", "_DummyVoronoiVector",
]
FIELDS = {}

_LIB = None


class ErodError(RuntimeError):
    pass


def library_path():
    return os.path.join(_HERE, "liberod.so")


def load_library():
    """Load liberod.so (built by elasticapp_amd/build.py); raise if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise ErodError("liberod.so is not built: run `python -m elasticapp_amd.build`")
    L = ctypes.CDLL(path)
    u64, i32, dbl, vp = ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    P = ctypes.POINTER
    sig = {
        "erod_last_error": (ctypes.c_char_p, []),
        "erod_field_shape": (i32, [i32, P(i32), P(i32)]),
        "erod_batch_create": (i32, [P(u64), u64, i32, P(vp)]),
        "erod_batch_destroy": (i32, [vp]),
        "erod_batch_device": (i32, [vp, P(i32)]),
        "erod_batch_slots": (i32, [vp, P(u64)]),
        "erod_rod_start": (i32, [vp, u64, P(u64)]),
        "erod_upload_field": (i32, [vp, i32, P(dbl)]),
        "erod_download_field": (i32, [vp, i32, P(dbl)]),
        "erod_upload_rod": (i32, [vp, i32, u64, P(dbl)]),
        "erod_download_rod": (i32, [vp, i32, u64, P(dbl)]),
        "erod_fill_ghosts": (i32, [vp]),
        "erod_kinematic_update": (i32, [vp, dbl]),
        "erod_kinematic_update_by": (i32, [vp, dbl]),
        "erod_fill_field": (i32, [vp, i32, dbl]),
        "erod_compute_internal_dynamics": (i32, [vp, dbl]),
        "erod_dynamic_update": (i32, [vp, dbl]),
        "erod_add_clamp": (i32, [vp, u64, u64, P(dbl), P(dbl)]),
        "erod_add_end_force": (i32, [vp, u64, P(dbl), i32, dbl, dbl]),
        "erod_set_gravity": (i32, [vp, P(dbl)]),
        "erod_set_plane_contact": (i32, [vp, dbl, dbl, dbl, dbl, dbl, dbl]),
        "erod_set_plane_normal": (i32, [vp, P(dbl)]),
        "erod_apply_post_kinematic": (i32, [vp, dbl]),
        "erod_apply_pre_dynamic": (i32, [vp, dbl]),
        "erod_apply_post_dynamic": (i32, [vp, dbl]),
        "erod_step": (i32, [vp, P(dbl), dbl, u64]),
        "erod_set_rod_dt": (i32, [vp, P(dbl)]),
        "erod_get_rod_time": (i32, [vp, P(dbl)]),
        "erod_set_rod_time": (i32, [vp, P(dbl)]),
        "erod_set_fused_config": (i32, [vp, i32, i32, i32]),
        "erod_get_fused_config": (i32, [vp, P(i32), P(i32), P(i32)]),
        "erod_set_fused_split": (i32, [vp, i32, i32]),
        "erod_get_fused_split": (i32, [vp, P(i32), P(i32), P(u64)]),
        "erod_set_fused_pieces": (i32, [vp, i32, i32, i32]),
        "erod_get_fused_pieces": (i32, [vp, P(i32), P(i32), P(i32), P(u64), P(u64)]),
        "erod_autotune_fused": (i32, [vp, i32, i32, P(dbl)]),
        "erod_step_fused": (i32, [vp, i32, u64]),
        "erod_step_fused_graph": (i32, [vp, i32, u64, u64]),
        "erod_capture_fused_graph": (i32, [vp, i32, u64]),
        "erod_reorthonormalize": (i32, [vp]),
        "erod_set_reortho_interval": (i32, [vp, u64]),
        "erod_get_reortho_interval": (i32, [vp, P(u64), P(u64)]),
        "erod_last_kernel_ms": (i32, [vp, P(dbl), P(u64)]),
        "erod_synchronize": (i32, [vp]),
        "erod_check_finite": (i32, [vp, P(ctypes.c_uint8)]),
        "erod_check_finite_noalloc": (i32, [vp, P(ctypes.c_uint8)]),
        "erod_save": (i32, [vp, ctypes.c_char_p]),
        "erod_load": (i32, [ctypes.c_char_p, P(vp)]),
        "erod_set_rod_active_length": (i32, [vp, u64, u64]),
        "erod_get_rod_active_length": (i32, [vp, u64, P(u64)]),
        "erod_set_rod_step_budget": (i32, [vp, u64, u64]),
        "erod_get_rod_step_budget": (i32, [vp, u64, P(u64)]),
        "erod_copy_rods": (i32, [vp, P(u64), vp, P(u64), u64]),
        "erod_server_create": (i32, [vp, P(u64), P(u64), i32, i32, i32, dbl, P(vp)]),
        "erod_server_create_budgeted": (i32, [vp, P(u64), i32, i32, i32, dbl, P(vp)]),
        "erod_server_pool_info": (i32, [vp, P(u64), P(u64), P(dbl)]),
        "erod_batch_rods": (i32, [vp, P(u64), P(u64)]),
        "erod_server_destroy": (i32, [vp]),
        "erod_server_pool": (i32, [vp, P(vp)]),
        "erod_server_submit": (i32, [vp, u64, P(u64), u64, u64]),
        "erod_server_tick": (i32, [vp, P(u64), u64, P(u64)]),
        "erod_server_read": (i32, [vp, u64, u64, i32, P(dbl)]),
        "erod_server_release": (i32, [vp, u64]),
        "erod_server_status": (i32, [vp, P(u64), P(u64), P(u64), P(u64)]),
        "erod_server_stats": (i32, [vp, P(u64), P(u64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    for i, name in enumerate(_FIELD_NAMES):
        nc, st = ctypes.c_int(), ctypes.c_int()
        if L.erod_field_shape(i, ctypes.byref(nc), ctypes.byref(st)) != 0:
            raise ErodError("field table mismatch")
        FIELDS[name] = (i, nc.value, st.value)
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        msg = _LIB.erod_last_error().decode() if _LIB is not None else ""
        raise ErodError("erod call failed (%d): %s" % (rc, msg))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class RodSlice:
    """blocks::get<tag>(slice) for one rod: `slice['Position']` returns a host
    copy (ncomp, n+1 | n | n-1); assignment uploads it."""

    def __init__(self, block, rod):
        self.block, self.rod = block, rod

    def __getitem__(self, tag):
        return self.block.get_rod(tag, self.rod)

    def __setitem__(self, tag, value):
        self.block.set_rod(tag, self.rod, value)


class DeviceBlock:
    """A device-resident ghost-padded block of Cosserat rods."""

    def __init__(self, n_elems, damped=True):
        L = load_library()
        self.n_elems = [int(n) for n in n_elems]
        arr = (ctypes.c_uint64 * len(self.n_elems))(*self.n_elems)
        h = ctypes.c_void_p()
        _check(L.erod_batch_create(arr, len(self.n_elems), 1 if damped else 0, ctypes.byref(h)))
        self._h = h
        e = ctypes.c_uint64()
        _check(L.erod_batch_slots(self._h, ctypes.byref(e)))
        self.element_slots = e.value
        self.starts = []
        for r in range(len(self.n_elems)):
            s = ctypes.c_uint64()
            _check(L.erod_rod_start(self._h, r, ctypes.byref(s)))
            self.starts.append(s.value)
        self.damped = damped
        self.time = 0.0

    # -- lifetime --
    def close(self):
        if getattr(self, "_h", None):
            _LIB.erod_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n_rods(self):
        return len(self.n_elems)

    def slots(self, stagger):
        E = self.element_slots
        return E + 1 if stagger == NODE else (E if stagger == ELEMENT else E - 1)

    def rod_slots(self, tag, rod):
        _, _, st = FIELDS[tag]
        n = self.active_length(rod)
        return n + 1 if st == NODE else (n if st == ELEMENT else n - 1)

    # -- data movement --
    def get_field(self, tag):
        fid, nc, st = FIELDS[tag]
        out = np.empty((nc, self.slots(st)))
        _check(_LIB.erod_download_field(self._h, fid, _dp(out)))
        return out

    def set_field(self, tag, value):
        fid, nc, st = FIELDS[tag]
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(nc, self.slots(st))
        _check(_LIB.erod_upload_field(self._h, fid, _dp(a)))

    def get_rod(self, tag, rod):
        fid, nc, _ = FIELDS[tag]
        out = np.empty((nc, self.rod_slots(tag, rod)))
        _check(_LIB.erod_download_rod(self._h, fid, rod, _dp(out)))
        return out

    def set_rod(self, tag, rod, value):
        fid, nc, _ = FIELDS[tag]
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(nc, self.rod_slots(tag, rod))
        _check(_LIB.erod_upload_rod(self._h, fid, rod, _dp(a)))

    def rod(self, i):
        return RodSlice(self, i)

    def fill_ghosts(self):
        _check(_LIB.erod_fill_ghosts(self._h))

    # -- the three per-block stepper calls (op-by-op FP64) --
    def kinematic_update(self, dt):
        _check(_LIB.erod_kinematic_update(self._h, float(dt)))

    def kinematic_update_by(self, c):
        """x += c v, Q <- exp(-c [w]x) Q (kinematic_states() += c * kinematic_rates())."""
        _check(_LIB.erod_kinematic_update_by(self._h, float(c)))

    def fill_field(self, tag, value):
        """Every slot of a field, ghosts included, set to `value`."""
        _check(_LIB.erod_fill_field(self._h, FIELDS[tag][0], float(value)))

    def compute_internal_dynamics(self, time=0.0):
        _check(_LIB.erod_compute_internal_dynamics(self._h, float(time)))

    def dynamic_update(self, dt):
        _check(_LIB.erod_dynamic_update(self._h, float(dt)))

    # -- hooks --
    def add_clamp(self, rod, x, Q, index=0):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(3)
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(9)
        _check(_LIB.erod_add_clamp(self._h, rod, index, _dp(x), _dp(Q)))

    def add_end_force(self, rod, F, ramp=RAMP_NONE, start=0.0, interval=1.0):
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(3)
        _check(_LIB.erod_add_end_force(self._h, rod, _dp(F), int(ramp), float(start), float(interval)))

    def set_gravity(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(3)
        _check(_LIB.erod_set_gravity(self._h, _dp(g)))

    def set_plane_contact(self, k, c, mu, ct, z0=0.0, offset=0.0, normal=None):
        """Per-node contact with the plane normal . x = z0 (normal default (0, 0, 1))."""
        _check(_LIB.erod_set_plane_contact(self._h, k, c, mu, ct, z0, offset))
        if normal is not None:
            n = np.ascontiguousarray(normal, dtype=np.float64).reshape(3)
            _check(_LIB.erod_set_plane_normal(self._h, _dp(n)))

    def apply_post_kinematic(self, t):
        _check(_LIB.erod_apply_post_kinematic(self._h, float(t)))

    def apply_pre_dynamic(self, t):
        _check(_LIB.erod_apply_pre_dynamic(self._h, float(t)))

    def apply_post_dynamic(self, t):
        _check(_LIB.erod_apply_post_dynamic(self._h, float(t)))

    # -- whole steps --
    def step(self, dt, n=1):
        """n op-by-op position-Verlet steps (plain FP64 path)."""
        t = ctypes.c_double(self.time)
        _check(_LIB.erod_step(self._h, ctypes.byref(t), float(dt), int(n)))
        self.time = t.value

    def set_rod_dt(self, dt):
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n_rods,)))
        _check(_LIB.erod_set_rod_dt(self._h, _dp(dt)))

    def rod_time(self):
        out = np.empty(self.n_rods)
        _check(_LIB.erod_get_rod_time(self._h, _dp(out)))
        return out

    def set_rod_time(self, t):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.n_rods,)))
        _check(_LIB.erod_set_rod_time(self._h, _dp(t)))

    def set_fused_config(self, lanes_per_tile=0, elements_per_thread=0, steps_per_launch=0):
        _check(_LIB.erod_set_fused_config(self._h, int(lanes_per_tile), int(elements_per_thread),
                                          int(steps_per_launch)))

    def set_fused_split(self, lanes_per_tile=0, elements_per_thread=0):
        """Rods of at most `lanes_per_tile` elements that find no room beside
        longer rods run in a second group of tiles of that shape (0 = off)."""
        _check(_LIB.erod_set_fused_split(self._h, int(lanes_per_tile), int(elements_per_thread)))

    def fused_split(self):
        """(lanes, elements per thread, tiles) of the current plan's second group."""
        a, b, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        _check(_LIB.erod_get_fused_split(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t)))
        return a.value, b.value, t.value

    def set_fused_pieces(self, lanes_per_tile=0, elements_per_thread=0, halo=0):
        """Rods longer than the main tile (or than `lanes_per_tile`) run as
        pieces of that shape recomputing `halo` elements of each neighbour."""
        _check(_LIB.erod_set_fused_pieces(self._h, int(lanes_per_tile), int(elements_per_thread), int(halo)))

    def fused_pieces(self):
        """(lanes, elements per thread, halo, tiles, pieces) of the current plan's pieces."""
        a, b, h, t, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(_LIB.erod_get_fused_pieces(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(h), ctypes.byref(t),
                                          ctypes.byref(p)))
        return a.value, b.value, h.value, t.value, p.value

    def fused_config(self):
        """(lanes_per_tile, elements_per_thread, steps_per_launch) of the current plan."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_LIB.erod_get_fused_config(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def autotune(self, precision=FP32_BULK, trial_steps=20):
        """Pick the fastest tile shape on this batch; returns its ms per step."""
        ms = ctypes.c_double()
        _check(_LIB.erod_autotune_fused(self._h, int(precision), int(trial_steps), ctypes.byref(ms)))
        return ms.value

    def step_fused(self, n, precision=FP64):
        _check(_LIB.erod_step_fused(self._h, int(precision), int(n)))

    def step_fused_graph(self, n, chunk, precision=FP64):
        _check(_LIB.erod_step_fused_graph(self._h, int(precision), int(n), int(chunk)))

    def capture_fused_graph(self, chunk, precision=FP64):
        """Capture the graph of one chunk of fused steps now, without stepping."""
        _check(_LIB.erod_capture_fused_graph(self._h, int(precision), int(chunk)))

    def reorthonormalize(self):
        _check(_LIB.erod_reorthonormalize(self._h))

    def set_reortho_interval(self, every):
        """Re-orthonormalise the directors of the rods with steps left every
        `every` fused steps (erod_set_reortho_interval; 0: never)."""
        if int(every) < 0:
            raise ValueError("the interval is a number of steps, not %r" % (every,))
        _check(_LIB.erod_set_reortho_interval(self._h, int(every)))

    def reortho_interval(self):
        """(interval, steps since the last scheduled re-orthonormalisation)"""
        e, p = ctypes.c_uint64(), ctypes.c_uint64()
        _check(_LIB.erod_get_reortho_interval(self._h, ctypes.byref(e), ctypes.byref(p)))
        return e.value, p.value

    def last_kernel_ms(self):
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _check(_LIB.erod_last_kernel_ms(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        _check(_LIB.erod_synchronize(self._h))

    # -- checkpoints --
    def save(self, path):
        """Write the whole batch (fields, clocks, rod records, budgets, environment)."""
        _check(_LIB.erod_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """A new batch continuing from a checkpoint written by `save`."""
        from .checkpoint import read_header
        L = load_library()
        damped, n_elems = read_header(path)
        h = ctypes.c_void_p()
        _check(L.erod_load(os.fsencode(path), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.n_elems = n_elems
        self.damped = damped
        e = ctypes.c_uint64()
        _check(L.erod_batch_slots(self._h, ctypes.byref(e)))
        self.element_slots = e.value
        self.starts = []
        for r in range(len(n_elems)):
            s_ = ctypes.c_uint64()
            _check(L.erod_rod_start(self._h, r, ctypes.byref(s_)))
            self.starts.append(s_.value)
        self.time = float(self.rod_time()[0])  # the op-by-op clock
        return self

    # -- slots, budgets, rod copies --
    def set_active_length(self, rod, n):
        _check(_LIB.erod_set_rod_active_length(self._h, rod, n))

    def active_length(self, rod):
        n = ctypes.c_uint64()
        _check(_LIB.erod_get_rod_active_length(self._h, rod, ctypes.byref(n)))
        return n.value

    def set_step_budget(self, rod, steps):
        """Steps the fused path may still advance `rod` (None: unlimited)."""
        _check(_LIB.erod_set_rod_step_budget(self._h, rod, (1 << 64) - 1 if steps is None else int(steps)))

    def step_budget(self, rod):
        n = ctypes.c_uint64()
        _check(_LIB.erod_get_rod_step_budget(self._h, rod, ctypes.byref(n)))
        return None if n.value == (1 << 64) - 1 else n.value

    def copy_rods_from(self, src, dst_rods, src_rods):
        """Device copy of rods `src_rods` of `src` into this batch's slots `dst_rods`."""
        d = np.ascontiguousarray(dst_rods, dtype=np.uint64)
        s_ = np.ascontiguousarray(src_rods, dtype=np.uint64)
        assert d.shape == s_.shape
        _check(_LIB.erod_copy_rods(self._h, _u64p(d), src._h, _u64p(s_), len(d)))

    def check_finite(self):
        out = np.zeros(self.n_rods, dtype=np.uint8)
        _check(_LIB.erod_check_finite(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out


def straight_rod_state(n, radius, length, origin, direction, normal, density, youngs, shear,
                       force_damping, torque_damping):
    """Host-side straight-rod initializer (A14 of the reference:
    Initializers/detail/Geometry.hpp:110-214, Orthogonal.cpp:33-116,
    CosseratRodSpanwiseGeometry.hpp:525-545, CosseratRodCrossSectionInterface.hpp:266-307,
    CircleCrossSectionOperations.hpp:83-132, LinearHyperElasticityFacade.hpp:260-354,
    WithRodKinematics.hpp:330-436) returning a dict tag -> per-rod array, in
    the same operation order as the reference so the constants are
    bit-identical to what its initializer computes."""
    def norm(a):
        a = np.asarray(a, dtype=np.float64)
        ln = np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
        return a * (1.0 / ln)

    def cross(a, b):
        return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])

    d, nr = norm(direction), norm(normal)
    v1 = norm(nr)
    v2 = cross(d, nr)
    v3 = cross(v1, v2)
    d2 = norm(cross(v3, v1))
    d3 = norm(v3)
    rows = np.concatenate([v1, d2, d3])
    origin = np.asarray(origin, dtype=np.float64)
    end = origin + length * d
    step = (end - origin) / float(n)
    x = np.stack([origin[c] + np.arange(n + 1, dtype=np.float64) * step[c] for c in range(3)])
    Q = np.repeat(rows[:, None], n, axis=1)
    dx = x[:, 1:] - x[:, :-1]
    l0 = np.sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) / 1.0
    D0 = 0.5 * (l0[:-1] + l0[1:])
    l = np.sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2])
    r0 = np.full(n, float(radius))
    vol = (np.pi * (r0 * r0)) * l0
    r = np.sqrt(M_1_PI * vol / l)
    A = np.pi * (r * r)
    I = (0.25 * M_1_PI) * (A * A)
    moi = np.stack([I, I, 2.0 * I])
    EEG = np.array([youngs, youngs, shear])
    B = ((l0[None, :-1] * EEG[:, None]) * moi[:, :-1] + (l0[None, 1:] * EEG[:, None]) * moi[:, 1:]) / (l0[:-1] + l0[1:])
    alpha = 13.5 / 14.0
    S = np.stack([shear * (A * alpha), shear * (A * alpha), youngs * A])
    J = (moi * density) * l0
    iJ = 1.0 / ((moi * density) * l0)
    me = density * vol
    # m_k = (0 + me_{k-1}/2) + me_k/2, the accumulation order of the reference loop
    m = np.zeros(n + 1)
    m[1:] = 0.5 * me
    m[:-1] += 0.5 * me
    return {
        "Position": x, "Director": Q, "ReferenceElementLength": l0[None, :],
        "ReferenceVoronoiLength": D0[None, :], "ElementVolume": vol[None, :], "ElementDimension": r[None, :],
        "ElementLength": l[None, :], "ElementDilatation": np.ones((1, n)),
        "BendingTwistRigidityMatrix": B, "ShearStretchRigidityMatrix": S,
        "MassSecondMomentOfInertia": J, "InvMassSecondMomentOfInertia": iJ,
        "Mass": m[None, :], "InvMass": (1.0 / m)[None, :],
        "ForceDampingRate": np.full((1, n + 1), float(force_damping)),
        "TorqueDampingRate": np.full((1, n), float(torque_damping)),
        "Velocity": np.zeros((3, n + 1)), "AngularVelocity": np.zeros((3, n)),
        "ReferenceCurvature": np.zeros((3, n - 1)), "ReferenceShearStretchStrain": np.zeros((3, n)),
    }


class _PoolView(DeviceBlock):
    """Non-owning DeviceBlock view of a server's pool batch."""

    def __init__(self, handle, n_elems, damped):
        self._h = handle
        self.n_elems = list(n_elems)
        self.damped = damped
        self.time = 0.0
        e = ctypes.c_uint64()
        _check(_LIB.erod_batch_slots(self._h, ctypes.byref(e)))
        self.element_slots = e.value
        self.starts = []
        for r in range(len(self.n_elems)):
            s_ = ctypes.c_uint64()
            _check(_LIB.erod_rod_start(self._h, r, ctypes.byref(s_)))
            self.starts.append(s_.value)

    def close(self):
        self._h = None


class Server:
    """Continuous batching of requests over a pool of rod slots (erod_server_*).

    `classes` is a list of (slot capacity in elements, slot count): a fixed
    pool (erod_server_create).  `Server.budgeted(library, capacities, ...)`
    makes a server whose pool grows within its HBM budget instead
    (erod_server_create_budgeted).  A request is a list of rods of `library`
    (a DeviceBlock holding the template rods, with their dt set) and a number
    of steps."""

    def __init__(self, library, classes, precision=1, steps_per_tick=100, hbm_budget_bytes=200e9, _budgeted=False):
        L = load_library()
        h = ctypes.c_void_p()
        if _budgeted:
            caps = np.array([int(c) for c in classes], dtype=np.uint64)
            _check(L.erod_server_create_budgeted(library._h, _u64p(caps), len(caps), precision, steps_per_tick,
                                                 float(hbm_budget_bytes), ctypes.byref(h)))
        else:
            caps = np.array([c for c, _ in classes], dtype=np.uint64)
            cnts = np.array([k for _, k in classes], dtype=np.uint64)
            _check(L.erod_server_create(library._h, _u64p(caps), _u64p(cnts), len(classes), precision, steps_per_tick,
                                        float(hbm_budget_bytes), ctypes.byref(h)))
        self._h = h
        self.library = library
        self._pool = None
        self._done = np.zeros(4096, dtype=np.uint64)
        self._lib_n = [library.active_length(r) for r in range(len(library.n_elems))]
        self._sizes = {}

    @classmethod
    def budgeted(cls, library, capacities, precision=1, steps_per_tick=100, hbm_budget_bytes=200e9):
        """A server whose pool of slots (classes of the given capacities)
        grows on demand while it fits `hbm_budget_bytes`."""
        return cls(library, capacities, precision, steps_per_tick, hbm_budget_bytes, _budgeted=True)

    @property
    def pool(self):
        """A view of the current pool batch (a new one after the pool grew)."""
        p = ctypes.c_void_p()
        _check(_LIB.erod_server_pool(self._h, ctypes.byref(p)))
        if self._pool is None or self._pool._h is None or self._pool._h.value != p.value:
            if self._pool is not None:
                self._pool.close()
            n = ctypes.c_uint64()
            _check(_LIB.erod_batch_rods(p, ctypes.byref(n), None))
            caps = np.zeros(n.value, dtype=np.uint64)
            _check(_LIB.erod_batch_rods(p, ctypes.byref(n), _u64p(caps)))
            self._pool = _PoolView(p, [int(c) for c in caps], self.library.damped)
        return self._pool

    def pool_info(self):
        """{"slots", "growths", "hbm_budget_bytes"} of the pool."""
        n, g, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        _check(_LIB.erod_server_pool_info(self._h, ctypes.byref(n), ctypes.byref(g), ctypes.byref(b)))
        return {"slots": n.value, "growths": g.value, "hbm_budget_bytes": b.value}

    def close(self):
        if getattr(self, "_h", None):
            # the pool view goes first: its handle dies with the server
            if getattr(self, "_pool", None) is not None:
                self._pool.close()
                self._pool = None
            _LIB.erod_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, rid, library_rods, steps):
        r = np.ascontiguousarray(library_rods, dtype=np.uint64)
        _check(_LIB.erod_server_submit(self._h, rid, _u64p(r), len(r), int(steps)))
        self._sizes[rid] = [self._lib_n[int(x)] for x in r]

    def tick(self):
        """One tick; returns the ids of requests that finished."""
        n = ctypes.c_uint64()
        _check(_LIB.erod_server_tick(self._h, _u64p(self._done), len(self._done), ctypes.byref(n)))
        return [int(x) for x in self._done[:n.value]]

    def read(self, rid, rod, tag):
        """`tag` of rod `rod` (index within the request) of a finished request."""
        fid, nc, st = FIELDS[tag]
        out = np.empty((nc, self._count(rid, rod, st)))
        _check(_LIB.erod_server_read(self._h, rid, rod, fid, _dp(out)))
        return out

    def _count(self, rid, rod, st):
        n = self._sizes[rid][rod]
        return n + 1 if st == NODE else (n if st == ELEMENT else n - 1)

    def release(self, rid):
        _check(_LIB.erod_server_release(self._h, rid))
        self._sizes.pop(rid, None)

    def status(self):
        q, r, e, b = (ctypes.c_uint64() for _ in range(4))
        _check(_LIB.erod_server_status(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(e), ctypes.byref(b)))
        return {"queued": q.value, "running": r.value, "active_elements": e.value, "pool_bytes": b.value}

    def stats(self):
        """(ticks launched, element-steps advanced) since creation."""
        t, e = ctypes.c_uint64(), ctypes.c_uint64()
        _check(_LIB.erod_server_stats(self._h, ctypes.byref(t), ctypes.byref(e)))
        return t.value, e.value


def assemble(block, rod_states):
    """Upload a list of per-rod state dicts (tag -> (ncomp, count) array) into
    a DeviceBlock field by field, ghosts left at their ghost values.  A
    container with a `fill_block(block)` method writes itself."""
    if hasattr(rod_states, "fill_block"):
        rod_states.fill_block(block)
        return
    tags = set()
    for st in rod_states:
        tags.update(st.keys())
    for tag in sorted(tags):
        full = block.get_field(tag)
        for r, st in enumerate(rod_states):
            if tag in st:
                s = block.starts[r]
                a = np.asarray(st[tag], dtype=np.float64)
                full[:, s:s + a.shape[1]] = a
        block.set_field(tag, full)

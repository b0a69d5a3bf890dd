"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes driver for the C++20 CPU restatement of Elastica++'s FP64 Cosserat-rod
timestep (cosserat_oracle.hpp).  Only the test suite, the smoke check and the
benchmark's CPU-baseline leg use this module; the product package never
imports it.

A System is a list of ghost-padded blocks (the reference's
blocks::Block<CosseratRod>, elastica/Systems/CosseratRods/Block.hpp) plus hook
descriptors (clamp / end force / gravity / plane contact).  Fields are exposed
as numpy views of shape (ncomp, slots) on a block, or per rod as the slice the
reference's blocks::get<tag>(rod) would return.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NODE_FIELDS = ("Position", "Velocity", "Acceleration", "Mass", "InvMass",
               "InternalLoads", "ExternalLoads", "ForceDampingRate")
VORONOI_FIELDS = ("ReferenceVoronoiLength", "VoronoiLength", "VoronoiDilatation",
                  "Curvature", "ReferenceCurvature", "BendingTwistRigidityMatrix",
                  "InternalCouple", "_DummyVoronoiVector")


# builds of the restatement (oracle/Makefile): the C library's sin / cos /
# acos (the reference's), the device library's, and the C library's with a
# quarter of the results moved by one ulp
VARIANTS = {"libm": "liboracle.so", "devmath": "liboracle_devmath.so", "perturbed": "liboracle_perturbed.so"}
_LOADED = {}


def use_library(path=None):
    """Load the oracle from `path` (default: oracle/liboracle.so) for every
    System created afterwards without a `variant`; the mutation tests load
    deliberately broken builds this way."""
    global _LIB
    _LIB = None
    return _lib(path)


def variant_library(variant):
    """The build of the restatement named in VARIANTS."""
    path = os.path.join(_HERE, VARIANTS[variant])
    if path not in _LOADED:
        _LOADED[path] = _load(path)
    return _LOADED[path]


def _lib(path=None):
    global _LIB
    if _LIB is None:
        _LIB = _load(path or os.path.join(_HERE, "liboracle.so"))
    return _LIB


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError("%s is not built: run `make -C oracle`" % os.path.relpath(path, _HERE))
    L = ctypes.CDLL(path)
    u64, vp, dp, cp = ctypes.c_uint64, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p
    L.oracle_system_create.restype = vp
    L.oracle_system_create.argtypes = [ctypes.POINTER(u64), u64, u64, ctypes.c_int]
    L.oracle_system_destroy.argtypes = [vp]
    L.oracle_system_n_blocks.restype = u64
    L.oracle_system_n_blocks.argtypes = [vp]
    L.oracle_rod_location.argtypes = [vp, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.oracle_block_field.restype = dp
    L.oracle_block_field.argtypes = [vp, u64, cp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.oracle_block_fill_ghosts.argtypes = [vp, u64]
    L.oracle_initialize_straight_rod.argtypes = [vp, u64, dp]
    for name in ("oracle_compute_internal_forces", "oracle_compute_internal_torques",
                 "oracle_update_acceleration"):
        getattr(L, name).argtypes = [vp]
    L.oracle_kinematic_update.argtypes = [vp, ctypes.c_double]
    L.oracle_add_clamp.argtypes = [vp, u64, u64, dp, dp]
    L.oracle_add_end_force.argtypes = [vp, u64, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.oracle_set_gravity.argtypes = [vp, dp]
    L.oracle_set_plane_contact.argtypes = [vp] + [ctypes.c_double] * 6
    L.oracle_set_plane_normal.argtypes = [vp, dp]
    for name in ("oracle_fp_flags", "oracle_fe_invalid", "oracle_fe_divbyzero", "oracle_fe_overflow"):
        getattr(L, name).restype = ctypes.c_int
    L.oracle_fp_flags.argtypes = [ctypes.c_int]
    L.oracle_ramp_value.restype = ctypes.c_double
    L.oracle_ramp_value.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.oracle_time.restype = ctypes.c_double
    L.oracle_time.argtypes = [vp]
    L.oracle_set_time.argtypes = [vp, ctypes.c_double]
    L.oracle_step.argtypes = [vp, ctypes.c_double, u64, ctypes.c_int]
    L.oracle_step_per_block_dt.argtypes = [vp, dp, u64, ctypes.c_int]
    L.oracle_two_point_difference.argtypes = [dp, u64, dp]
    L.oracle_quadrature.argtypes = [dp, u64, dp]
    L.oracle_inv_rotate_divide.argtypes = [dp, u64, dp, dp]
    L.oracle_so3_add_assign.argtypes = [dp, u64, dp]
    return L


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def stagger_of(name):
    if name in NODE_FIELDS:
        return "node"
    if name in VORONOI_FIELDS:
        return "voronoi"
    return "element"


def rod_slots(name, n_elem):
    s = stagger_of(name)
    return n_elem + 1 if s == "node" else (n_elem - 1 if s == "voronoi" else n_elem)


class System:
    """Rods packed first-fit into ghost-padded blocks of <= max_block_elems
    elements (0: one block per rod)."""

    def __init__(self, n_elems, max_block_elems=1024, damped=True, variant=None):
        # a System keeps the library it was created with
        L = self._L = _lib() if variant is None else variant_library(variant)
        self.n_elems = [int(n) for n in n_elems]
        arr = (ctypes.c_uint64 * len(self.n_elems))(*self.n_elems)
        self._h = L.oracle_system_create(arr, len(self.n_elems), int(max_block_elems), 1 if damped else 0)
        if not self._h:
            raise ValueError("oracle_system_create failed")

    def close(self):
        if self._h:
            self._L.oracle_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_blocks(self):
        return int(self._L.oracle_system_n_blocks(self._h))

    def location(self, rod):
        b, s = ctypes.c_uint64(), ctypes.c_uint64()
        self._L.oracle_rod_location(self._h, rod, ctypes.byref(b), ctypes.byref(s))
        return b.value, s.value

    def block_field(self, block, name):
        nc, ns = ctypes.c_uint64(), ctypes.c_uint64()
        p = self._L.oracle_block_field(self._h, block, name.encode(), ctypes.byref(nc), ctypes.byref(ns))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(nc.value, ns.value))

    def rod_field(self, rod, name):
        """View of one rod's slice: shape (ncomp, slots_of_stagger)."""
        b, s = self.location(rod)
        full = self.block_field(b, name)
        return full[:, s:s + rod_slots(name, self.n_elems[rod])]

    def fill_ghosts(self):
        for b in range(self.n_blocks):
            self._L.oracle_block_fill_ghosts(self._h, b)

    def init_straight_rod(self, rod, radius, length, origin, direction, normal,
                          density, youngs, shear, force_damping, torque_damping):
        p = np.array([radius, length, *origin, *direction, *normal, density, youngs, shear,
                      force_damping, torque_damping], dtype=np.float64)
        self._L.oracle_initialize_straight_rod(self._h, rod, _dptr(p))

    def compute_internal_forces(self):
        self._L.oracle_compute_internal_forces(self._h)

    def compute_internal_torques(self):
        self._L.oracle_compute_internal_torques(self._h)

    def update_acceleration(self):
        self._L.oracle_update_acceleration(self._h)

    def kinematic_update(self, dt):
        self._L.oracle_kinematic_update(self._h, dt)

    def add_clamp(self, rod, x, Q, index=0):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(3)
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(9)
        if self._L.oracle_add_clamp(self._h, rod, index, _dptr(x), _dptr(Q)) != 0:
            raise IndexError("clamp at index %d of rod %d of a system of %d rods (index past the rod's last node, "
                             "or no such rod)" % (index, rod, len(self.n_elems)))

    def add_end_force(self, rod, F, ramp=0, start=0.0, interval=1.0):
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(3)
        if self._L.oracle_add_end_force(self._h, rod, _dptr(F), int(ramp), float(start), float(interval)) != 0:
            raise IndexError("end force on rod %d of a system of %d rods" % (rod, len(self.n_elems)))

    def set_gravity(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(3)
        self._L.oracle_set_gravity(self._h, _dptr(g))

    def set_plane_contact(self, k, c, mu, ct, z0=0.0, offset=0.0, normal=None):
        """Contact with the plane normal . x = z0 (normal default (0, 0, 1))."""
        self._L.oracle_set_plane_contact(self._h, k, c, mu, ct, z0, offset)
        if normal is not None:
            n = np.ascontiguousarray(normal, dtype=np.float64).reshape(3)
            if self._L.oracle_set_plane_normal(self._h, _dptr(n)) != 0:
                raise ValueError("plane normal must be finite and non-zero")

    @property
    def time(self):
        return self._L.oracle_time(self._h)

    @time.setter
    def time(self, t):
        self._L.oracle_set_time(self._h, float(t))

    def step(self, dt, n=1, threads=1):
        self._L.oracle_step(self._h, float(dt), int(n), int(threads))

    def step_per_block_dt(self, block_dt, n=1, threads=1):
        block_dt = np.ascontiguousarray(block_dt, dtype=np.float64)
        assert block_dt.size == self.n_blocks
        rc = self._L.oracle_step_per_block_dt(self._h, _dptr(block_dt), int(n), int(threads))
        if rc != 0:
            raise ValueError("per-block dt stepping supports constant forces only")


def two_point_difference(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros((3, a.shape[1] + 1))
    _lib().oracle_two_point_difference(_dptr(a), a.shape[1], _dptr(out))
    return out


def quadrature(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros((3, a.shape[1] + 1))
    _lib().oracle_quadrature(_dptr(a), a.shape[1], _dptr(out))
    return out


def inv_rotate_divide(Q, D0):
    """Q: (3, 3, n) tensor (page, row, element); D0: (n-1,) -> kappa (3, n-1)."""
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    D0 = np.ascontiguousarray(D0, dtype=np.float64)
    n = Q.shape[2]
    out = np.zeros((3, n - 1))
    _lib().oracle_inv_rotate_divide(_dptr(Q), n, _dptr(D0), _dptr(out))
    return out


def so3_add_assign(Q, u):
    """In-place Q <- exp(-[u]x) Q per element; Q (3,3,n), u (3,n)."""
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    _lib().oracle_so3_add_assign(_dptr(Q), Q.shape[2], _dptr(u))
    return Q


def ramp_value(kind, t, start, interval):
    """The ramp factor of an end force at time t (0 none, 1 linear, 2 sin)."""
    return _lib().oracle_ramp_value(int(kind), float(t), float(start), float(interval))


def fp_exceptions(clear=True):
    """Floating-point exceptions (invalid, divide-by-zero, overflow) raised by
    the oracle's steps and force evaluations on the calling thread since the
    last read, by name."""
    L = _lib()
    f = L.oracle_fp_flags(1 if clear else 0)
    names = (("invalid", L.oracle_fe_invalid()), ("divide-by-zero", L.oracle_fe_divbyzero()),
             ("overflow", L.oracle_fe_overflow()))
    return [n for n, bit in names if f & bit]

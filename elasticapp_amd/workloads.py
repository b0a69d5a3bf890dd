"""Seeded synthetic rod workloads (host side, FP64).

Each builder returns a Workload: per-rod element counts, per-rod state dicts
(tag -> (ncomp, count) arrays, ready for `assemble`), a timestep, and hook
descriptors.  The same arrays feed the GPU paths and the CPU oracle, so the
random generator does not affect parity.

  timoshenko_cantilever  1 rod x 100 elements, clamped, tip force (0, -15, 0)
  uniform_1024x64        1024 rods x 64 elements, random rest curvature and velocity
  ragged_4096            4096 rods, n log-uniform in [16, 1024], random materials,
                         constant tip forces
  rods_on_plane          100,000 rods x 128 elements lying above the plane z = 0
                         under gravity (contact)
  serving_library        template rods for the serving workload, and
  serving_requests       requests of 1-64 template rods with horizons of 1e3-1e5 steps

Material conventions: G = E / (1 + 0.99) (a Poisson ratio of 0.99 as stated
for the Timoshenko configuration); damping rates for forces and torques are
equal.  The reference's timestep rule dt = 0.01 L / n
(examples/timoshenko/TimoshenkoCaseStudy.hpp:136-143) is tuned for the axial
wave speed sqrt(E / rho) = sqrt(1e6 / 5e3); for other materials the same
Courant number is kept: dt_i = 0.01 (L_i / n_i) sqrt(1e6 / 5e3) / sqrt(E_i / rho_i),
and a batch uses the smallest dt of its rods (the reference steps every
system with one dt).  The ragged batch also carries each rod's own dt
(`info["rod_dt"]`) for the fused paths, which step every rod with its own.

A rod stepped with its own dt (the serving library) must also respect two
limits of the explicit step that the axial rule does not see when the radius
r is small against L / n, taken at a tenth of their critical values:
  shear-rotation mode   omega = (2 / r) sqrt(G alpha / rho), alpha = 27/28:
                        dt < 2 / omega = r sqrt(rho / (G alpha))
  torque damping        d(w)/dt = -gamma (m / J) w with m / J = 4 / r^2:
                        dt < 2 / (4 gamma / r^2) = r^2 / (2 gamma)
"""
from dataclasses import dataclass, field

import numpy as np

from . import DeviceBlock, assemble, straight_rod_state

C_REF = np.sqrt(1e6 / 5e3)


@dataclass
class Workload:
    name: str
    n_elems: list
    states: list
    dt: float
    steps: int
    clamps: list = field(default_factory=list)   # (rod, x[3], Q[9]) or (rod, x[3], Q[9], index)
    forces: list = field(default_factory=list)   # (rod, F[3], ramp, start, interval)
    gravity: tuple = None
    contact: tuple = None                        # (k, c, mu, ct, z0, offset[, normal[3]])
    info: dict = field(default_factory=dict)

    @property
    def total_elements(self):
        return int(sum(self.n_elems))


def stable_dt(length, n, youngs, density, radius=None, shear=None, damping=None):
    """The axial Courant rule; with radius, shear modulus and torque damping
    rate also a tenth of the shear-rotation and damping limits."""
    dt = 0.01 * (np.asarray(length) / n) * C_REF / np.sqrt(np.asarray(youngs) / density)
    if radius is not None:
        r = np.asarray(radius)
        rot = r * np.sqrt(np.asarray(density) / (np.asarray(shear) * (27.0 / 28.0)))
        damp = r * r / (2.0 * np.maximum(np.asarray(damping), 1e-300))
        dt = np.minimum(dt, 0.1 * np.minimum(rot, damp))
    return dt


def timoshenko_cantilever(steps=50000):
    n, L, r, rho, E = 100, 3.0, 0.25, 5000.0, 1e6
    G = E / (1.0 + 0.99)
    st = straight_rod_state(n, r, L, [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], rho, E, G, 0.1, 0.1)
    dt = 0.01 * (L / n)
    clamps = [(0, st["Position"][:, 0].copy(), st["Director"][:, 0].copy())]
    forces = [(0, np.array([0.0, -15.0, 0.0]), 0, 0.0, 1.0)]
    return Workload("timoshenko_cantilever", [n], [st], dt, steps, clamps=clamps, forces=forces,
                    info=dict(length=L, radius=r, density=rho, youngs=E, shear=G, damping=0.1))


def _randomize(st, rng, kappa_std=0.5, v_std=1e-2):
    n = st["Director"].shape[1]
    st["ReferenceCurvature"] = rng.normal(0.0, kappa_std, (3, n - 1))
    st["Velocity"] = rng.normal(0.0, v_std, (3, n + 1))
    return st


def uniform_1024x64(n_rods=1024, n=64, steps=10000, seed=1001):
    rng = np.random.default_rng(seed)
    L, r, rho, E = 3.0, 0.25, 5000.0, 1e6
    G = E / (1.0 + 0.99)
    states = []
    for i in range(n_rods):
        st = straight_rod_state(n, r, L, [0.0, 1.0 * i, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], rho, E, G, 0.1, 0.1)
        states.append(_randomize(st, rng))
    dt = 0.01 * (L / n)
    return Workload("uniform_1024x64", [n] * n_rods, states, dt, steps,
                    info=dict(length=L, radius=r, density=rho, youngs=E, shear=G, damping=0.1))


def ragged_4096(n_rods=4096, steps=10000, seed=2002, n_min=16, n_max=1024):
    rng = np.random.default_rng(seed)
    n_elems = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_rods))).astype(int)
    n_elems = np.clip(n_elems, n_min, n_max)
    L = rng.uniform(0.5, 3.0, n_rods)
    r = rng.uniform(0.005, 0.05, n_rods)
    E = np.exp(rng.uniform(np.log(1e5), np.log(1e7), n_rods))
    rho = rng.uniform(500.0, 5000.0, n_rods)
    gam = rng.uniform(0.01, 0.5, n_rods)
    F = rng.normal(0.0, 1.0, (n_rods, 3))
    states, forces = [], []
    for i in range(n_rods):
        st = straight_rod_state(int(n_elems[i]), r[i], L[i], [0.0, 0.2 * i, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0],
                                rho[i], E[i], E[i] / (1.0 + 0.99), gam[i], gam[i])
        states.append(_randomize(st, rng))
        forces.append((i, F[i].copy(), 0, 0.0, 1.0))
    dts = stable_dt(L, n_elems, E, rho)
    rod_dt = stable_dt(L, n_elems, E, rho, r, E / (1.0 + 0.99), gam)
    return Workload("ragged_4096", [int(x) for x in n_elems], states, float(dts.min()), steps, forces=forces,
                    info=dict(rod_dt_min=float(dts.min()), rod_dt_max=float(dts.max()), rod_dt=rod_dt))


class UniformRods:
    """The per-rod states of equal rods built from one base rod plus a
    per-rod position offset and per-rod rest curvature, made on demand.

    Indexing gives the same per-rod dicts a list of states would (for the
    oracle at test sizes); `fill_block` writes every rod into a DeviceBlock
    with one whole-field upload per tag (for 1e5 rods)."""

    def __init__(self, base, offsets, kappa0):
        self.base = base
        self.offsets = np.asarray(offsets, dtype=np.float64)    # (n_rods, 3)
        self.kappa0 = np.asarray(kappa0, dtype=np.float64)      # (n_rods, 3, n - 1)

    def __len__(self):
        return len(self.offsets)

    def _rod(self, i):
        st = {k: v.copy() for k, v in self.base.items()}
        st["Position"] = self.base["Position"] + self.offsets[i][:, None]
        st["ReferenceCurvature"] = self.kappa0[i].copy()
        return st

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self._rod(j) for j in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        return self._rod(i)

    def __iter__(self):
        return (self._rod(i) for i in range(len(self)))

    def rod_array(self, tag):
        """(n_rods, ncomp, count) values of `tag` over all rods."""
        if tag == "Position":
            return self.base["Position"][None] + self.offsets[:, :, None]
        if tag == "ReferenceCurvature":
            return self.kappa0
        b = self.base[tag]
        return np.broadcast_to(b, (len(self),) + b.shape)

    def fill_block(self, block):
        # equal rods start every n + 2 slots: view each field as
        # (ncomp, rods, n + 2) and write the rods' leading slots in one go
        n = self.base["Director"].shape[1]
        period = n + 2
        rods = len(self)
        starts = np.asarray(block.starts, dtype=np.int64)
        if not np.array_equal(starts, period * np.arange(rods)):
            raise ValueError("UniformRods.fill_block: block does not hold these rods in order")
        for tag in sorted(self.base):
            full = block.get_field(tag)
            arr = self.rod_array(tag)
            nc, count = arr.shape[1], arr.shape[2]
            buf = np.empty((nc, rods * period))
            buf[:, :full.shape[1]] = full
            buf.reshape(nc, rods, period)[:, :, :count] = arr.transpose(1, 0, 2)
            block.set_field(tag, buf[:, :full.shape[1]])


def rods_on_plane(n_rods=100000, n=128, steps=2000, seed=4004, k=1e3, c=10.0, mu=0.3, ct=10.0):
    """Rods of length 1 m and radius 0.01 m (rho 1000, E 1e6, damping 0.1)
    lying along x at z = radius, 5 cm apart in y, so that they touch the plane
    z = 0 (contact surface z0 + offset with offset = radius); rest curvature
    N(0, 0.5) per component, gravity (0, 0, -9.81), per-node plane contact
    k = 1e3 N/m, c = 10 N s/m, mu = 0.3, tangential damping ct = 10 N s/m for
    the regularised friction.  dt = 3.5e-5 s by the stable-dt rule; with a
    node mass of 2.45e-3 kg both dampers stay explicit-stable and non-stiff
    (c dt / m = ct dt / m = 0.14)."""
    rng = np.random.default_rng(seed)
    L, r, rho, E = 1.0, 0.01, 1000.0, 1e6
    G = E / (1.0 + 0.99)
    base = straight_rod_state(n, r, L, [0.0, 0.0, r], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], rho, E, G, 0.1, 0.1)
    offsets = np.zeros((n_rods, 3))
    offsets[:, 1] = 0.05 * np.arange(n_rods)
    states = UniformRods(base, offsets, rng.normal(0.0, 0.5, (n_rods, 3, n - 1)))
    dt = stable_dt(L, n, E, rho)
    return Workload("rods_on_plane", [n] * n_rods, states, float(dt), steps, gravity=(0.0, 0.0, -9.81),
                    contact=(k, c, mu, ct, 0.0, r), info=dict(length=L, radius=r, density=rho, youngs=E))


def serving_library(n_templates=2048, seed=3003, n_min=32, n_max=512):
    """Template rods for the serving workload: n log-uniform in [32, 512],
    materials and perturbations as in the ragged batch, a constant tip force
    N(0, 1) N, and each rod's own stable dt (returned per rod; the axial rule
    and the rotational limits)."""
    rng = np.random.default_rng(seed)
    n_elems = np.clip(np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_templates))).astype(int), n_min, n_max)
    L = rng.uniform(0.5, 3.0, n_templates)
    r = rng.uniform(0.005, 0.05, n_templates)
    E = np.exp(rng.uniform(np.log(1e5), np.log(1e7), n_templates))
    rho = rng.uniform(500.0, 5000.0, n_templates)
    gam = rng.uniform(0.01, 0.5, n_templates)
    F = rng.normal(0.0, 1.0, (n_templates, 3))
    states, forces = [], []
    for i in range(n_templates):
        st = straight_rod_state(int(n_elems[i]), r[i], L[i], [0.0, 0.2 * i, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0],
                                rho[i], E[i], E[i] / (1.0 + 0.99), gam[i], gam[i])
        states.append(_randomize(st, rng))
        forces.append((i, F[i].copy(), 0, 0.0, 1.0))
    dts = stable_dt(L, n_elems, E, rho, r, E / (1.0 + 0.99), gam)
    wl = Workload("serving_library", [int(x) for x in n_elems], states, float(dts.min()), 0, forces=forces)
    return wl, dts


def serving_requests(n_requests, n_templates, seed=3004, max_rods=64, min_steps=1000, max_steps=100000):
    """Requests of U{1..64} template rods (drawn with replacement) and a
    horizon log-uniform in [1e3, 1e5] steps."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_requests):
        k = int(rng.integers(1, max_rods + 1))
        rods = rng.integers(0, n_templates, k)
        steps = int(np.rint(np.exp(rng.uniform(np.log(min_steps), np.log(max_steps)))))
        out.append((rods, steps))
    return out


def apply_hooks(target, wl):
    """Register the workload's hooks on a DeviceBlock or an oracle System."""
    for rod, x, Q, *index in wl.clamps:
        target.add_clamp(rod, x, Q, index[0] if index else 0)
    for rod, F, ramp, start, interval in wl.forces:
        target.add_end_force(rod, F, ramp, start, interval)
    if wl.gravity is not None:
        target.set_gravity(wl.gravity)
    if wl.contact is not None:
        target.set_plane_contact(*wl.contact)


def build_block(wl, damped=True):
    """A DeviceBlock holding the workload's rods, hooks and (uniform) dt."""
    blk = DeviceBlock(wl.n_elems, damped=damped)
    assemble(blk, wl.states)
    apply_hooks(blk, wl)
    blk.set_rod_dt(wl.dt)
    return blk

"""Host-side workload builders (no GPU): the whole-field writer of equal-rod
batches puts every rod's values at the slots the per-rod path uses, and the
seeded builders are reproducible."""
import numpy as np

import elasticapp_amd as ea
from elasticapp_amd.workloads import UniformRods, rods_on_plane, ragged_4096


class FakeBlock:
    """Field store with the ghost-padded slot rule s_{i+1} = s_i + n_i + 2."""

    def __init__(self, n_elems):
        ea.load_library()
        self.n_elems = list(n_elems)
        self.starts = list(np.concatenate([[0], np.cumsum(np.asarray(n_elems) + 2)[:-1]]))
        self.E = int(sum(n_elems) + 2 * len(n_elems) - 2)
        self.fields = {}

    def _slots(self, st):
        return self.E + 1 if st == 0 else (self.E if st == 1 else self.E - 1)

    def get_field(self, tag):
        if tag not in self.fields:
            _, nc, st = ea.FIELDS[tag]
            self.fields[tag] = np.full((nc, self._slots(st)), -7.0)  # stand-in ghost value
        return self.fields[tag].copy()

    def set_field(self, tag, a):
        assert a.shape == self.get_field(tag).shape
        self.fields[tag] = np.array(a)


def test_fill_block_matches_per_rod_assembly():
    wl = rods_on_plane(n_rods=7, n=12)
    assert isinstance(wl.states, UniformRods)
    a, b = FakeBlock(wl.n_elems), FakeBlock(wl.n_elems)
    ea.assemble(a, wl.states)            # whole-field path
    ea.assemble(b, list(wl.states))      # per-rod path
    assert a.fields.keys() == b.fields.keys()
    for tag in a.fields:
        assert np.array_equal(a.fields[tag], b.fields[tag]), tag
    # ghosts untouched, rods where the slot rule puts them
    x = a.fields["Position"]
    assert x[0, wl.n_elems[0] + 1] == -7.0
    assert np.array_equal(x[:, a.starts[3]:a.starts[3] + 13], wl.states[3]["Position"])


def test_contact_c4_geometry_and_dt():
    wl = rods_on_plane(n_rods=5, n=128)
    st = wl.states[4]
    assert np.allclose(st["Position"][2], 0.01) and np.allclose(st["Position"][1], 0.2)
    assert np.isclose(st["ReferenceElementLength"].sum(), 1.0)
    # 0.01 (L / n) sqrt(1e6 / 5e3) / sqrt(E / rho) with L = 1, n = 128, E = 1e6, rho = 1e3
    assert np.isclose(wl.dt, 0.01 / 128 * np.sqrt(200.0) / np.sqrt(1e3))
    assert wl.contact == (1e3, 10.0, 0.3, 10.0, 0.0, 0.01) and wl.gravity == (0.0, 0.0, -9.81)
    # same seed, same rods; rods differ in their rest curvature
    again = rods_on_plane(n_rods=5, n=128)
    assert np.array_equal(again.states[2]["ReferenceCurvature"], wl.states[2]["ReferenceCurvature"])
    assert not np.array_equal(wl.states[1]["ReferenceCurvature"], wl.states[2]["ReferenceCurvature"])


def test_ragged_c2_is_seeded():
    a, b = ragged_4096(n_rods=64), ragged_4096(n_rods=64)
    assert a.n_elems == b.n_elems and a.dt == b.dt
    assert np.array_equal(a.states[10]["Velocity"], b.states[10]["Velocity"])
    assert min(a.n_elems) >= 16 and max(a.n_elems) <= 1024


def test_per_rod_dt_respects_the_rotational_limits():
    """The per-rod rule keeps dt * omega below 0.2 for the shear-rotation mode,
    omega = (2 / r) sqrt(G alpha / rho), and dt * gamma * 4 / r^2 below 0.2
    for the torque damping; for the cantilever's rod (r > L / n) it is the
    axial rule alone."""
    from elasticapp_amd.workloads import serving_library, stable_dt
    L, n, E, rho, r, gam = 1.9, 17, 1.26e5, 2256.0, 0.0077, 0.3
    G = E / 1.99
    dt = float(stable_dt(L, n, E, rho, r, G, gam))
    assert dt < float(stable_dt(L, n, E, rho))
    omega = 2.0 / r * np.sqrt(G * 27.0 / 28.0 / rho)
    assert dt * omega <= 0.2 * (1 + 1e-12) and dt * gam * 4.0 / r ** 2 <= 0.2 * (1 + 1e-12)
    assert float(stable_dt(3.0, 100, 1e6, 5e3, 0.25, 1e6 / 1.99, 0.1)) == float(stable_dt(3.0, 100, 1e6, 5e3))
    wl, dts = serving_library(n_templates=64)
    assert dts.shape == (64,) and (dts > 0).all() and float(dts.min()) == wl.dt

"""Shared builders for the GPU tests: the same per-rod arrays go into a
DeviceBlock (product) and an oracle System (checker)."""
import numpy as np

from elasticapp_amd import straight_rod_state
from elasticapp_amd.workloads import Workload, apply_hooks, build_block
from oracle import pyoracle

STATE = ("Position", "Velocity", "Director", "AngularVelocity")


def _turn_directors(d, R):
    # director lines 3 p + r hold Q(p, r); each director row d_p turns to R d_p
    return (d.T.reshape(-1, 3, 3) @ R.T).reshape(-1, 9).T.copy()


def rotated(wl, R):
    """The workload turned by the rotation R: world-frame quantities
    (positions, velocities, director rows, gravity, end forces, clamp poses,
    the contact plane's normal) rotate; material-frame ones (angular
    velocities, rest strains and curvatures, constants) do not."""
    states = []
    for st in wl.states:
        s = dict(st)
        s["Position"] = R @ st["Position"]
        s["Velocity"] = R @ st["Velocity"]
        s["Director"] = _turn_directors(st["Director"], R)
        states.append(s)
    clamps = [(r, R @ x, _turn_directors(np.asarray(Q).reshape(9, 1), R)[:, 0], *i) for r, x, Q, *i in wl.clamps]
    forces = [(r, R @ np.asarray(F), ramp, t0, T) for r, F, ramp, t0, T in wl.forces]
    gravity = None if wl.gravity is None else tuple(R @ np.asarray(wl.gravity))
    contact = None
    if wl.contact is not None:
        n = np.asarray(wl.contact[6]) if len(wl.contact) > 6 else np.array([0.0, 0.0, 1.0])
        contact = tuple(wl.contact[:6]) + (tuple(R @ n),)
    return Workload(wl.name + "-rotated", list(wl.n_elems), states, wl.dt, wl.steps, clamps=clamps, forces=forces,
                    gravity=gravity, contact=contact, info=dict(wl.info))


def random_workload(n_rods, n_min, n_max, seed, steps=100, forces=True, kappa=0.5, vstd=1e-2, lengths=None):
    """Randomly bent, moving, loaded rods; `lengths` fixes the element counts
    (n_rods, n_min and n_max are then unused)."""
    rng = np.random.default_rng(seed)
    n_elems = rng.integers(n_min, n_max + 1, n_rods) if lengths is None else np.array(lengths)
    n_rods = len(n_elems)
    states, fl = [], []
    dts = []
    for i in range(n_rods):
        n = int(n_elems[i])
        L = rng.uniform(0.3, 2.0)
        r = rng.uniform(0.005, 0.04)
        E = np.exp(rng.uniform(np.log(1e5), np.log(1e7)))
        rho = rng.uniform(500.0, 5000.0)
        g = rng.uniform(0.01, 0.5)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        nrm = np.cross(d, rng.normal(size=3))
        nrm /= np.linalg.norm(nrm)
        st = straight_rod_state(n, r, L, rng.normal(size=3), d, nrm, rho, E, E / 1.5, g, g)
        st["ReferenceCurvature"] = rng.normal(0.0, kappa, (3, n - 1))
        st["ReferenceShearStretchStrain"] = rng.normal(0.0, 1e-3, (3, n))
        st["Velocity"] = rng.normal(0.0, vstd, (3, n + 1))
        st["AngularVelocity"] = rng.normal(0.0, 10 * vstd, (3, n))
        states.append(st)
        if forces:
            fl.append((i, rng.normal(0.0, 0.1, 3), 0, 0.0, 1.0))
        dts.append(0.01 * (L / n) * np.sqrt(1e6 / 5e3) / np.sqrt(E / rho))
    return Workload("random%d" % seed, [int(x) for x in n_elems], states, float(min(dts)), steps, forces=fl)


def device_block(wl, damped=True):
    return build_block(wl, damped=damped)


def oracle_system(wl, damped=True, max_block_elems=0, variant=None):
    """The workload in an oracle System; `variant` names a build of the
    restatement (pyoracle.VARIANTS; default: the C library's functions)."""
    s = pyoracle.System(wl.n_elems, max_block_elems=max_block_elems, damped=damped, variant=variant)
    for r, st in enumerate(wl.states):
        for tag, a in st.items():
            s.rod_field(r, tag)[...] = a
    apply_hooks(s, wl)
    return s


def rel_state_error(blk, sysm, tags=STATE, rods=None):
    """max over rods and tags of |gpu - cpu| / max|cpu| (per tag and rod)."""
    worst = {}
    for r in (range(len(blk.n_elems)) if rods is None else rods):
        for t in tags:
            g = blk.get_rod(t, r)
            c = sysm.rod_field(r, t)
            scale = max(np.abs(c).max(), 1e-300)
            e = np.abs(g - c).max() / scale
            worst[t] = max(worst.get(t, 0.0), e)
    return worst

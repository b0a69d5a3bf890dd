"""The reference-side binding (include/erod_elastica.hpp) driven by a
restatement of the reference's stepper call sites (tests/stepper_surface.cpp:
PositionVerlet.hpp:49-88, VelocityVerlet.hpp:50-90, SymplecticSteppers.hpp:152-216).

CPU: built with g++ against a recording stand-in of the C ABI
(tests/erod_stub.cpp), the stepper issues, per block and step, the device
calls of one position-Verlet step with the reference's coefficients and
times; velocity Verlet's substeps pass their own coefficients (dt/2 on the
dynamic states, dt on the kinematic ones); a failed call ends the step with
its message.
GPU: built against liberod.so, 60 steps through the binding equal 60 steps
of erod_step bitwise.
"""
import ctypes
import os
import signal
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")


def _build(tmp, link_real):
    so = os.path.join(tmp, "libsurface_%s.so" % ("erod" if link_real else "stub"))
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-pedantic", "-fPIC", "-shared", "-o", so,
           os.path.join(TESTS, "stepper_surface.cpp")]
    if link_real:
        cmd += ["-L", os.path.join(ROOT, "elasticapp_amd"), "-lerod", "-Wl,-rpath," + os.path.join(ROOT, "elasticapp_amd")]
    else:
        cmd += [os.path.join(TESTS, "erod_stub.cpp")]
    subprocess.run(cmd, check=True)
    L = ctypes.CDLL(so)
    vp, dbl = ctypes.c_void_p, ctypes.c_double
    L.surface_position_verlet.argtypes = [ctypes.POINTER(vp), ctypes.c_int, dbl, dbl, ctypes.c_int, ctypes.POINTER(dbl)]
    L.surface_velocity_verlet_substeps.argtypes = [vp, dbl, dbl]
    if not link_real:
        L.stub_log.restype = ctypes.c_char_p
    return L


@pytest.fixture(scope="module")
def stub(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("surf")), False)


def _log(L):
    out = []
    for line in L.stub_log().decode().splitlines():
        w = line.split()
        out.append((w[0], int(w[1])) + tuple(float(x) for x in w[2:]))
    L.stub_clear()
    return out


def test_position_verlet_issues_the_reference_calls(stub):
    EXT_LOADS, EXT_TORQUES = 6, 24
    handles = (ctypes.c_void_p * 2)(0x1000, 0x2000)
    t0, dt, n = 0.25, 1e-3, 3
    t_out = ctypes.c_double()
    stub.stub_clear()
    assert stub.surface_position_verlet(handles, 2, t0, dt, n, ctypes.byref(t_out)) == 0
    got = _log(stub)
    want = []
    t = t0
    for _ in range(n):
        want += [("kinematic_update_by", h, 0.5 * dt) for h in (0x1000, 0x2000)]
        t += 0.5 * dt
        want += [("apply_post_kinematic", h, t) for h in (0x1000, 0x2000)]
        for h in (0x1000, 0x2000):
            want += [("fill_field", h, EXT_LOADS, 0.0), ("fill_field", h, EXT_TORQUES, 0.0),
                     ("compute_internal_dynamics", h, t)]
        want += [("apply_pre_dynamic", h, t) for h in (0x1000, 0x2000)]
        want += [("dynamic_update", h, dt) for h in (0x1000, 0x2000)]
        want += [("apply_post_dynamic", h, t) for h in (0x1000, 0x2000)]
        t += 0.5 * dt
        want += [("kinematic_update_by", h, 0.5 * dt) for h in (0x1000, 0x2000)]
    assert got == want
    assert t_out.value == t


def test_velocity_verlet_coefficients(stub):
    stub.stub_clear()
    assert stub.surface_velocity_verlet_substeps(0x3000, 1.0, 2e-3) == 0
    assert _log(stub) == [("dynamic_update", 0x3000, 1e-3), ("kinematic_update_by", 0x3000, 2e-3)]


def test_a_failed_call_ends_the_step_with_its_message(tmp_path):
    """A failing C call throws out of the binding; inside the stepper's
    noexcept substeps that ends the program with the call's message, as a
    failed assertion ends the reference's (a stand-in whose kinematic update
    fails, run in a child process)."""
    src = open(os.path.join(TESTS, "erod_stub.cpp")).read()
    bad = src.replace('return rec("kinematic_update_by", b, "%.17g", c);', "return EROD_ERR_HIP;")
    assert bad != src
    p = tmp_path / "bad_stub.cpp"
    p.write_text(bad.replace('#include "../include/erod.h"', '#include "%s"' % os.path.join(ROOT, "include", "erod.h")))
    so = str(tmp_path / "libbad.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", so, os.path.join(TESTS, "stepper_surface.cpp"),
                    str(p)], check=True)
    child = ("import ctypes\n"
             "L = ctypes.CDLL(%r)\n"
             "h = (ctypes.c_void_p * 1)(0x1000)\n"
             "t = ctypes.c_double()\n"
             "L.surface_position_verlet(h, 1, ctypes.c_double(0.0), ctypes.c_double(1e-3), 1, ctypes.byref(t))\n"
             "print('returned')\n" % so)
    r = subprocess.run([sys.executable, "-c", child], capture_output=True, text=True, timeout=60)
    assert r.returncode == -signal.SIGABRT and "returned" not in r.stdout
    assert "erod: stub failure" in r.stderr


@pytest.mark.gpu
def test_binding_steps_equal_erod_step(tmp_path):
    from tests.gpu_helpers import STATE, device_block, random_workload
    L = _build(str(tmp_path), True)
    wl = random_workload(12, 3, 90, seed=61, steps=60)
    for r in range(3):
        wl.clamps.append((r, wl.states[r]["Position"][:, 0].copy(), wl.states[r]["Director"][:, 0].copy()))
    a, b = device_block(wl), device_block(wl)
    handles = (ctypes.c_void_p * 1)(a._h.value)
    t = ctypes.c_double()
    assert L.surface_position_verlet(handles, 1, 0.0, wl.dt, wl.steps, ctypes.byref(t)) == 0
    b.step(wl.dt, wl.steps)
    assert t.value == b.time
    for r in range(len(wl.n_elems)):
        for tag in STATE + ("InternalLoads", "InternalTorques", "Curvature"):
            assert np.array_equal(a.get_rod(tag, r), b.get_rod(tag, r)), (r, tag)

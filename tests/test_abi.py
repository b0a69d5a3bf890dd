"""The C-ABI library loads without a GPU and exports every entry point that
include/erod.h declares; the Python field table agrees with the header."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "erod.h")
LIB = os.path.join(ROOT, "elasticapp_amd", "liberod.so")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(erod_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        if not os.path.exists("/opt/rocm/bin/hipcc"):
            pytest.skip("liberod.so not built and no hipcc here")
        from elasticapp_amd import build
        build.build()
    return LIB


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 30
    for must in ("erod_batch_create", "erod_kinematic_update", "erod_compute_internal_dynamics",
                 "erod_dynamic_update", "erod_step_fused_graph", "erod_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(erod_\w+)$", out, re.M))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_field_shapes(lib):
    import elasticapp_amd as ea
    L = ea.load_library()
    assert ea.FIELDS["Director"][1:] == (9, ea.ELEMENT)
    assert ea.FIELDS["Position"][1:] == (3, ea.NODE)
    assert ea.FIELDS["BendingTwistRigidityMatrix"][1:] == (3, ea.VORONOI)
    assert len(ea.FIELDS) == 36
    # error path without touching the device: bad field id
    import ctypes
    nc, st = ctypes.c_int(), ctypes.c_int()
    assert L.erod_field_shape(99, ctypes.byref(nc), ctypes.byref(st)) == -1
    assert b"field" in L.erod_last_error()


def test_no_cpu_fallback_without_gpu(lib):
    """Creating a batch on a machine without a visible GPU fails loudly."""
    import elasticapp_amd as ea
    ea.load_library()
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is visible")
    with pytest.raises(ea.ErodError):
        ea.DeviceBlock([4, 5])


def test_bad_arguments_fail_before_any_device_call(lib):
    """Arguments the library can refuse on the host are refused with a status
    and a message, without touching a device (so this runs with no GPU)."""
    import ctypes
    import elasticapp_amd as ea
    L = ea.load_library()
    out = ctypes.c_void_p()
    one = (ctypes.c_uint64 * 1)(1)
    assert L.erod_batch_create(one, 1, 1, ctypes.byref(out)) != 0 and not out.value
    assert b"< 2 elements" in L.erod_last_error()
    two = (ctypes.c_uint64 * 1)(8)
    assert L.erod_batch_create(two, 0, 1, ctypes.byref(out)) != 0
    # a rod count beyond 32-bit rod ids is refused before the array is read
    assert L.erod_batch_create(two, 1 << 32, 1, ctypes.byref(out)) != 0 and b"2^32" in L.erod_last_error()
    assert L.erod_batch_destroy(None) == 0
    t = ctypes.c_double(0.0)
    assert L.erod_step(None, ctypes.byref(t), 1e-3, 1) != 0
    assert L.erod_step(None, None, 1e-3, 1) != 0 and b"null time" in L.erod_last_error()
    n = (ctypes.c_double * 3)(0.0, 0.0, 1.0)
    assert L.erod_set_plane_normal(None, n) != 0
    done = ctypes.c_uint64()
    assert L.erod_server_tick(None, None, 0, ctypes.byref(done)) != 0 and b"null server" in L.erod_last_error()
    assert L.erod_load(b"/nonexistent/erod.ckpt", ctypes.byref(out)) != 0 and b"cannot open" in L.erod_last_error()


def test_header_is_plain_c(tmp_path):
    """include/erod.h compiles as C99 with -pedantic: a cgo or FFI binding can
    include it as it is."""
    import shutil
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "uses_header.c"
    src.write_text('#include "erod.h"\nint main(void) { return erod_batch_destroy((erod_batch*)0); }\n')
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

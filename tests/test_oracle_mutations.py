"""The golden vectors and KATs pin the oracle: a copy of the oracle with one
plausible mistake planted (a dropped term, a wrong sign, a wrong index, a
transposed operand, a wrong power) fails them, while an unchanged copy built
the same way passes.

Each mutation is one textual replacement in a copy of cosserat_oracle.hpp,
compiled on the CPU into a temporary directory; nothing in oracle/ changes.
"""
import concurrent.futures
import os
import shutil
import subprocess

import pytest

from oracle import pyoracle
from tests import test_oracle_golden as golden

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")

# (what the oracle says, the planted mistake, what kind of mistake)
MUTATIONS = [
    ("tau[c * E + k] += d1[c * E + k] * l[k];", "(void)0;", "dropped unsteady-dilatation torque"),
    ("o[nin] = T(0.5) * a[nin - 1];", "o[nin] = T(0.25) * a[nin - 1];", "wrong quadrature end weight"),
    ("- q[6 + j] * u[1]);", "+ q[6 + j] * u[1]);", "sign in the Rodrigues update"),
    ("const T em = T(0.5) * (m[k] + m[k + 1]);", "const T em = T(0.5) * (m[k] + m[k]);", "index in the torque damping mass"),
    ("D[i] = T(1.0) / (eps[i] * eps[i] * eps[i]);", "D[i] = T(1.0) / (eps[i] * eps[i]);", "power of the dilatation"),
    ("cross(qt, nv, o);", "cross(nv, qt, o);", "transposed cross-product operands"),
    ("d1[c * E + k] = d2[c * E + k] / e[k];", "d1[c * E + k] = d2[c * E + k] * e[k];", "J w / e inverted"),
]

# mistakes in the environment hooks, which the golden vectors do not reach;
# the contact KATs must catch them
ENV_MUTATIONS = [
    ("const T vt[3] = {v[i] - vnv * nx, v[N + i] - vnv * ny, v[2 * N + i] - vnv * nz};",
     "const T vt[3] = {v[i], v[N + i], T(0)};", "slip taken in the x-y plane"),
    ("for (int c = 0; c < 3; ++c) fe[c * N + i] += fn * nv[c];", "fe[2 * N + i] += fn;", "normal force along z"),
    ("std::min(T(s.plane.ct) * vtn, T(s.plane.mu) * fn)", "std::max(T(s.plane.ct) * vtn, T(s.plane.mu) * fn)",
     "friction not capped at mu f_N"),
    ("T(s.plane.k) * delta + T(s.plane.c) * vn", "T(s.plane.k) * delta - T(s.plane.c) * vn", "sign of the contact damping"),
    ("fe[c * N + i] += T(s.g[c]) * m[i];", "fe[c * N + i] += T(s.g[c]);", "gravity without the mass"),
]

# mistakes in the clamps at an index, which the golden vectors (clamp-free)
# do not reach; the clamped-clamped beam's closed form must catch them
CLAMP_MUTATIONS = [
    ("const std::size_t e = std::min(c.index, b.n_elems[ri] - 1);\n      for (int d = 0; d < 3; ++d) b.get(Position)",
     "const std::size_t e = c.index;\n      for (int d = 0; d < 3; ++d) b.get(Position)", "last end held at its node alone"),
    ("b.get(Position)[d * N + st + c.index] = T(c.x[d]);", "b.get(Position)[d * N + st + e] = T(c.x[d]);",
     "the element's index for the node"),
]

FLAGS = ["-O1", "-ffp-contract=off", "-std=c++20", "-fPIC", "-pthread", "-shared"]


def _build(tmp, i, mutation):
    d = os.path.join(tmp, "m%d" % i)
    os.makedirs(d)
    for f in ("cosserat_oracle.hpp", "oracle_capi.cpp"):
        shutil.copy(os.path.join(ORACLE, f), d)
    if mutation is not None:
        p = os.path.join(d, "cosserat_oracle.hpp")
        src = open(p).read()
        assert src.count(mutation[0]) >= 1, mutation[0]
        open(p, "w").write(src.replace(mutation[0], mutation[1], 1))
    so = os.path.join(d, "liboracle.so")
    subprocess.run(["g++"] + FLAGS + ["-o", so, os.path.join(d, "oracle_capi.cpp")], check=True)
    return so


def _failing_golden_cases():
    fails = 0
    for p in golden._step_cases():
        case, section = p.values
        try:
            golden.test_golden_steps(case, section)
        except AssertionError:
            fails += 1
    try:
        golden.test_golden_single_evaluation()
    except AssertionError:
        fails += 1
    return fails


def _failing_env_kats():
    from tests import test_oracle_kats as kats
    fails = 0
    for t in (kats.test_free_fall_is_exact_under_position_verlet, kats.test_resting_contact_carries_the_weight,
              kats.test_sliding_friction_decelerates_at_mu_g, kats.test_plane_contact_is_rotation_equivariant):
        try:
            t()
        except AssertionError:
            fails += 1
    return fails


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mut"))
    jobs = [None] + MUTATIONS + ENV_MUTATIONS + CLAMP_MUTATIONS
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(lambda im: _build(tmp, im[0], im[1]), enumerate(jobs)))
    yield libs
    pyoracle.use_library(None)


def test_unchanged_copy_passes(builds):
    pyoracle.use_library(builds[0])
    try:
        assert _failing_golden_cases() == 0
        assert _failing_env_kats() == 0
    finally:
        pyoracle.use_library(None)


@pytest.mark.parametrize("i", range(len(MUTATIONS)), ids=[m[2] for m in MUTATIONS])
def test_planted_mistake_is_caught(builds, i):
    pyoracle.use_library(builds[i + 1])
    try:
        assert _failing_golden_cases() > 0, MUTATIONS[i][2]
    finally:
        pyoracle.use_library(None)


@pytest.mark.parametrize("i", range(len(ENV_MUTATIONS)), ids=[m[2] for m in ENV_MUTATIONS])
def test_planted_environment_mistake_is_caught(builds, i):
    pyoracle.use_library(builds[1 + len(MUTATIONS) + i])
    try:
        assert _failing_env_kats() > 0, ENV_MUTATIONS[i][2]
    finally:
        pyoracle.use_library(None)


def _clamp_kat_fails():
    from tests import test_oracle_kats as kats
    try:
        kats.test_beam_clamped_at_both_ends_converges()
    except AssertionError:
        return True
    return False


def test_unchanged_copy_passes_the_clamp_kat(builds):
    pyoracle.use_library(builds[0])
    try:
        assert not _clamp_kat_fails()
    finally:
        pyoracle.use_library(None)


@pytest.mark.parametrize("i", range(len(CLAMP_MUTATIONS)), ids=[m[2] for m in CLAMP_MUTATIONS])
def test_planted_clamp_mistake_is_caught(builds, i):
    pyoracle.use_library(builds[1 + len(MUTATIONS) + len(ENV_MUTATIONS) + i])
    try:
        assert _clamp_kat_fails(), CLAMP_MUTATIONS[i][2]
    finally:
        pyoracle.use_library(None)

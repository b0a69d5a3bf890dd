"""The CPU oracle against the reference's PyElastica golden vectors.

Mirrors tests/Unit/Systems/CosseratRods/FullRodTest/Test_PerturbedRodStep.cpp,
Test_PerturbedRodStepWithoutDamping.cpp and Test_PerturbedRod.cpp: build the
rods with the straight-rod initializer, overwrite the state with the
*_REPLACED arrays in the reference's assignment order, step with position
Verlet (dt = 2e-5) and compare all 27 variables with the reference's own
relative tolerances (4e-13 ... 4e-12; 1e-12 for the single evaluation).
"""
import numpy as np
import pytest

from oracle import pyoracle
from tests.golden_fixtures import as_field, load, rel_err_ok


def _build(case):
    p = case["init"]
    n = [p["n_elem"]] * case["n_rods"]
    # RestrictSizeAcrossBlockTypes{8} puts each 5-element rod in its own block
    max_block = 8 if case["separate_blocks"] else 1024
    damped = case["plugin"] == "CosseratRod"
    sysm = pyoracle.System(n, max_block_elems=max_block, damped=damped)
    for r in range(case["n_rods"]):
        sysm.init_straight_rod(r, p["radius"], p["length"], p["origin"], p["direction"], p["normal"],
                               p["density"], p["youngs_modulus"], p["shear_modulus"],
                               p["force_damping"] or 0.0, p["torque_damping"] or 0.0)
    for a in case["assign"]:
        entry = case["replaced"][str(a["rod"])][a["tag"]]
        sysm.rod_field(a["rod"], a["tag"])[...] = as_field(entry, a["tag"], a["shrink_diagonal"])
    return sysm


def _compare(sysm, expected, eps):
    bad = []
    for rod, vars_ in expected.items():
        for tag, entry in vars_.items():
            exp = as_field(entry, tag)
            got = sysm.rod_field(int(rod), tag)
            ok = rel_err_ok(got, exp, eps)
            if not ok.all():
                rel = np.max(np.abs(got - exp) / np.maximum(np.abs(exp), 1e-300))
                bad.append("rod%s %s: %d/%d off, max rel %.3e" % (rod, tag, (~ok).sum(), ok.size, rel))
    return bad


def _step_cases():
    out = []
    for f in ("fullrod_PerturbedRodStep.json", "fullrod_PerturbedRodStepWithoutDamping.json"):
        for case in load(f)["cases"]:
            for sec in case["sections"]:
                out.append(pytest.param(case, sec, id="%s-%s" % (case["name"].split(".")[-2] + "." +
                                                                  case["name"].split(".")[-1], sec["name"])))
    return out


@pytest.mark.parametrize("case,section", _step_cases())
def test_golden_steps(case, section):
    """Every variable the reference compares, at its tolerance; and, as its
    test runner enables floating-point exceptions, none raised."""
    sysm = _build(case)
    pyoracle.fp_exceptions()
    for _ in range(section["steps"]):
        sysm.step(case["dt"], 1)
    assert pyoracle.fp_exceptions() == []
    bad = _compare(sysm, section["expected"], section["epsilon"])
    assert not bad, "\n".join(bad)


def test_golden_single_evaluation():
    """InitializeStraightRod: the initializer's derived constants (A14) plus one
    compute_internal_forces / compute_internal_torques / update_acceleration."""
    case = [c for c in load("fullrod_PerturbedRod.json")["cases"] if c["name"].endswith("InitializeStraightRod")][0]
    sysm = _build(case)
    sec = case["sections"][0]
    assert sec["calls"] == ["compute_internal_forces", "compute_internal_torques", "update_acceleration"]
    pyoracle.fp_exceptions()
    sysm.compute_internal_forces()
    sysm.compute_internal_torques()
    sysm.update_acceleration()
    assert pyoracle.fp_exceptions() == []
    bad = _compare(sysm, sec["expected"], sec["epsilon"])
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("name", ["WithDynamicAndExternal", "FullWithDirector"])
def test_golden_two_rods_one_block_layout(name):
    """Two rods in ONE ghost-padded block: what is written through a rod slice
    reads back through it (Test_PerturbedRod.cpp:267-803)."""
    case = [c for c in load("fullrod_PerturbedRod.json")["cases"] if c["name"].endswith(name)][0]
    sysm = _build(case)
    assert sysm.n_blocks == 1
    # rod 1 starts after rod 0's 5 elements and the 2 ghost elements of the gap
    assert sysm.location(1) == (0, 7)
    sec = case["sections"][0]
    assigned = {(a["rod"], a["tag"]) for a in case["assign"]}
    expected = {r: {t: e for t, e in v.items() if (int(r), t) in assigned} for r, v in sec["expected"].items()}
    bad = _compare(sysm, expected, sec["epsilon"])
    assert not bad, "\n".join(bad)

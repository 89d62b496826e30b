"""The closed-form pins of the oracle's wall friction (contactLinear_wall,
src/contactmodels.cuh:10-109), integrate's fixvel branch
(src/integration.cuh:65-77) and contact model 3 (contactHertz,
src/contactmodels.cuh:499-666) catch plausible mistakes.  Each case builds the
oracle with one planted mistake (-DORC_MUTANT=k: a flipped sign, a swapped
threshold, a dropped term), runs the pins (tests/test_oracle_physics.py or
tests/test_oracle_hertz.py) against that build in a child process and
requires the named pin to fail.  The mistakes
live only in those builds; the oracle's default build is the restatement."""

import os
import re
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")

MUTANTS = {
    1: ("viscous wall friction pushes along the slip", ["test_floor_friction_regimes_one_step[0.01-1.0-viscous]"]),
    2: ("static and dynamic wall friction swapped at 1 mm/s",
        ["test_floor_friction_regimes_one_step[0.0011-100000.0-dynamic]",
         "test_floor_friction_regimes_one_step[0.0009-100000.0-static]"]),
    3: ("wall friction torque with the wrong sign", ["test_sliding_sphere_on_floor_rolls_at_five_sevenths"]),
    4: ("scalar wall velocity added to z only",
        ["test_moving_wall_adds_its_scalar_velocity_to_every_component"]),
    5: ("Coulomb-capped wall friction pushes along the slip",
        ["test_sliding_sphere_on_floor_rolls_at_five_sevenths"]),
    6: ("fixvel ignored", ["test_fixvel_free_flight_keeps_velocities_to_the_bit"]),
    7: ("fixvel keeps gravity on z", ["test_fixvel_free_flight_keeps_velocities_to_the_bit",
                                      "test_fixvel_sphere_moves_in_z_by_contact_force_only"]),
    8: ("fixvel >= 2 does not hold v_z", ["test_fixvel_hold_z_extension_keeps_the_whole_velocity"]),
    9: ("fixvel spheres spin up", ["test_fixvel_sphere_moves_in_z_by_contact_force_only"]),
}


# contact model 3 (contactHertz, src/contactmodels.cuh:499-666), pinned by
# tests/test_oracle_hertz.py
HERTZ_MUTANTS = {
    10: ("Hertz normal force linear in the overlap", ["test_static_overlap_under_a_load",
                                                      "test_elastic_collision_time_and_peak_overlap"]),
    11: ("rolling resistance along the relative spin", ["test_rolling_resistance[0.0001-viscous]"]),
    12: ("rolling resistance without its mu_r cap", ["test_rolling_resistance[10.0-capped]"]),
    13: ("normal dashpot without the xi^(1/4) factor", ["test_damped_collision_against_the_ode"]),
    14: ("tangential spring without the xi^(1/2) factor", ["test_tangential_stick_and_slip"]),
}


@pytest.mark.parametrize("k", sorted(MUTANTS) + sorted(HERTZ_MUTANTS),
                         ids=[MUTANTS[k][0] for k in sorted(MUTANTS)] + [HERTZ_MUTANTS[k][0] for k in sorted(HERTZ_MUTANTS)])
def test_pins_catch_planted_mistake(k):
    what, pins = MUTANTS[k] if k in MUTANTS else HERTZ_MUTANTS[k]
    suite = "test_oracle_physics.py" if k in MUTANTS else "test_oracle_hertz.py"
    subprocess.run(["make", "-s", "-C", ORACLE, "mutant", f"M={k}"], check=True)
    env = dict(os.environ, SPH_ORACLE_MUTANT_LIB=os.path.join(ORACLE, "build", f"liboracle_dem_mut{k}.so"))
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-rf", "-p", "no:cacheprovider", "-p", "no:xdist",
                          os.path.join(HERE, suite)],
                         capture_output=True, text=True, env=env, cwd=ROOT, timeout=600)
    failed = set(re.findall(r"FAILED \S+::(\S+)", res.stdout))
    assert res.returncode != 0, f"mutant {k} ({what}) passes every pin"
    assert failed & set(pins), f"mutant {k} ({what}): expected one of {pins} to fail, failed {sorted(failed)}"


def test_default_build_is_not_a_mutant():
    """The oracle every other test loads is built without ORC_MUTANT."""
    src = open(os.path.join(ORACLE, "Makefile")).read()
    all_rule = src.split("mutant:")[0]
    assert "ORC_MUTANT" not in all_rule

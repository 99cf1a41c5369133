"""Planted mistakes in the oracle must fail its pins.

Each mutant replaces one exact piece of oracle/rainbow_oracle.cc (the piece
must occur exactly once, so a mutant cannot silently stop applying when the
oracle changes), is compiled with the oracle's own flags into a temporary
library, and the pin named beside it must then raise AssertionError.  This
shows that the pins of the compensated sums, the potential lanes, the FP64
direct sum and the FP64 conserved quantities are not restatements of the
code they check: a dropped term, a wrong sign, a swapped branch or a wrong
index in those functions fails one of them.
"""

import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

import oracle as orc
import test_oracle_compensation as comp
import test_oracle_potential as pot

SOURCE = Path(orc.ORACLE_DIR) / "rainbow_oracle.cc"
FLAGS = ["-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared"]

# name: (original text, replacement, pin that must fail)
MUTANTS = {
    "neumaier-branch-swapped": (
        "if (std::fabs(s) >= std::fabs(t))", "if (std::fabs(s) < std::fabs(t))", comp.check_known_answers),
    "kahan-without-branch": (
        "    if (std::fabs(s) >= std::fabs(t))\n      c = c + ((s - u) + t);\n    else\n      c = c + ((t - u) + s);",
        "    c = c + ((s - u) + t);", comp.check_known_answers),
    "compensation-overwritten": (
        "      c = c + ((s - u) + t);", "      c = ((s - u) + t);", comp.check_known_answers),
    "value-drops-compensation": (
        "R value() const { return s + c; }", "R value() const { return s; }", comp.check_known_answers),
    "near-sum-restarts-from-zero": (
        "    CompField<R> comp;\n    if (kComp) comp.reset(result);\n    near_stencil(",
        "    CompField<R> comp;\n    FieldT<R> zero_init;\n    zero_field(zero_init);\n"
        "    if (kComp) comp.reset(zero_init);\n    near_stencil(",
        comp.check_compensated_solve_matches_plain_and_f64),
    "compensated-own-leaf-dropped": (
        "      if (has_self) s.add(self[k]);\n", "", comp.check_compensated_solve_matches_plain_and_f64),
    "compensated-expansion-wrong-index": (
        "{field.j0[1], field.j0[3], field.j1[0]},", "{field.j0[1], field.j0[3], field.j0[2]},",
        comp.check_compensated_solve_matches_plain_and_f64),
    "m2l-compensated-store-dropped": (
        "    if (kComp) comp.store(result);\n    fields[g] = result;\n  }\n}\n\n// near_all",
        "    fields[g] = result;\n  }\n}\n\n// near_all", comp.check_solve_lanes_are_exact_sums),
    "l2l-potential-sign": (
        "inherited.a[3] = parent_field.a[3] - dot3(parent_field.a, delta);",
        "inherited.a[3] = parent_field.a[3] + dot3(parent_field.a, delta);", pot.check_l2l_potential_telescopes),
    "l2l-potential-dropped": (
        "inherited.a[3] = parent_field.a[3] - dot3(parent_field.a, delta);",
        "inherited.a[3] = parent_field.a[3];", pot.check_l2l_potential_telescopes),
    "body-potential-half-term-sign": (
        "- R(0.5) * dot3(d, jd)) + self_phi;", "+ R(0.5) * dot3(d, jd)) + self_phi;",
        pot.check_body_potential_gradient),
    "body-potential-own-leaf-dropped": (
        "- R(0.5) * dot3(d, jd)) + self_phi;", "- R(0.5) * dot3(d, jd));", pot.check_body_potential_closed_form),
    "body-potential-linear-sign": (
        "*phi = ((field.a[3] - dot3(field.a, d))", "*phi = ((field.a[3] + dot3(field.a, d))",
        pot.check_body_potential_closed_form),
    "direct-sum-self-included": (
        "      if (j == i) continue;\n      const double dx", "      const double dx",
        pot.check_direct_sum_closed_forms),
    "direct-sum-unsoftened": (
        "const double r2 = dx * dx + dy * dy + dz * dz + e2;", "const double r2 = dx * dx + dy * dy + dz * dz;",
        pot.check_direct_sum_closed_forms),
    "direct-sum-sign": ("      ay += dy * inv3;", "      ay -= dy * inv3;", pot.check_direct_sum_closed_forms),
    "direct-sum-mass-dropped": (
        "out[4 * t + 0] = gravity * m * ax;", "out[4 * t + 0] = gravity * ax;", pot.check_direct_sum_closed_forms),
    "direct-sum-potential-not-gradient": (
        "      phi -= inv;", "      phi -= inv * inv;", pot.check_direct_sum_third_law_and_gradient),
    "angular-momentum-cross-sign": (
        "ly += m * (z * vx - x * vz);", "ly += m * (x * vz - z * vx);", pot.check_conserved_rigid_motions),
    "kinetic-energy-half-dropped": ("ke += 0.5 * m * (", "ke += m * (", pot.check_conserved_rigid_motions),
    "momentum-mass-dropped": ("px += m * vx;", "px += vx;", pot.check_conserved_rigid_motions),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = SOURCE.read_text()
    out = tmp_path_factory.mktemp("oracle_mutants")

    def build(item):
        name, (old, new, _) = item
        assert src.count(old) == 1, f"mutant {name}: its original text occurs {src.count(old)} times"
        cc = out / f"{name}.cc"
        so = out / f"{name}.so"
        cc.write_text(src.replace(old, new))
        subprocess.run(["g++", *FLAGS, "-o", str(so), str(cc)], check=True, capture_output=True)
        return name, so

    with ThreadPoolExecutor(max_workers=8) as pool:
        return dict(pool.map(build, MUTANTS.items()))


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_its_pin(name, mutant_libs):
    check = MUTANTS[name][2]
    with orc.use_library(mutant_libs[name]):
        with pytest.raises(AssertionError):
            check()
    # and the unmutated oracle passes the same pin
    check()

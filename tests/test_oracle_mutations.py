"""The oracle's pins catch planted mistakes.

Every case copies oracle/ss_oracle.c, plants one plausible mistake by a
textual substitution (the pattern must occur exactly once, so a change of the
oracle that moves it fails here loudly), compiles the copy with gcc into a
temporary directory and runs the pin that should catch it: the pin must raise
AssertionError on the mutant and pass on the built oracle.
"""
import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_pins as pins
from oracle_lib import ORACLE_DIR, Oracle

SRC = os.path.join(ORACLE_DIR, "ss_oracle.c")

# (name, pattern, replacement, pin that must fail)
MUTATIONS = [
    ("diffusion_drops_volume_exclusion", "(p->diffusion_translational + p->volume_exclusion * rho)",
     "(p->diffusion_translational + 0.0 * rho)", pins.pin_translational_diffusion_amplitude),
    ("diffusion_drops_factor_two", "sqrt(2. * dt * (p->diffusion_translational",
     "sqrt(dt * (p->diffusion_translational", pins.pin_translational_diffusion_known_answer),
    ("diffusion_amplitude_without_sqrt", "double diff = sqrt(2. * dt * (p->diffusion_translational + p->volume_exclusion * rho));",
     "double diff = 2. * dt * (p->diffusion_translational + p->volume_exclusion * rho);",
     pins.pin_translational_diffusion_amplitude),
    ("strain_projection_sign", "for (int c = 0; c < 3; ++c) f[c] -= n[c] * fn;",
     "for (int c = 0; c < 3; ++c) f[c] += n[c] * fn;", pins.pin_jeffrey_strain_closed_forms),
    ("strain_without_projection", "for (int c = 0; c < 3; ++c) f[c] -= n[c] * fn;",
     "for (int c = 0; c < 3; ++c) f[c] -= 0.0 * n[c] * fn;", pins.pin_jeffrey_strain_closed_forms),
    ("strain_without_shape", "for (int c = 0; c < 3; ++c) f[c] *= p->shape;",
     "for (int c = 0; c < 3; ++c) f[c] *= 1.0;", pins.pin_jeffrey_strain_closed_forms),
    ("strain_transposed", "matvec(E, n, f);\n        for (int c = 0; c < 3; ++c) f[c] *= p->shape;",
     "{ double Et[9] = {E[0], E[3], E[6], E[1], E[4], E[7], E[2], E[5], E[8] + 0.5};"
     " matvec(Et, n, f); }\n        for (int c = 0; c < 3; ++c) f[c] *= p->shape;",
     pins.pin_jeffrey_strain_closed_forms),
    ("rods_zz_sign", "s[8] = -2. * ct * (st * st);", "s[8] = 2. * ct * (st * st);",
     pins.pin_rods_stress_special_angles),
    ("rods_xy_single_angle", "s[1] = ct * (st * st) * cos(2. * phi);", "s[1] = ct * (st * st) * cos(phi);",
     pins.pin_rods_stress_special_angles),
    ("rods_yy_single_angle", "s[4] = st * sin(2. * theta) * (sp * sp);", "s[4] = st * sin(theta) * (sp * sp);",
     pins.pin_rods_stress_special_angles),
    ("rods_xz_wrong_factor", "s[2] = cos(2. * theta) * cp * st;", "s[2] = cos(2. * theta) * sp * st;",
     pins.pin_rods_stress_special_angles),
    ("flow_projection_sign", "double ar = (sr[i] - kn[i] * kr) * fac;", "double ar = (sr[i] + kn[i] * kr) * fac;",
     pins.pin_compare_implementations),
    ("flow_missing_normalisation", "double fac = (s->ik2[idx] + screening) / norm;",
     "double fac = (s->ik2[idx] + screening) / (norm + 1.0);", pins.pin_compare_implementations),
]


def _build(tmp, name, pattern, repl):
    text = open(SRC).read()
    assert text.count(pattern) == 1, f"{name}: pattern occurs {text.count(pattern)} times"
    src = os.path.join(tmp, f"{name}.c")
    with open(src, "w") as f:
        f.write(text.replace(pattern, repl))
    so = os.path.join(tmp, f"{name}.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-std=c11", "-shared", f"-I{ORACLE_DIR}", "-o", so, src,
                    "-lm"], check=True, capture_output=True)
    return so


@pytest.fixture(scope="module")
def mutants():
    with tempfile.TemporaryDirectory(prefix="oracle_mut_") as tmp:
        with ThreadPoolExecutor(max_workers=4) as ex:
            futs = {m[0]: ex.submit(_build, tmp, *m[:3]) for m in MUTATIONS}
            yield {k: v.result() for k, v in futs.items()}


@pytest.mark.parametrize("pin", pins.PINS, ids=lambda f: f.__name__)
def test_pin_passes_on_oracle(oracle, pin):
    pin(oracle)


@pytest.mark.parametrize("mutation", MUTATIONS, ids=lambda m: m[0])
def test_pin_fails_on_planted_mistake(mutants, mutation):
    name, _, _, pin = mutation
    mutant = Oracle(so_path=mutants[name])
    with pytest.raises(AssertionError):
        pin(mutant)

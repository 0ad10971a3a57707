"""Pins of the oracle's non-periodic closure rows, and proof that they bite.

`closure_pins` checks the condition that fixes each closure row of cd06,
cd10, cf90 and the one-sided cd06stagg ends (row-by-row order of accuracy,
global conservation of the first-derivative closures, the compact-stencil
family members, constant preservation and Nyquist removal of the cf90 rows).

The mutation tests rebuild the oracle from its C source into a scratch
directory with one closure constant perturbed by 2 % (the constant and every
value derived from it, as a transcription error in that constant would do),
and require that at least one pin fails.  The source edit is anchored on the
routine that defines the constant, so a renamed or moved constant makes the
test fail loudly rather than silently test nothing.

Constants left out, with the reason:
* the row weights w1..w3 (cd06), w1..w4 (cd10) scale a row of A and of B
  together and do not change the operator A^-1 B; they are still covered by
  the conservation pin, so they are mutated below; the staggered weights
  w0s, w1s (cd06stagg.F90:56) have no such identity and are output-invisible;
* alpha1 = 37/183 of the staggered E2C row 2
  (STAGG_CD06_files/D1RHS_E2C_common.F90:2) belongs to a family that is
  fourth order for every alpha1 and is neither conservative nor compact for
  any member; no mathematical condition fixes it, so it is pinned only by
  the literal value, read against the reference line.
"""
import os
import shutil
import subprocess

import pytest

import closure_pins as C
import oracle as O

SRC = os.path.join(os.path.dirname(O.__file__), "padeops_oracle.c")

CD06 = "static cd06_np_t cd06_np(void)"
CD10 = "static cd10_np_t cd10_np(void)"
CF90 = "static const double b2_alpha90"
STG_TRI = "static int stg_tri("
STG_RHS = "static void stg_rhs("
STG_LELE = "static void stg_lele("
STG_DEF = "#define SG_ALPHA"

# name -> list of (anchor, old text after the anchor, new text)
MUTATIONS = {
    # cd06 first-derivative closure (cd06.F90:20-53)
    "cd06 alpha": [(CD06, "c.alpha = 3.0", "c.alpha = 3.06")],
    "cd06 p": [(CD06, "c.p = 17.0 / 6.0", "c.p = 1.02 * 17.0 / 6.0")],
    "cd06 q": [(CD06, "c.q = 3.0 / 2.0", "c.q = 1.02 * 3.0 / 2.0")],
    "cd06 r": [(CD06, "c.r = 3.0 / 2.0", "c.r = 1.02 * 3.0 / 2.0")],
    "cd06 s": [(CD06, "c.s = -1.0 / 6.0", "c.s = -1.02 / 6.0")],
    "cd06 q'": [(CD06, "c.q_p = 3.0 / 4.0", "c.q_p = 1.02 * 3.0 / 4.0")],
    "cd06 alpha'": [(CD06, "c.alpha_p = 1.0 / 4.0", "c.alpha_p = 1.02 / 4.0")],
    "cd06 alpha''": [(CD06, "c.alpha_pp = (", "c.alpha_pp = 1.02 * (")],
    "cd06 q''": [(CD06, "c.q_pp = (1.0 / 3.0)", "c.q_pp = (1.02 / 3.0)")],
    "cd06 r''": [(CD06, "c.r_pp = (1.0 / 12.0)", "c.r_pp = (1.02 / 12.0)")],
    "cd06 w1": [(CD06, "c.w1 = (", "c.w1 = 1.02 * (")],
    "cd06 w2": [(CD06, "c.w2 = (", "c.w2 = 1.02 * (")],
    "cd06 w3": [(CD06, "c.w3 = (", "c.w3 = 1.02 * (")],
    # cd10 first-derivative closure (cd10.F90:34-78)
    "cd10 alpha": [(CD10, "c.alpha = 3.0", "c.alpha = 3.06")],
    "cd10 p": [(CD10, "c.p = -17.0 / 6.0", "c.p = -1.02 * 17.0 / 6.0")],
    "cd10 q": [(CD10, "c.q = 3.0 / 2.0", "c.q = 1.02 * 3.0 / 2.0")],
    "cd10 r": [(CD10, "c.r = 3.0 / 2.0", "c.r = 1.02 * 3.0 / 2.0")],
    "cd10 s": [(CD10, "c.s = -1.0 / 6.0", "c.s = -1.02 / 6.0")],
    "cd10 q'": [(CD10, "c.q_p = 3.0 / 4.0", "c.q_p = 1.02 * 3.0 / 4.0")],
    "cd10 alpha'": [(CD10, "c.alpha_p = 1.0 / 4.0", "c.alpha_p = 1.02 / 4.0")],
    "cd10 alpha'''": [(CD10, "c.alpha_ppp = (", "c.alpha_ppp = 1.02 * (")],
    "cd10 beta'''": [(CD10, "c.beta_ppp = (1.0 / 20.0)", "c.beta_ppp = (1.02 / 20.0)")],
    "cd10 q'''": [(CD10, "c.q_ppp = (1.0 / 12.0)", "c.q_ppp = (1.02 / 12.0)")],
    "cd10 r'''": [(CD10, "c.r_ppp = (1.0 / 600.0)", "c.r_ppp = (1.02 / 600.0)")],
    "cd10 s'''": [(CD10, "c.s_ppp = (1.0 / 300.0)", "c.s_ppp = (1.02 / 300.0)")],
    "cd10 alpha''": [(CD10, "c.alpha_pp = (", "c.alpha_pp = 1.02 * (")],
    "cd10 beta''": [(CD10, "c.beta_pp = (1.0 / 12.0)", "c.beta_pp = (1.02 / 12.0)")],
    "cd10 q''": [(CD10, "c.q_pp = (2.0 / 18.0)", "c.q_pp = (2.04 / 18.0)")],
    "cd10 r''": [(CD10, "c.r_pp = (1.0 / 72.0)", "c.r_pp = (1.02 / 72.0)")],
    "cd10 w1": [(CD10, "c.w1 = (", "c.w1 = 1.02 * (")],
    "cd10 w2": [(CD10, "c.w2 = (", "c.w2 = 1.02 * (")],
    "cd10 w3": [(CD10, "c.w3 = (", "c.w3 = 1.02 * (")],
    "cd10 w4": [(CD10, "c.w4 = sh", "c.w4 = 1.02 * sh")],
    # cd10 second-derivative closure (cd10.F90:88-104)
    "cd10 d2 b1 alpha": [(CD10, "c.b1_alpha = 11.0", "c.b1_alpha = 11.22")],
    "cd10 d2 b1 a": [(CD10, "c.b1_a = (", "c.b1_a = 1.02 * (")],
    "cd10 d2 b1 b": [(CD10, "c.b1_b = -(", "c.b1_b = -1.02 * (")],
    "cd10 d2 b1 c": [(CD10, "c.b1_c = (", "c.b1_c = 1.02 * (")],
    "cd10 d2 b1 d": [(CD10, "c.b1_d = (", "c.b1_d = 1.02 * (")],
    "cd10 d2 b1 e": [(CD10, "c.b1_e = (11 - c.b1_alpha) / 12", "c.b1_e = (11 - c.b1_alpha) / 12 + 0.02")],
    "cd10 d2 b2 alpha": [(CD10, "c.b2_alpha = 1.0 / 10.0", "c.b2_alpha = 1.02 / 10.0")],
    "cd10 d2 b2 a": [(CD10, "c.b2_a = (", "c.b2_a = 1.02 * (")],
    "cd10 d2 b3 alpha": [(CD10, "c.b3_alpha = 344.0 / 1179.0", "c.b3_alpha = 1.02 * 344.0 / 1179.0")],
    "cd10 d2 b3 beta": [(CD10, "c.b3_beta = (", "c.b3_beta = 1.02 * (")],
    "cd10 d2 b3 a": [(CD10, "c.b3_a = (", "c.b3_a = 1.02 * (")],
    "cd10 d2 b3 b": [(CD10, "c.b3_b = (", "c.b3_b = 1.02 * (")],
    # staggered one-sided closures (STAGG_CD06_files/*, cd06stagg.F90:18-57)
    "stagg SG alpha": [(STG_DEF, "#define SG_ALPHA 3.0", "#define SG_ALPHA 3.06")],
    "stagg SG p": [(STG_DEF, "#define SG_P (17.0 / 6.0)", "#define SG_P (1.02 * 17.0 / 6.0)")],
    "stagg SG q": [(STG_DEF, "#define SG_Q (3.0 / 2.0)", "#define SG_Q (1.02 * 3.0 / 2.0)")],
    "stagg SG r": [(STG_DEF, "#define SG_R (3.0 / 2.0)", "#define SG_R (1.02 * 3.0 / 2.0)")],
    "stagg SG s": [(STG_DEF, "#define SG_S (-1.0 / 6.0)", "#define SG_S (-1.02 / 6.0)")],
    "stagg alpha''": [(STG_LELE, "*alpha_pp = (", "*alpha_pp = 1.02 * (")],
    "stagg q'": [(STG_LELE, "*q_p = 3.0 / 4.0", "*q_p = 1.02 * 3.0 / 4.0")],
    "stagg alpha'": [(STG_LELE, "*alpha_p = 1.0 / 4.0", "*alpha_p = 1.02 / 4.0")],
    "stagg E2C alpha0": [(STG_TRI, "alpha0 = -1.0", "alpha0 = -1.02"), (STG_RHS, "alpha0 = -1.0", "alpha0 = -1.02")],
    "stagg E2C a0": [(STG_RHS, "a0 = (1.0 / 24.0)", "a0 = (1.02 / 24.0)")],
    "stagg E2C b0": [(STG_RHS, "b0 = (1.0 / 8.0)", "b0 = (1.02 / 8.0)")],
    "stagg E2C c0": [(STG_RHS, "c0 = (1.0 / 8.0)", "c0 = (1.02 / 8.0)")],
    "stagg E2C d0": [(STG_RHS, "d0 = -(1.0 / 24.0) * (alpha0 + 1.0)", "d0 = -(1.0 / 24.0) * (alpha0 + 1.02)")],
    "stagg E2C a1": [(STG_RHS, "a1 = (3.0 / 8.0)", "a1 = (3.06 / 8.0)")],
    "stagg E2C b1": [(STG_RHS, "b1 = (1.0 / 8.0)", "b1 = (1.02 / 8.0)")],
    "stagg C2E alpha1": [(STG_TRI, "1.0 / 22.0; dup[N - 1] = 1.0 / 22.0", "1.02 / 22.0; dup[N - 1] = 1.02 / 22.0"),
                         (STG_TRI, "dup[2] = 1.0 / 22.0; ddn[2] = 1.0 / 22.0",
                          "dup[2] = 1.02 / 22.0; ddn[2] = 1.02 / 22.0")],
    "stagg C2E a0": [(STG_RHS, "a0 = -71.0 / 24.0", "a0 = -1.02 * 71.0 / 24.0")],
    "stagg C2E b0": [(STG_RHS, "b0 = 47.0 / 8.0", "b0 = 1.02 * 47.0 / 8.0")],
    "stagg C2E c0": [(STG_RHS, "c0 = -31.0 / 8.0", "c0 = -1.02 * 31.0 / 8.0")],
    "stagg C2E d0": [(STG_RHS, "d0 = 23.0 / 24.0", "d0 = 1.02 * 23.0 / 24.0")],
    "stagg C2E a1": [(STG_RHS, "a1 = 12.0 / 11.0", "a1 = 1.02 * 12.0 / 11.0")],
    "stagg interp E2C alpha0": [(STG_TRI, "const double alpha0 = 1.0;", "const double alpha0 = 1.02;"),
                                (STG_RHS, "alpha0 = 1.0;", "alpha0 = 1.02;")],
    "stagg interp E2C a0": [(STG_RHS, "a0 = (1.0 / 16.0) * (5.0", "a0 = (1.02 / 16.0) * (5.0")],
    "stagg interp E2C b0": [(STG_RHS, "b0 = (1.0 / 16.0) * (9.0", "b0 = (1.02 / 16.0) * (9.0")],
    "stagg interp E2C c0": [(STG_RHS, "c0 = (1.0 / 16.0) * (9.0", "c0 = (1.02 / 16.0) * (9.0")],
    "stagg interp E2C d0": [(STG_RHS, "d0 = (1.0 / 16.0) * (1.0 - alpha0)", "d0 = (1.0 / 16.0) * (1.02 - alpha0)")],
    "stagg interp C2E a0": [(STG_RHS, "a0 = 15.0 / 8.0", "a0 = 1.02 * 15.0 / 8.0")],
    "stagg interp C2E b0": [(STG_RHS, "b0 = -5.0 / 4.0", "b0 = -1.02 * 5.0 / 4.0")],
    "stagg interp C2E c0": [(STG_RHS, "c0 = 3.0 / 8.0", "c0 = 1.02 * 3.0 / 8.0")],
    "stagg interp C2E alpha1": [(STG_TRI, "const double alpha1 = 1.0 / 6.0;", "const double alpha1 = 1.02 / 6.0;"),
                                (STG_RHS, "alpha1 = 1.0 / 6.0", "alpha1 = 1.02 / 6.0")],
    "stagg interp C2E a1": [(STG_RHS, "a1 = (1.0 / 8.0) * (9.0", "a1 = (1.02 / 8.0) * (9.0")],
}

# cf90 closure constants (cf90.F90:27-48), each perturbed on its own
for _name in ("b2_alpha90", "b2_a90", "b2_b90", "b3_alpha90", "b3_beta90", "b3_a90", "b3_b90", "b3_c90",
              "b4_alpha90", "b4_beta90", "b4_a90", "b4_b90", "b4_c90", "b4_d90"):
    MUTATIONS["cf90 " + _name] = [(CF90, _name + " = ", _name + " = 1.02 * ")]


def _mutate(src, edits):
    for anchor, old, new in edits:
        assert src.count(anchor) == 1, "anchor %r not unique" % anchor
        a = src.index(anchor)
        i = src.find(old, a)
        assert i >= 0, "%r not found after %r" % (old, anchor)
        src = src[:i] + new + src[i + len(old):]
    return src


@pytest.fixture(scope="module")
def oracle_source():
    with open(SRC) as f:
        return f.read()


def test_closure_conditions_hold_on_the_oracle():
    assert C.failures() == []


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the perturbed oracle")
@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_perturbed_closure_constant_is_detected(name, oracle_source, tmp_path):
    src = _mutate(oracle_source, MUTATIONS[name])
    assert src != oracle_source
    c = tmp_path / "mutant.c"
    so = tmp_path / "libmutant.so"
    c.write_text(src)
    subprocess.check_call(["gcc", "-O0", "-fPIC", "-shared", "-std=gnu11", "-o", str(so), str(c), "-lm"])
    failed = C.failures(O.load(str(so)))
    assert failed, "perturbing %s is not caught by any closure pin" % name

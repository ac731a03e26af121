"""The oracle's pivot choice pinned against LAPACK and hand-derived known answers
(tests/oracle_pins.py), and the pins shown to catch plausible transcription slips: each
mutation below is applied to a copy of oracle/oracle.c, built into a temporary folder, and the
pins are run against it in a child process (SDS_ORACLE_LIB)."""
import os
import subprocess
import sys

import pytest

import oracle_pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_C = os.path.join(ROOT, "oracle", "oracle.c")


@pytest.mark.parametrize("name", sorted(oracle_pins.PINS))
def test_pin(name):
    oracle_pins.PINS[name]()


# (what the slip is, the source text, its mutated form, a pin that must fail)
MUTATIONS = [
    ("lambda searched over all f rows instead of the block's w",
     "for (int64_t i = k + 1; i < w; ++i) {\n        double a = fabs(F_(i, k));\n        if (a > lambda) {",
     "for (int64_t i = k + 1; i < f; ++i) {\n        double a = fabs(F_(i, k));\n        if (a > lambda) {",
     "pin_bk_cb_rows"),
    ("sigma searched over all f rows instead of the block's w",
     "for (int64_t i = k; i < w; ++i)\n            if (i != rr) sigma = jmax(sigma, fabs(F_(i, rr)));",
     "for (int64_t i = k; i < f; ++i)\n            if (i != rr) sigma = jmax(sigma, fabs(F_(i, rr)));",
     "pin_bk_cb_rows"),
    ("the two Bunch-Kaufman branches after the sigma search swapped",
     "if (akk * sigma >= alpha * lambda * lambda)\n            bc = k;\n        else if (fabs(F_(rr, rr)) >= alpha * sigma)\n            bc = rr;",
     "if (fabs(F_(rr, rr)) >= alpha * sigma)\n            bc = rr;\n        else if (akk * sigma >= alpha * lambda * lambda)\n            bc = k;",
     "pin_bk_single_front"),
    ("the 2x2 partner swapped into column k instead of k+1",
     "swap_front(F, f, k + 1, r, w, lp);",
     "swap_front(F, f, k, r, w, lp);",
     "pin_bk_single_front"),
    ("best 1x1 takes the first acceptable column, not the largest",
     "if (best >= 0 && !(a > bv)) continue;",
     "if (best >= 0) continue;",
     "pin_known_answer_diagonal"),
    ("the 1x1 threshold test reads the block rows only",
     "static int threshold_ok(const double *F, int64_t f, int64_t c, int64_t k, double u, double d) {\n    double m = 0.0;\n    for (int64_t i = k; i < f; ++i) {",
     "static int threshold_ok(const double *F, int64_t f, int64_t c, int64_t k, double u, double d) {\n    double m = 0.0;\n    for (int64_t i = k; i < f - 1; ++i) {",
     "pin_known_answer_bk_fallback"),
    ("LU gamma taken over the fully summed rows only",
     "for (int64_t i = k; i < f; ++i) g = jmax(g, fabs(F_(i, k)));",
     "for (int64_t i = k; i < w; ++i) g = jmax(g, fabs(F_(i, k)));",
     "pin_known_answer_lu_threshold"),
    ("LU keeps the last of equal largest rows instead of the first",
     "        if (a > best) {\n            r = i;\n            best = a;",
     "        if (a >= best) {\n            r = i;\n            best = a;",
     "pin_known_answer_lu_threshold"),
    ("LU row search stops one row early",
     "    for (int64_t i = k + 1; i < w; ++i) {\n        double a = fabs(F_(i, k));\n        if (a > best) {",
     "    for (int64_t i = k + 1; i < w - 1; ++i) {\n        double a = fabs(F_(i, k));\n        if (a > best) {",
     "pin_lu_single_front"),
]


@pytest.mark.parametrize("slip,old,new,pin", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pins_catch_mutation(tmp_path, slip, old, new, pin):
    src = open(ORACLE_C).read()
    assert src.count(old) == 1, f"the mutation site of '{slip}' is not unique in oracle.c"
    (tmp_path / "oracle.c").write_text(src.replace(old, new))
    so = tmp_path / "liboracle.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-ffp-contract=off", "-std=c11", "-shared", "-o", str(so),
                    str(tmp_path / "oracle.c"), "-lm"], check=True)
    env = dict(os.environ, SDS_ORACLE_LIB=str(so),
               PYTHONPATH=os.pathsep.join([os.path.join(ROOT, "sparsedirectsolver.jl_amd"),
                                           os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]))
    out = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "oracle_pins.py")], env=env,
                         capture_output=True, text=True, timeout=300)
    done = [ln for ln in out.stdout.splitlines() if ln.startswith("PINS-DONE")]
    assert done, f"the pins did not run to the end on the mutated oracle:\n{out.stdout[-2000:]}\n{out.stderr[-2000:]}"
    failed = done[0].split(" ", 1)[1].split(",") if " " in done[0] else []
    assert pin in failed, f"'{slip}' was not caught by {pin}; failed pins: {failed}"

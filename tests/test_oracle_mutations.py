"""The branch tests of test_oracle.py must catch a misreading of the terms
they pin: each case compiles a copy of the oracle with one term mutated (on
the CPU, into a temporary directory), runs those tests against it in a
subprocess, and requires them to fail.  The real oracle passes them (the
last case)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pamo_oracle.cpp")
TESTS = "restatement or closed_form or flip or restore or quotient"

MUTATIONS = {
    "skinny weight x5 dropped": ("if (!sp) Qa = 5.0f * Qa;", "if (!sp) Qa = 1.0f * Qa;"),
    "divide by num_tri dropped": ("Qa / (float)nt * tau", "Qa * tau"),
    "normal flip not tested": ("if (vdot(on, nn) < 0) return false;", ""),
    "restore threshold formed in float": ("float thr = (float)(1e-6 * (double)ml * (double)ml);",
                                          "float thr = 1e-6f * ml * ml;"),
    "unguarded quotient in float (D10)": ("float term = sp ? (float)(", "float term = false ? (float)("),
}


def _build(tmp, src_text):
    path = os.path.join(tmp, "pamo_oracle.cpp")
    with open(path, "w") as fp:
        fp.write(src_text)
    out = os.path.join(tmp, "libmut.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
                    "-I", os.path.join(ROOT, "oracle"), "-o", out, path], check=True, capture_output=True)
    return out


def _run_branch_tests(lib):
    env = dict(os.environ, PAMO_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle.py"), "-q", "-x",
                        "-p", "no:cacheprovider", "-k", TESTS], cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode, r.stdout[-400:]


@pytest.mark.parametrize("name", list(MUTATIONS) + ["unmutated"])
def test_branch_tests_catch_mutation(tmp_path, name):
    src = open(SRC).read()
    if name != "unmutated":
        old, new = MUTATIONS[name]
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    lib = _build(str(tmp_path), src)
    rc, tail = _run_branch_tests(lib)
    if name == "unmutated":
        assert rc == 0, tail
    else:
        assert rc != 0 and "failed" in tail, (name, tail)

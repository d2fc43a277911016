"""The flip filter (pamo_amd/csrc/flip_filter.h) answers the reference's
normalise-then-dot flip test (simp_cuda/src/cusimp_free.cu:520-526,560-566)
whenever it decides: checked by a g++ -ffp-contract=off build of
tests/flip_filter_check.cpp over random, near-perpendicular, wide-exponent
and special inputs.  A filter with a loosened margin must be caught."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "pamo_amd", "csrc")


def _build(tmp_path, defines=()):
    exe = str(tmp_path / "flip_check")
    cmd = ["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I", CSRC]
    cmd += ["-D%s" % d for d in defines]
    cmd += ["-o", exe, os.path.join(HERE, "flip_filter_check.cpp")]
    subprocess.run(cmd, check=True)
    return exe


def _run(exe, n):
    r = subprocess.run([exe, str(n)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    rows = {}
    for line in r.stdout.splitlines():
        name, total, decided, bad = line.split()
        rows[name] = (int(total), int(decided), int(bad))
    return r.returncode, rows


def test_filter_never_disagrees(tmp_path):
    code, rows = _run(_build(tmp_path), 400000)
    assert code == 0, rows
    assert set(rows) == {"random", "near_perpendicular", "wide_exponent", "specials"}
    for name, (total, decided, bad) in rows.items():
        assert bad == 0, name
    # it must also decide almost every ordinary case, or it saves nothing
    total, decided, _ = rows["random"]
    assert decided > 0.97 * total
    # and the near-perpendicular set reaches both sides of the margin
    total, decided, _ = rows["near_perpendicular"]
    assert 0.2 * total < decided < 0.95 * total


def test_loosened_margin_is_caught(tmp_path):
    """The same check against a copy of the filter whose relative margin is
    2^-30 instead of 2^-20 (too small for the rounding of the normalised
    dot): it must report mismatches."""
    src = open(os.path.join(CSRC, "flip_filter.h")).read()
    assert "0x1p-20f * tc" in src
    d = tmp_path / "inc"
    d.mkdir()
    (d / "flip_filter.h").write_text(src.replace("0x1p-20f * tc", "0x1p-30f * tc"))
    exe = str(tmp_path / "flip_loose")
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I", str(d), "-o", exe,
                    os.path.join(HERE, "flip_filter_check.cpp")], check=True)
    code, rows = _run(exe, 400000)
    assert code != 0
    assert rows["near_perpendicular"][2] > 0

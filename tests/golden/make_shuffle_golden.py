"""Writes tests/golden/shuffle_thrust.json: rocThrust's thrust::shuffle
permutations (host backend) for a set of sizes.  Needs hipcc and the
rocThrust headers of /opt/rocm; run from the repository root."""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = [1, 2, 3, 5, 8, 15, 16, 17, 31, 32, 33, 100, 255, 256, 257, 1000, 1023, 1025, 4097, 20480]


def main():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "shuffle_thrust")
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, os.path.join(HERE, "shuffle_thrust.cpp")])
        out = subprocess.check_output([exe] + [str(m) for m in SIZES], text=True)
    res = {}
    for line in out.strip().splitlines():
        x = [int(t) for t in line.split()]
        res[str(x[0])] = x[1:]
    with open(os.path.join(HERE, "shuffle_thrust.json"), "w") as f:
        json.dump(res, f)


if __name__ == "__main__":
    main()

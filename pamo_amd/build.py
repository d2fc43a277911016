"""Build recipe for the in-tree native libraries.

libpamo_hip.so   the gfx950 HIP kernels + C ABI (hipcc)
libpamo_synth.so the seeded SDF mesh generator used by benchmarks (g++)
libpamo_verify.so the exact mesh verifier (g++)

Float flags matter for parity with the reference: no FMA contraction
(the reference is built with nvcc --fmad=false, simp_cuda/setup.py:34),
correctly rounded fp32 division and sqrt, no denormal flushing.  Packed FP32
instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("PAMO_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: %s" % " ".join(cmd))


def build_hip():
    src = os.path.join(HERE, "csrc", "pamo_hip.hip")
    out = os.path.join(HERE, "libpamo_hip.so")
    _run(["/opt/rocm/bin/hipcc"] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", out, src])
    return out


def build_synth():
    src = os.path.join(HERE, "csrc", "synth_mc.cpp")
    out = os.path.join(HERE, "libpamo_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-o", out, src])
    return out


def build_verify():
    """The exact verifier (FP64 filtered predicates with an exact fallback)."""
    src = os.path.join(HERE, "csrc", "verify.cpp")
    out = os.path.join(HERE, "libpamo_verify.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-o", out, src])
    return out


def build_all():
    build_hip()
    build_synth()
    build_verify()


if __name__ == "__main__":
    build_all()

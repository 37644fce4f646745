"""Build recipe for the native parts (run on the CPU build host; gfx950 code is cross-compiled).

  libknp_mi355x.so   HIP kernels + C ABI (knp_amd/csrc/knp_mi355x.hip)
The checker's own CPU restatement (oracle/) is built separately by oracle/Makefile.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, so float expressions round like the strict-IEEE reference
# arithmetic they restate; -packed-fp32-ops: no packed FP32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_library():
    src = os.path.join(HERE, "csrc", "knp_mi355x.hip")
    out = os.path.join(HERE, "libknp_mi355x.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, src])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_library()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)

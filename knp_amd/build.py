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


CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-std=c++17", "-O2", "-Wall", "-I" + os.path.join(HERE, "host", "include"),
             "-I" + os.path.join(ROOT, "include")]


def build_backend_plugin(name="libknp-mi355x-backend.so", factory="plugin_single_threaded.cpp"):
    """The knp::core::Backend implementation exporting create_knp_backend: libknp-mi355x-backend.so (beside the
    single-threaded CPU backend) or libknp-mi355x-mt-backend.so (beside the multi-threaded one)."""
    out = os.path.join(HERE, name)
    _run([CXX, *CXX_FLAGS, "-fPIC", "-shared", os.path.join(HERE, "host", "src", "mi355x_backend.cpp"),
          os.path.join(HERE, "host", "src", factory), "-L" + HERE, "-lknp_mi355x", "-Wl,-rpath,$ORIGIN",
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cpp_tests():
    """bin/mi355x_backend_test: the reference's backend tests written against the knp::core API."""
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    out = os.path.join(HERE, "bin", "mi355x_backend_test")
    _run([CXX, *CXX_FLAGS, os.path.join(ROOT, "tests", "cpp", "mi355x_backend_test.cpp"),
          os.path.join(HERE, "host", "src", "mi355x_backend.cpp"), "-L" + HERE, "-lknp_mi355x", "-ldl",
          "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_float_repeat_test():
    """bin/float_repeat_test: the closed-form repeated float addition (csrc/float_repeat.h, used by the training
    kernels) against additions one by one, run on the host."""
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    out = os.path.join(HERE, "bin", "float_repeat_test")
    _run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(ROOT, "tests", "cpp", "float_repeat_test.hip"),
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    return [build_library(), build_backend_plugin(),
            build_backend_plugin("libknp-mi355x-mt-backend.so", "plugin_multi_threaded.cpp"), build_cpp_tests(),
            build_float_repeat_test()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)

"""In-tree build of the gfx950 library (libczw.so) and the host tools library.

Compiles with hipcc directly (no JIT cache) so the built .so files live in the
tree and travel with it.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no implicit FMA contraction (the reference is built
#   without it; explicit fma() calls are kept where products are exact).
# -fno-gpu-flush-denormals-to-zero: binary32 denormals are IEEE, as on x86.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libczw.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "czw_kernels.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_hip()]


if __name__ == "__main__":
    for p in build_all():
        print(p)

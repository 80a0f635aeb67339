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
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "czw_kernels.hip"), os.path.join(CSRC, "czw_half.hip"),
          os.path.join(CSRC, "czw_transform.hip"),
          "-o", tmp])
    os.replace(tmp, out)
    return out


HDF5_PREFIX = os.environ.get("CZ_HDF5_PREFIX", "/opt/conda")
BINDIR = os.path.join(HERE, "bin")
TOOLS = {"hdf2cz": 1, "cz2hdf": 2, "cz2diff": 3}


def build_host():
    """libczhost.so: .cz container writer/reader (host C++, zlib, OpenMP) and the
    device-field writer that overlaps zlib with the device compressor (links
    libczw.so and the HIP runtime)."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libczhost.so")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-Wall", "-D__HIP_PLATFORM_AMD__",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(rocm, "include"),
          os.path.join(CSRC, "cz_file.cpp"), os.path.join(CSRC, "cz_lz4.cpp"), os.path.join(CSRC, "cz_pipeline.cpp"),
          "-o", out + ".tmp",
          "-L" + LIBDIR, "-lczw", "-lz", "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def have_hdf5():
    return os.path.exists(os.path.join(HDF5_PREFIX, "include", "hdf5.h")) and \
        os.path.exists(os.path.join(HDF5_PREFIX, "lib", "libhdf5.so"))


def build_tools():
    """hdf2cz / cz2hdf / cz2diff executables in cubismz_amd/bin."""
    os.makedirs(BINDIR, exist_ok=True)
    extra, libs = [], []
    if have_hdf5():
        # link libhdf5 through a private directory holding only a link to it,
        # so that the HDF5 prefix's own (older) libstdc++ is never picked up
        real = os.path.realpath(os.path.join(HDF5_PREFIX, "lib", "libhdf5.so"))
        hdir = os.path.join(LIBDIR, "hdf5")
        os.makedirs(hdir, exist_ok=True)
        soname = os.path.basename(real).split(".so")[0] + ".so." + os.path.basename(real).split(".so.")[1].split(".")[0]
        for name in ("libhdf5.so", soname):
            link = os.path.join(hdir, name)
            if os.path.lexists(link):
                os.remove(link)
            os.symlink(real, link)
        extra = ["-DCZ_HAVE_HDF5", "-isystem", os.path.join(HDF5_PREFIX, "include")]
        libs = ["-L" + hdir, "-lhdf5", "-Wl,-rpath,$ORIGIN/../lib/hdf5"]
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    outs = []
    for name, k in TOOLS.items():
        out = os.path.join(BINDIR, name)
        # host-only C++ against the HIP runtime; all device code is in libczw.so
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-DCZ_TOOL={k}", *extra,
              "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(rocm, "include"),
              os.path.join(CSRC, "cz_tools.cpp"), "-o", out + ".tmp",
              "-L" + LIBDIR, "-lczw", "-lczhost", "-lz", *libs,
              "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_checks():
    """tests/native/adapter_check: the per-block C++ adapter (include/czw_compressor.hpp)
    driven like the reference's writer and reader; run by the GPU tests."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    ndir = os.path.join(ROOT, "tests", "native")
    out = os.path.join(ndir, "adapter_check")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(rocm, "include"),
          os.path.join(ndir, "adapter_check.cpp"), "-o", out + ".tmp", "-L" + LIBDIR, "-lczw",
          "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, ndir)])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    return [build_hip(), build_host(), *build_tools(), build_checks()]


if __name__ == "__main__":
    for p in build_all():
        print(p)

"""In-tree build of the HIP library and the Python extension (gfx950).

  libpmpm_hip.so   hand-written HIP kernels + the C ABI of include/pmpm.h
  _core*.so        C++ host classes (ParticlesContainer, NodesContainer, USL,
                   MUSL, materials, boundary conditions) bound with pybind11;
                   links libpmpm_hip.so

Both are compiled by hipcc (cross-compiles without a GPU). Run:
    python pyroclastmpm_amd/build.py
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["partition.hip", "transfer.hip", "step.hip", "rigid.hip", "batch.hip", "timing.hip", "noalloc.hip"]
HOST_SOURCES = ["host.cpp", "bind.cpp", "output.cpp"]

DEVICE_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 (v_pk_*_f32) stays off on this hardware pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile_hip(src, verbose):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
    return obj


def build_library(verbose=False, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile_hip(s, verbose), HIP_SOURCES))
    out = os.path.join(HERE, "libpmpm_hip.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_extension(verbose=False, jobs=2):
    import pybind11

    os.makedirs(BUILD, exist_ok=True)
    flags = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
             "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
             "-I/opt/rocm/include"]

    def cc(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".host.o")
        _run(["g++", *flags, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(cc, HOST_SOURCES))
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_core" + ext)
    _run(["g++", "-shared", "-fPIC", *objs, "-o", out + ".tmp",
          "-L" + HERE, "-lpmpm_hip", "-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build(verbose=False):
    lib = build_library(verbose)
    ext = build_extension(verbose)
    return lib, ext


if __name__ == "__main__":
    v = "-v" in sys.argv
    if "--lib-only" in sys.argv:
        print(build_library(v))
    else:
        print(build(v))

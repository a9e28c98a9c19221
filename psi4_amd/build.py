"""In-tree build of the native libraries (gfx950).

  libp4a_ints.so       host integral engine + host FP64 linear algebra (g++, OpenMP)
  libpsi4amd_dfjk.so   HIP kernels + host orchestration + the C ABI (hipcc, gfx950; links
                       rocSOLVER/rocBLAS for the eigen fit of the metric)
  _core*.so            pybind11 module over the C ABI (g++)

Run ``python -m psi4_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("PSI4AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-fopenmp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build_ints(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, "ints", f) for f in ("ints.cpp", "ints_capi.cpp", "hostlinalg.cpp")]
    out = os.path.join(LIB, "libp4a_ints.so")
    _run(["g++", *HOST_FLAGS, "-shared", f"-I{INC}", *srcs, "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


DFJK_SOURCES = ("kernels.hip", "kernels_k.hip", "kernels_ints.hip", "kernels_ints_l0.hip", "kernels_ints_l1.hip",
                "kernels_ints_l2.hip", "kernels_ints_l3.hip", "kernels_ints_l4.hip", "kernels_schwarz.hip",
                "kernels_schwarz_f.hip", "dfjk.hip")


def build_dfjk(verbose=False):
    """Each HIP source to an object in parallel (the integral classes dominate the build), then link."""
    objdir = os.path.join(LIB, "obj")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    jobs = int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))
    procs, objs = [], []
    for f in DFJK_SOURCES:
        obj = os.path.join(objdir, f.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *HIPCC_FLAGS, "-c", f"-I{INC}", os.path.join(CSRC, "dfjk", f), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        while sum(p.poll() is None for p, _ in procs) >= max(1, jobs):
            procs[[p.poll() is None for p, _ in procs].index(True)][0].wait()
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), cmd))
    for p, cmd in procs:
        out_txt = p.communicate()[0]
        if p.returncode != 0:
            sys.stderr.write(out_txt)
            raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-3]}")
    out = os.path.join(LIB, "libpsi4amd_dfjk.so")
    # rocSOLVER (dsyevd) for psi4's eigen fit of the metric; rocBLAS carries its handle
    _run([hipcc, *HIPCC_FLAGS, "-shared", *objs, f"-L{LIB}", "-lp4a_ints", f"-L{ROCM}/lib", "-lrocsolver",
          "-lrocblas", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib", "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_pybind(verbose=False):
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_core" + suffix)
    src = os.path.join(CSRC, "pybind", "core.cpp")
    pyinc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{INC}", f"-I{pybind11.get_include()}", f"-I{pyinc}",
          f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__", src, f"-L{LIB}", "-lpsi4amd_dfjk", "-lp4a_ints",
          "-Wl,-rpath,$ORIGIN/lib", "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build(verbose=False):
    build_ints(verbose)
    build_dfjk(verbose)
    build_pybind(verbose)


if __name__ == "__main__":
    build(verbose=True)

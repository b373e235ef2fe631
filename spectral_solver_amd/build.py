"""Build libspectral_hip.so (gfx950) and the taylor_green CLI in-tree.

    python -m spectral_solver_amd.build        # or spectral_solver_amd.build.build()

Objects go to spectral_solver_amd/_build/; the shared library and the CLI are
written next to this file so that they travel with the source tree.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libspectral_hip.so")
CLI = os.path.join(HERE, "taylor_green")
CPP_TESTS = os.path.join(HERE, "spectral_ns_tests")
NS_PARITY = os.path.join(HERE, "spectral_ns_parity")
MPI_CHECK = os.path.join(HERE, "mpi_comm_check")
INDEX_TESTS = os.path.join(HERE, "index_tests")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("SPH_HOST_CXX", "/opt/rocm/bin/amdclang++")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INCLUDE, "-I" + CSRC]
UNITS = ["passes_f32.hip", "passes_f64.hip", "passes_f32_b.hip", "passes_f64_b.hip", "passes_vmax.hip", "library.hip", "noalloc.hip", "passes_ens.hip"]


def find_mpi():
    """(include dir, library dir) of an MPI installation with <mpi.h> and
    libmpi, or None: $MPI_HOME, then the conda MPICH, then the Debian MPICH
    and Open MPI locations. Nothing is fetched."""
    homes = [os.environ.get("MPI_HOME"), "/opt/conda", "/usr"]
    pairs = [(os.path.join(h, "include"), os.path.join(h, "lib")) for h in homes if h]
    pairs += [("/usr/include/x86_64-linux-gnu/mpich", "/usr/lib/x86_64-linux-gnu"),
              ("/usr/lib/x86_64-linux-gnu/openmpi/include", "/usr/lib/x86_64-linux-gnu/openmpi/lib")]
    for inc, lib in pairs:
        if os.path.exists(os.path.join(inc, "mpi.h")) and any(
                os.path.exists(os.path.join(lib, f)) for f in ("libmpi.so", "libmpi.so.12", "libmpi.so.40")):
            return inc, lib
    return None


def _run(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        errors = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append((c, "\n".join(errors)))
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + e for c, e in failed)
        raise RuntimeError("build failed:\n" + msg)


def build_id() -> str:
    """Hash of the library's sources and public headers: the autotuner's
    cache file (SPH_TUNE_CACHE) keys its entries by it."""
    h = hashlib.sha1()
    for d in (CSRC, INCLUDE):
        for name in sorted(os.listdir(d)):
            if name.endswith((".hip", ".hpp", ".h")):
                with open(os.path.join(d, name), "rb") as f:
                    h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    bid = '-DSPH_BUILD_ID="' + build_id() + '"'
    for u in UNITS:
        obj = os.path.join(BUILD, u.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, bid, "-c", os.path.join(CSRC, u), "-o", obj])
    if verbose:
        print("[build] compiling", ", ".join(UNITS), "for", ARCH, flush=True)
    _run(cmds)
    tmp = LIB + ".tmp"
    _run([[HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]])
    os.replace(tmp, LIB)
    # host programs over the C ABI (plain C++: no HIP code, no HIP headers):
    # the CLI, the reference-style C++ test, the legacy-API parity driver
    host = [CXX, "-O2", "-std=c++17", "-I" + INCLUDE, "-Wl,-rpath,$ORIGIN", "-L" + HERE, "-lspectral_hip"]
    _run([[*host[:4], os.path.join(CSRC, "main_taylor_green.cpp"), "-o", CLI, *host[4:]],
          [*host[:4], os.path.join(CSRC, "spectral_ns_tests.cpp"), "-o", CPP_TESTS, *host[4:]],
          [*host[:4], os.path.join(CSRC, "spectral_ns_parity.cpp"), "-o", NS_PARITY, *host[4:]],
          # host checks of the kernels' index maps (no GPU needed to run)
          [HIPCC, *HIP_FLAGS, os.path.join(CSRC, "index_tests.hip"), "-o", INDEX_TESTS]])
    # the backend with a real MPI communicator, where an MPI is installed
    mpi = find_mpi()
    if mpi:
        inc, lib = mpi
        # the MPI library directory may hold an older libstdc++ than the HIP
        # runtime needs: the compiler's own libstdc++ directory comes first
        stdcxx = subprocess.run([CXX, "-print-file-name=libstdc++.so.6"], capture_output=True, text=True).stdout
        first = ["-Wl,-rpath," + os.path.dirname(stdcxx.strip())] if os.path.isabs(stdcxx.strip()) else []
        _run([[*host[:4], "-I" + inc, os.path.join(CSRC, "mpi_comm_check.cpp"), "-o", MPI_CHECK, *host[4:],
               *first, "-L" + lib, "-Wl,-rpath," + lib, "-lmpi"]])
    elif os.path.exists(MPI_CHECK):
        os.remove(MPI_CHECK)
    if verbose:
        print("[build] wrote", LIB, "| MPI:", mpi[0] if mpi else "not found", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

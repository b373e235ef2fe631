"""The C++ backend class's communicator type (include/spectral_hip.hpp),
compiled on the host without a GPU (syntax only):

  - without MPI, Comm is the one-rank stand-in;
  - with MPICH's <mpi.h> (found by the build's probe), Comm is MPI_Comm, an
    int there;
  - with a pointer-typed MPI_Comm, as Open MPI declares it (a stub header
    written here declaring only what the class uses), the class and the
    legacy solver compile unchanged, constructed with MPI_COMM_SELF and
    handing fft().comm() to MPI_Allreduce (src/fftw_fft.hpp:15,26;
    src/spectral_solver.cpp:317,334).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
CXX = shutil.which("g++")

USE = r"""
#include <type_traits>
int use_backend() {
    spectral::SpectralNavierStokes* s = nullptr;
    double local = 1.0, global = 0.0;
    if (s) MPI_Allreduce(&local, &global, 1, MPI_DOUBLE, MPI_SUM, s->fft().comm());
    spectral::Comm c = MPI_COMM_SELF;
    (void)c;
    return 0;
}
"""

POINTER_MPI_H = r"""
#pragma once
#define MPI_VERSION 3
struct stub_communicator;
typedef struct stub_communicator* MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;
extern struct stub_communicator stub_comm_self;
#define MPI_COMM_SELF (&stub_comm_self)
#define MPI_DOUBLE 1
#define MPI_SUM 1
#define MPI_MAX 2
extern "C" int MPI_Initialized(int*);
extern "C" int MPI_Comm_size(MPI_Comm, int*);
extern "C" int MPI_Allreduce(const void*, void*, int, MPI_Datatype, MPI_Op, MPI_Comm);
"""


def compile_only(tmp_path, src, extra=()):
    f = tmp_path / "t.cpp"
    f.write_text(src)
    r = subprocess.run([CXX, "-std=c++17", "-fsyntax-only", "-I" + INCLUDE, *extra, str(f)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_without_mpi_comm_is_the_stand_in(tmp_path):
    compile_only(tmp_path, '#include "spectral_hip.hpp"\n#include <type_traits>\n'
                           'static_assert(std::is_same_v<spectral::Comm, int>);\n'
                           'spectral::Comm c = spectral::kCommSelf;\n')


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_with_mpich_comm_is_mpi_comm(tmp_path):
    import sys
    sys.path.insert(0, ROOT)
    from spectral_solver_amd.build import find_mpi
    mpi = find_mpi()
    if not mpi:
        pytest.skip("no MPI installation")
    compile_only(tmp_path, '#include <mpi.h>\n#include "spectral_hip.hpp"\n'
                           'static_assert(std::is_same_v<spectral::Comm, MPI_Comm>);\n' + USE, ["-I" + mpi[0]])


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_with_pointer_typed_communicator(tmp_path):
    stub = tmp_path / "stubmpi"
    stub.mkdir()
    (stub / "mpi.h").write_text(POINTER_MPI_H)
    compile_only(tmp_path, '#include <mpi.h>\n#include "spectral_hip.hpp"\n'
                           'static_assert(std::is_pointer_v<spectral::Comm>);\n'
                           'static_assert(std::is_same_v<spectral::Comm, MPI_Comm>);\n' + USE, ["-I" + str(stub)])
    # the build-time switch without an explicit include
    compile_only(tmp_path, '#include "spectral_hip.hpp"\n'
                           'static_assert(std::is_same_v<spectral::Comm, MPI_Comm>);\n' + USE,
                 ["-DSPECTRAL_HIP_WITH_MPI", "-I" + str(stub)])

"""The host programs the build writes (CLI, C++ known-answer test, legacy-API
parity driver, the MPI communicator check) resolve every shared library they
need on this host: a library directory earlier in a program's search path can
hold an older libstdc++ than the HIP runtime requires, which only shows when
the program starts (no GPU needed to see it)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "spectral_solver_amd")
PROGRAMS = ["taylor_green", "spectral_ns_tests", "spectral_ns_parity", "mpi_comm_check", "libspectral_hip.so"]


@pytest.mark.skipif(shutil.which("ldd") is None, reason="no ldd")
@pytest.mark.parametrize("name", PROGRAMS)
def test_program_resolves_its_libraries(name):
    path = os.path.join(PKG, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built")
    out = subprocess.run(["ldd", path], capture_output=True, text=True, timeout=60)
    text = out.stdout + out.stderr
    assert "not found" not in text, text

"""Host checks of the pass kernels' index maps (LDS layouts, butterfly and
slot orders, the LDS stash of the fused forward-z kernel, the ky-blocked
spectral layout's mode map and the blocked y tiles' value order): each must
be a bijection and fit the kernel's LDS. The checks are compiled from the same
headers as the kernels (spectral_solver_amd/csrc/index_tests.hip) and run on
the host."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spectral_solver_amd", "index_tests")


@pytest.mark.skipif(not os.path.exists(EXE), reason="index_tests not built (python -m spectral_solver_amd.build)")
def test_index_maps_are_bijections():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "all index-map checks passed" in r.stdout

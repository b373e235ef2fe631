"""Host-side checks of the C ABI (no GPU needed, no compute calls).

The library must load, export every function include/spectral_hip.h declares,
and reject invalid input with the reference's messages before touching a device
(src/ava_taylor_green_impl.hpp:704-709, src/fftw_fft.cpp:12-24).
"""
import ctypes
import os
import re

import numpy as np
import pytest

import spectral_solver_amd as sph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spectral_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sph.load_library()
    names = declared_functions()
    assert len(names) >= 20
    bound = {s[0] for s in sph.SYMBOLS}
    for name in names:
        assert hasattr(lib, name), f"{name} not exported"
        assert name in bound, f"{name} has no Python binding"


def test_integration_guide_names_every_entry_point():
    """The binding guide a maintainer follows names every declared call."""
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    missing = [name for name in declared_functions() if not re.search(rf"\b{name}\b", guide)]
    assert not missing, f"INTEGRATION.md does not name {missing}"


def test_integration_guide_c_examples_compile(tmp_path):
    """The guide's C examples are statement blocks against the header: each
    compiles as a function body (C11, every warning an error)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", guide, flags=re.S)
    assert blocks
    for i, body in enumerate(blocks):
        src = tmp_path / f"example_{i}.c"
        src.write_text('#include "spectral_hip.h"\nvoid example(void) {\n' + body + "}\n")
        r = subprocess.run([gcc, "-std=c11", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                            str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_integration_guide_cpp_examples_compile(tmp_path):
    """The guide's two C++ examples against include/spectral_hip.hpp with a
    plain host compiler: the backend switch as a translation unit, the step
    API as a function body (its free names t, init_seconds, run_seconds as
    parameters)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", guide, flags=re.S)
    assert len(blocks) == 2, "update this test with the guide's C++ examples"
    backend, step = blocks
    lines = step.splitlines()
    includes = [l for l in lines if l.startswith("#include")]
    body = [l for l in lines if not l.startswith("#include")]
    sources = {
        "backend.cpp": backend,
        "step.cpp": "\n".join(includes) + "\nvoid example(double t, double init_seconds, double run_seconds) {\n"
        + "\n".join(body) + "\n}\n",
    }
    for name, text in sources.items():
        src = tmp_path / name
        src.write_text(text)
        r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, f"{name}: {r.stderr}"


def test_supported_lengths():
    lib = sph.load_library()
    for n in (8, 12, 16, 24, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512, 640, 768, 1024):
        assert lib.sph_supported_length(n) == 1
    for n in (0, 1, 7, 100, 2048):
        assert lib.sph_supported_length(n) == 0


def test_radix_plans_of_the_z_type_passes():
    """Both radix orders of every compiled length factor the length into the
    butterflies the kernels have (2, 3, 4, 5, 6, 8, 16); the alternative
    order is the stated permutation of the default's stages (or, for
    all-radix-8 plans, the first two as 4*4*4; for 1024 the three-stage plan
    8*16*8), and differs from it wherever the autotuner offers it."""
    lengths = (8, 12, 16, 24, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512, 640, 768, 1024)
    expect = {1024: ([8, 4, 4, 8], [8, 16, 8]), 512: ([8, 8, 8], [4, 4, 4, 8]), 256: ([8, 4, 8], [4, 8, 8]),
              128: ([8, 4, 4], [4, 4, 8]), 96: ([4, 4, 6], [6, 4, 4]), 80: ([4, 4, 5], [5, 4, 4]),
              64: ([8, 8], [8, 8]), 48: ([4, 4, 3], [3, 4, 4]), 32: ([8, 4], [4, 8])}
    for n in lengths:
        a, b = sph.radix_plan(n, 0), sph.radix_plan(n, 1)
        for plan in (a, b):
            assert set(plan) <= {2, 3, 4, 5, 6, 8, 16}
            assert int(np.prod(plan)) == n, (n, plan)
        if n in expect:
            assert (a, b) == expect[n]
        if not all(r == 8 for r in a) and n != 1024:
            assert sorted(a) == sorted(b), (n, a, b)
    assert sph.radix_plan(100, 0) == [] and sph.radix_plan(2048, 1) == []
    with pytest.raises(RuntimeError, match="radix order"):
        sph.radix_plan(64, 2)


def _solver_error(n, re_, prec=sph.FP64):
    lib = sph.load_library()
    h = ctypes.c_void_p()
    rc = lib.sph_solver_create(n, re_, prec, ctypes.byref(h))
    assert rc != 0
    return lib.sph_last_error().decode()


def test_solver_argument_errors_match_reference():
    assert _solver_error(0, 1600.0) == "grid size must be positive"
    assert _solver_error(-4, 1600.0) == "grid size must be positive"
    assert _solver_error(64, 0.0) == "Reynolds number must be positive"
    assert _solver_error(64, -1.0) == "Reynolds number must be positive"
    assert "unsupported transform length" in _solver_error(100, 1600.0)
    assert "precision" in _solver_error(64, 1600.0, 7)


def test_fft_argument_errors_match_reference():
    with pytest.raises(RuntimeError, match="all grid sizes must be greater than one"):
        sph.HipFft3D((1, 8, 8), (1.0, 1.0, 1.0))
    with pytest.raises(RuntimeError, match="all domain lengths must be positive"):
        sph.HipFft3D((8, 8, 8), (1.0, 0.0, 1.0))
    with pytest.raises(RuntimeError, match="unsupported transform length"):
        sph.HipFft3D((8, 8, 100), (1.0, 1.0, 1.0))


def test_driver_argument_errors_match_reference():
    with pytest.raises(RuntimeError, match="--dt, --cfl, and --output-dt must be positive"):
        sph.run_taylor_green(sph.RunConfig(n=16, dt=0.0))
    with pytest.raises(RuntimeError, match="--dt, --cfl, and --output-dt must be positive"):
        sph.run_taylor_green(sph.RunConfig(n=16, cfl=-1.0))


def test_null_handles_are_rejected():
    lib = sph.load_library()
    assert lib.sph_solver_step_rk3(None, 1e-3) != 0
    assert "null handle" in lib.sph_last_error().decode()
    assert lib.sph_fft_forward(None, None, None, None) != 0


@pytest.mark.skipif(sph.device_count() > 0, reason="checks the no-GPU failure path")
def test_no_device_fails_loudly():
    with pytest.raises(RuntimeError, match="no HIP device"):
        sph.TaylorGreenSolver(16, 1600.0)


def test_pass_count_matches_header():
    """The timing calls write one double per pass; the binding's arrays must
    be that long."""
    text = open(HEADER).read()
    m = re.search(r"written to ms\[(\d+)\]", text)
    assert m and int(m.group(1)) == len(sph._PASS_NAMES)

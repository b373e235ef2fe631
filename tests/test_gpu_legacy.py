"""The legacy value-semantics API (spectral::SpectralNavierStokes over the GPU
FFT backend, include/spectral_hip.hpp) against the CPU restatement of the
reference's src/spectral_solver.cpp (oracle/legacy.hpp): non-cubic grids and
box lengths, phase-shift dealiasing on and off, a seeded random physical
velocity (so that every stored mode, the Nyquist planes included, is
exercised), 3 RK3 steps. The C++ driver spectral_ns_parity runs the shim;
its nonlinear term of the initial state, its final state and its diagnostics
after every step are compared with the oracle's at 1e-12 relative L2.

Also: the backend with a real MPI communicator (mpi_comm_check, built where
the build finds <mpi.h>) hands fft().comm() to MPI_Allreduce.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "spectral_solver_amd", "spectral_ns_parity")
MPI_EXE = os.path.join(ROOT, "spectral_solver_amd", "mpi_comm_check")
TWO_PI = 2.0 * math.pi


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.skipif(not os.path.exists(EXE), reason="spectral_ns_parity not built (python -m spectral_solver_amd.build)")
@pytest.mark.parametrize("shape,lengths", [((16, 12, 8), (TWO_PI, 2 * TWO_PI, 0.5 * TWO_PI)),
                                           ((24, 16, 32), (1.0, 2.0, 3.0))])
@pytest.mark.parametrize("dealias", [True, False])
def test_legacy_api_matches_oracle(tmp_path, shape, lengths, dealias):
    nx, ny, nz = shape
    nu, dt, steps = 0.05, 1e-3, 3
    print("case legacy", shape, lengths, "dealias", dealias, flush=True)
    rng = np.random.default_rng(nx * 10000 + ny * 100 + nz + (1 if dealias else 0))
    phys = [rng.standard_normal((nz, ny, nx)) for _ in range(3)]
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    np.stack(phys).astype(np.float64).tofile(inp)
    args = [EXE, str(nx), str(ny), str(nz)] + [repr(v) for v in lengths] + [
        repr(nu), "1" if dealias else "0", str(steps), repr(dt), str(inp), str(out)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    modes = (nx // 2 + 1) * ny * nz
    raw = np.fromfile(out, np.float64)
    assert raw.size == 5 * (steps + 1) + 2 * 3 * 2 * modes
    diag = raw[:5 * (steps + 1)].reshape(steps + 1, 5)
    fields = raw[5 * (steps + 1):].view(np.complex128).reshape(2, 3, nx // 2 + 1, nz, ny)
    rhs0, state = fields[0], fields[1]
    rd, rr, rs = oracle.legacy_run(shape, lengths, nu, dealias, steps, dt, phys)
    ref_rhs = np.stack([oracle.c2c_to_half(x) for x in rr])
    ref_state = np.stack([oracle.c2c_to_half(x) for x in rs])
    e_rhs, e_state = rel(rhs0, ref_rhs), rel(state, ref_state)
    print(f"  nonlinear term rel L2 {e_rhs:.2e}, state after {steps} steps {e_state:.2e}", flush=True)
    assert e_rhs < 1e-12 and e_state < 1e-12
    for col, name in ((1, "energy"), (2, "dissipation"), (3, "enstrophy")):
        err = float(np.max(np.abs(diag[:, col] - rd[:, col]) / np.abs(rd[:, col])))
        print(f"  {name}: max rel err {err:.2e}", flush=True)
        assert err < 1e-12, name
    # divergence after projection is rounding on both sides
    scale = np.linalg.norm(ref_state) * 2 * math.pi * max(n / L for n, L in zip(shape, lengths))
    assert np.all(diag[:, 4] < 1e-12 * scale) and np.all(rd[:, 4] < 1e-12 * scale)


@pytest.mark.skipif(not os.path.exists(MPI_EXE), reason="no MPI found at build time (mpi_comm_check not built)")
def test_backend_with_mpi_communicator():
    print("case mpi_comm_check", flush=True)
    r = subprocess.run([MPI_EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "comm_size=1" in r.stdout

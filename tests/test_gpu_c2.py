"""Config C2: Taylor-Green 256^3 FP64, nu = 1/1600, dt = 1e-3, 500 fixed RK3 steps,
against the CPU oracle's run of the same configuration (fixture written by
tests/golden/make_c2_fixture.py): kinetic energy and dissipation after every
step, and at step 500 the low modes |m| <= 4, the shell spectrum and the
squared norm of the state, all at 1e-12 relative (shells: relative L2 over
the spectrum, which spans many decades).
"""
import os
import time

import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tg256_fp64_500.npz")


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="fixture not generated (tests/golden/make_c2_fixture.py)")
def test_config_256cube_fp64_500_steps_matches_oracle():
    g = np.load(FIXTURE)
    n, dt, steps, K = int(g["n"]), float(g["dt"]), int(g["steps"]), int(g["low_k"])
    print("case C2", n, steps, flush=True)
    s = sph.TaylorGreenSolver(n, float(g["reynolds"]), "fp64")
    d = s.diagnostics(0.0)
    worst_e = abs(d.kinetic_energy - g["energy"][0]) / g["energy"][0]
    worst_eps = abs(d.dissipation - g["dissipation"][0]) / g["dissipation"][0]
    t0 = time.time()
    for k in range(steps):
        s.step_rk3(dt)
        d = s.diagnostics((k + 1) * dt)
        worst_e = max(worst_e, abs(d.kinetic_energy - g["energy"][k + 1]) / g["energy"][k + 1])
        worst_eps = max(worst_eps, abs(d.dissipation - g["dissipation"][k + 1]) / g["dissipation"][k + 1])
        if k % 100 == 99:
            print(f"  step {k + 1}: worst rel E {worst_e:.2e}, eps {worst_eps:.2e}, {time.time() - t0:.1f} s",
                  flush=True)
    assert worst_e < 1e-12 and worst_eps < 1e-12
    idx = [m % n for m in range(-K, K + 1)]
    u = [s.u_hat(c) for c in range(3)]
    for c in range(3):
        low = u[c][: K + 1][:, idx][:, :, idx]
        ref = g[f"low{c}"]
        assert np.linalg.norm(low - ref) <= 1e-12 * np.linalg.norm(ref), c
    full = [oracle.half_to_c2c(x) for x in u]
    norm2 = sum(float((np.abs(x) ** 2).sum()) for x in full)
    assert abs(norm2 - float(g["norm2"])) <= 1e-12 * float(g["norm2"])
    m = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    kk = np.rint(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)).astype(int)
    e = sum(np.abs(x) ** 2 for x in full)
    shells = np.bincount(kk.ravel(), weights=e.ravel(), minlength=len(g["shells"]))
    assert np.linalg.norm(shells - g["shells"]) <= 1e-12 * np.linalg.norm(g["shells"])

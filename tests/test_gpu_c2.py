"""Config C2: Taylor-Green 256^3 FP64, nu = 1/1600, dt = 1e-3, 500 fixed RK3 steps,
against the CPU oracle's run of the same configuration (fixture written by
tests/golden/make_c2_fixture.py):

  - kinetic energy and dissipation after every step, 1e-12 relative;
  - at steps 100, 200, ..., 500 the whole state through 32 seeded random
    linear sketches per component (tests/sketch.py; every stored mode enters
    every sketch with its own weight, so a phase, sign or index error in any
    mode moves them): relative L2 of the sketch vectors, 1e-12;
  - at step 500 the block of modes |mx|, |my|, |mz| <= 16, the squared norm
    and the shell spectrum, 1e-12 relative.
"""
import os
import time

import numpy as np
import pytest

import oracle
import sketch
import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tg256_fp64_500.npz")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="fixture not generated (tests/golden/make_c2_fixture.py)")
def test_config_256cube_fp64_500_steps_matches_oracle():
    g = np.load(FIXTURE)
    n, dt, steps = int(g["n"]), float(g["dt"]), int(g["steps"])
    checks = [int(k) for k in g["sketch_steps"]]
    ref_sk = g["sketches"]  # [checkpoint][component][sketch]
    print("case C2", n, steps, "sketches at", checks, flush=True)
    w = sketch.weights(n * n * (n // 2 + 1), ref_sk.shape[2], int(g["sketch_seed"]))
    s = sph.TaylorGreenSolver(n, float(g["reynolds"]), "fp64")
    d = s.diagnostics(0.0)
    worst_e = abs(d.kinetic_energy - g["energy"][0]) / g["energy"][0]
    worst_eps = abs(d.dissipation - g["dissipation"][0]) / g["dissipation"][0]
    worst_sk = 0.0
    t0 = time.time()
    for k in range(steps):
        s.step_rk3(dt)
        d = s.diagnostics((k + 1) * dt)
        worst_e = max(worst_e, abs(d.kinetic_energy - g["energy"][k + 1]) / g["energy"][k + 1])
        worst_eps = max(worst_eps, abs(d.dissipation - g["dissipation"][k + 1]) / g["dissipation"][k + 1])
        if k + 1 in checks:
            i = checks.index(k + 1)
            got = np.stack([sketch.sketch(w, s.u_hat(c)) for c in range(3)])
            err = rel(got, ref_sk[i])
            worst_sk = max(worst_sk, err)
            print(f"  step {k + 1}: sketch rel err {err:.2e}; worst rel E {worst_e:.2e}, eps {worst_eps:.2e}, "
                  f"{time.time() - t0:.1f} s", flush=True)
            assert err <= 1e-12, (k + 1, err)
    assert worst_e < 1e-12 and worst_eps < 1e-12
    u = [s.u_hat(c) for c in range(3)]
    K = int(g["block_k"])
    block = np.stack([sketch.low_block(x, K) for x in u])
    err_block = rel(block, g["block"])
    print(f"  step {steps}: modes |m| <= {K} rel err {err_block:.2e}", flush=True)
    assert err_block <= 1e-12
    full = [oracle.half_to_c2c(x) for x in u]
    norm2 = sum(float((np.abs(x) ** 2).sum()) for x in full)
    assert abs(norm2 - float(g["norm2"])) <= 1e-12 * float(g["norm2"])
    m = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    kk = np.rint(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)).astype(int)
    e = sum(np.abs(x) ** 2 for x in full)
    shells = np.bincount(kk.ravel(), weights=e.ravel(), minlength=len(g["shells"]))
    assert np.linalg.norm(shells - g["shells"]) <= 1e-12 * np.linalg.norm(g["shells"])
    s.close()

"""The x pass's prefetch of its second and third velocity pairs into LDS
(global-to-LDS loads while the previous pair's inverse transform runs on the
other tile with LDS-only barriers). It runs in FP32 in the line layout at the
lengths whose padded 16-line tiles hold the rows (256, 512); SPH_TILE_VARIANT=0
forces that tile. Against the same step with the prefetch-free x-pass forms
(variant 1: half-width register-direct tiles; variant 2: LDS-staged tiles) on
a seeded random solenoidal state: the same arithmetic up to multiply-add
contraction, so the states agree to FP32 rounding; and against the FP64 path.
"""
import numpy as np
import pytest

import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu


def run(monkeypatch, n, precision, variant, steps=2, dt=1e-3):
    monkeypatch.setenv("SPH_LAYOUT", "0")
    monkeypatch.setenv("SPH_TILE_VARIANT", variant)
    s = sph.TaylorGreenSolver(n, 400.0, precision)
    assert s.tuning()["x_products"][0] == int(variant)
    u0 = sph.random_solenoidal_field(n, 77)
    for c in range(3):
        s.set_u_hat(c, u0[c])
    for _ in range(steps):
        s.step_rk3(dt)
    u = [s.u_hat(c) for c in range(3)]
    s.close()
    return u


def rel(a, b):
    num = sum(np.linalg.norm(x - y) ** 2 for x, y in zip(a, b))
    den = sum(np.linalg.norm(y) ** 2 for y in b)
    return float(np.sqrt(num / den))


@pytest.mark.parametrize("n", [256, 512])
def test_prefetching_x_pass_agrees_with_the_other_forms(monkeypatch, n):
    print("case x prefetch", n, flush=True)
    pre = run(monkeypatch, n, "fp32", "0")
    for v in ("1", "2"):
        other = run(monkeypatch, n, "fp32", v)
        err = rel(pre, other)
        print(f"  variant 0 (prefetch) vs {v}: rel L2 {err:.2e}", flush=True)
        assert err < 1e-6
    ref = run(monkeypatch, n, "fp64", "0")
    err = rel(pre, ref)
    print(f"  FP32 with prefetch vs FP64: rel L2 {err:.2e}", flush=True)
    assert err < 2e-6

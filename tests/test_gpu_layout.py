"""The transposed work layout (k_plines_t and the x pass on [z][y][kx] pair
fields, SPH_LAYOUT=1) against the line layout (SPH_LAYOUT=0): the same
arithmetic on other addresses, so the states agree to rounding, and against
the CPU oracle in FP64. Also the adaptive step's fused velocity maximum in
that layout, and the layout the autotuner reports.
"""
import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu


def run(monkeypatch, layout, n, precision, steps, dt, variant=None):
    monkeypatch.setenv("SPH_LAYOUT", layout)
    if variant is not None:
        monkeypatch.setenv("SPH_TILE_VARIANT", variant)
    s = sph.TaylorGreenSolver(n, 1600.0, precision)
    assert s.layout()[0] == (layout == "1")
    for _ in range(steps):
        s.step_rk3(dt)
    u = [s.u_hat(c) for c in range(3)]
    d = s.diagnostics(steps * dt)
    s.close()
    return u, d


def rel(a, b):
    num = sum(np.linalg.norm(x - y) ** 2 for x, y in zip(a, b))
    den = sum(np.linalg.norm(y) ** 2 for y in b)
    return float(np.sqrt(num / den))


@pytest.mark.parametrize("n,precision,variant,tol", [(256, "fp64", None, 1e-14), (256, "fp64", "1", 1e-14),
                                                     (256, "fp64", "2", 1e-14), (256, "fp64", "3", 1e-14),
                                                     (512, "fp32", None, 1e-6),
                                                     (256, "fp32", "1", 1e-6)])
def test_transposed_layout_agrees_with_line_layout(monkeypatch, n, precision, variant, tol):
    print("case layout", n, precision, "variant", variant, flush=True)
    a, da = run(monkeypatch, "1", n, precision, 2, 1e-3, variant)
    b, db = run(monkeypatch, "0", n, precision, 2, 1e-3, variant)
    err = rel(a, b)
    print(f"  transposed vs line layout: rel L2 {err:.2e}, E {da.kinetic_energy:.15f} / {db.kinetic_energy:.15f}",
          flush=True)
    assert err <= tol
    assert abs(da.dissipation - db.dissipation) <= tol * db.dissipation


def test_transposed_layout_matches_oracle_fp64(monkeypatch):
    n, dt = 256, 2e-3
    print("case layout oracle", n, flush=True)
    u, _ = run(monkeypatch, "1", n, "fp64", 1, dt)
    o = oracle.TaylorGreenOracle(n, 1600.0)
    o.step_rk3(dt)
    ref = [oracle.c2c_to_half(o.u_hat(c)) for c in range(3)]
    err = rel(u, ref)
    print(f"  rel L2 vs oracle {err:.2e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_adaptive_step_in_transposed_layout(monkeypatch, precision, tol):
    print("case layout adaptive", precision, flush=True)
    monkeypatch.setenv("SPH_LAYOUT", "1")
    s = sph.TaylorGreenSolver(256, 1600.0, precision)
    assert s.layout()[0]
    for _ in range(2):
        v_ref = s.max_velocity()
        vmax, dt, fused = s.step_adaptive(0.5, 0.5)
        assert fused and abs(vmax - v_ref) <= tol * v_ref
    s.close()


def test_autotuner_reports_both_layouts():
    print("case layout report", flush=True)
    s = sph.TaylorGreenSolver(256, 1600.0, "fp32")
    xt, ms = s.layout()
    print("  transposed kept:", xt, "ms line / transposed:", ms, flush=True)
    assert ms[0] > 0 and ms[1] > 0
    assert xt == (ms[1] < ms[0])
    s.close()

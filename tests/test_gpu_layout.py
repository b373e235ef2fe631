"""The work layouts against each other and against the CPU oracle:
SPH_LAYOUT=1, the transposed layout (k_plines_t and the x pass on [z][y][kx]
pair fields), and SPH_LAYOUT=2, the ky-blocked layout (state and spectral pair
fields in blocks of the z tile's ky count, k_plines_b between them and the x
pass's fields in [z][kx][y], or [kx][z][y] with SPH_X_ZMAJOR=0), against
SPH_LAYOUT=0, the line layout: the same arithmetic on other addresses, so the states agree to rounding, and the FP64
states agree with the oracle. Also state access and a random state with
non-Hermitian Nyquist modes in the ky-blocked layout, an ensemble of two
members in it, the adaptive step's fused velocity maximum in both, and the
layout the autotuner reports.
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
    assert s.layout()[0] == int(layout)
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


@pytest.mark.parametrize("layout", ["1", "2"])
@pytest.mark.parametrize("n,precision,variant,tol", [(256, "fp64", None, 1e-14), (256, "fp64", "1", 1e-14),
                                                     (256, "fp64", "2", 1e-14), (256, "fp64", "3", 1e-14),
                                                     (512, "fp32", None, 1e-6),
                                                     (256, "fp32", "1", 1e-6)])
def test_layout_agrees_with_line_layout(monkeypatch, layout, n, precision, variant, tol):
    print("case layout", layout, n, precision, "variant", variant, flush=True)
    a, da = run(monkeypatch, layout, n, precision, 2, 1e-3, variant)
    b, db = run(monkeypatch, "0", n, precision, 2, 1e-3, variant)
    err = rel(a, b)
    print(f"  layout {layout} vs line layout: rel L2 {err:.2e}, E {da.kinetic_energy:.15f} / "
          f"{db.kinetic_energy:.15f}", flush=True)
    assert err <= tol
    assert abs(da.dissipation - db.dissipation) <= tol * db.dissipation
    assert abs(da.kinetic_energy - db.kinetic_energy) <= tol * db.kinetic_energy


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-14), ("fp32", 1e-6)])
def test_blocked_layout_with_kx_major_x_fields(monkeypatch, precision, tol):
    """SPH_X_ZMAJOR=0: the ky-blocked layout with the x pass's fields in
    [kx][z][y] instead of [z][kx][y] (other strides in the x pass, its
    diagnostics transforms and the y passes' line side)."""
    print("case blocked layout, kx-major x fields", precision, flush=True)
    monkeypatch.setenv("SPH_X_ZMAJOR", "0")
    a, da = run(monkeypatch, "2", 256, precision, 2, 1e-3)
    monkeypatch.delenv("SPH_X_ZMAJOR")
    b, db = run(monkeypatch, "2", 256, precision, 2, 1e-3)
    c, dc = run(monkeypatch, "0", 256, precision, 2, 1e-3)
    for x, dx in ((a, da), (b, db)):
        err = rel(x, c)
        print(f"  vs line layout: rel L2 {err:.2e}", flush=True)
        assert err <= tol
        assert abs(dx.dissipation - dc.dissipation) <= tol * dc.dissipation


@pytest.mark.parametrize("layout", ["1", "2"])
def test_layout_matches_oracle_fp64(monkeypatch, layout):
    n, dt = 256, 2e-3
    print("case layout oracle", layout, n, flush=True)
    u, _ = run(monkeypatch, layout, n, "fp64", 1, dt)
    o = oracle.TaylorGreenOracle(n, 1600.0)
    o.step_rk3(dt)
    ref = [oracle.c2c_to_half(o.u_hat(c)) for c in range(3)]
    err = rel(u, ref)
    print(f"  rel L2 vs oracle {err:.2e}", flush=True)
    assert err < 1e-12


def random_state(n, seed):
    """A random state with every half-spectrum mode set (no projection, no
    Hermitian symmetry on the kx = 0 and n/2 planes): the step projects and
    truncates it, and the state access must put every value where the
    kernels look for it."""
    rng = np.random.default_rng(seed)
    shape = (n // 2 + 1, n, n)
    return [(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 1e-3 for _ in range(3)]


def test_blocked_layout_state_access_and_random_state(monkeypatch):
    n, dt = 256, 1e-3
    print("case blocked layout random state", n, flush=True)
    u0 = random_state(n, 7)
    out = {}
    for layout in ("2", "0"):
        monkeypatch.setenv("SPH_LAYOUT", layout)
        s = sph.TaylorGreenSolver(n, 1600.0, "fp64")
        assert s.layout()[0] == int(layout)
        for c in range(3):
            s.set_u_hat(c, u0[c])
        # state access round trip, bit for bit
        for c in range(3):
            assert np.array_equal(s.u_hat(c), u0[c])
        d0 = s.diagnostics(0.0)
        for _ in range(2):
            s.step_rk3(dt)
        out[layout] = ([s.u_hat(c) for c in range(3)], d0, s.diagnostics(2 * dt), s.max_velocity())
        s.close()
    (ua, d0a, da, va), (ub, d0b, db, vb) = out["2"], out["0"]
    err = rel(ua, ub)
    print(f"  blocked vs line: rel L2 {err:.2e}; divergence {d0a.max_spectral_divergence:.6e} / "
          f"{d0b.max_spectral_divergence:.6e}", flush=True)
    assert err <= 1e-14
    for x, y in ((d0a, d0b), (da, db)):
        assert abs(x.kinetic_energy - y.kinetic_energy) <= 1e-14 * y.kinetic_energy
        assert abs(x.enstrophy - y.enstrophy) <= 1e-14 * y.enstrophy
        assert abs(x.max_spectral_divergence - y.max_spectral_divergence) <= 1e-14 * max(y.max_spectral_divergence,
                                                                                         1e-300)
    assert abs(va - vb) <= 1e-14 * vb
    # the unprojected start state has a divergence, the stepped one none
    assert d0a.max_spectral_divergence > 1e-6
    assert da.max_spectral_divergence < 1e-12


def test_blocked_layout_ensemble_members(monkeypatch):
    print("case blocked layout ensemble", flush=True)
    out = {}
    for layout in ("2", "0"):
        monkeypatch.setenv("SPH_LAYOUT", layout)
        ens = sph.Ensemble([256, 256], [1e-2, 2e-2], seeds=[11, 12], precision="fp64")
        for _ in range(2):
            ens.step_rk3([1e-3, 2e-3])
        out[layout] = [[ens.u_hat(m, c) for c in range(3)] for m in range(2)]
        ens.close()
    for m in range(2):
        err = rel(out["2"][m], out["0"][m])
        print(f"  member {m}: blocked vs line rel L2 {err:.2e}", flush=True)
        assert err <= 1e-14


@pytest.mark.parametrize("layout", ["1", "2"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_adaptive_step_in_layout(monkeypatch, layout, precision, tol):
    print("case layout adaptive", layout, precision, flush=True)
    monkeypatch.setenv("SPH_LAYOUT", layout)
    s = sph.TaylorGreenSolver(256, 1600.0, precision)
    assert s.layout()[0] == int(layout)
    for _ in range(2):
        v_ref = s.max_velocity()
        vmax, dt, fused = s.step_adaptive(0.5, 0.5)
        assert fused and abs(vmax - v_ref) <= tol * v_ref
    s.close()


def test_autotuner_reports_every_layout():
    print("case layout report", flush=True)
    s = sph.TaylorGreenSolver(256, 1600.0, "fp32")
    kind, ms = s.layout()
    print("  layout kept:", kind, "ms line 1-3 / transposed 1-3 / blocked 1-4 / best other 1-4:", ms, flush=True)
    assert all(x > 0 for x in ms)
    assert ms[3] >= min(ms[0], ms[1])
    if kind == 2:
        assert ms[2] < ms[3]
    else:
        assert ms[2] >= ms[3]
        assert kind == (1 if ms[1] < ms[0] else 0)
    s.close()

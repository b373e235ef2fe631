"""The adaptive time step's fused max |u| (the stage-0 x pass reduces the
velocity maximum of the state it transforms, passes_vmax.hip) on every x-pass
form, in FP64 and FP32:

  - the GPU driver's CSV against the oracle driver's for a short adaptive-CFL
    run with each autotuner variant forced (SPH_TILE_VARIANT 0, 1 and 3 -- the
    alternative radix order -- run the fused reduction, 2 the LDS-staged x
    pass and max_velocity() instead);
  - sph_solver_step_adaptive: its max |u| equals max_velocity() of the same
    state and its dt is min(dt_max, cfl h / max|u|);
  - an FP32 adaptive run: every step's main_dt column follows from the
    max_velocity column (diagnostics' own transforms) of the state before it.
"""
import math
import os

import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu


def read_csv_rows(path):
    lines = open(path).read().strip().splitlines()
    return lines[0], np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])


@pytest.mark.parametrize("variant", ["0", "1", "2", "3"])
def test_adaptive_driver_matches_oracle_for_every_x_form(tmp_path, monkeypatch, variant):
    print("case adaptive driver, x-pass variant", variant, flush=True)
    monkeypatch.setenv("SPH_TILE_VARIANT", variant)
    probe = sph.TaylorGreenSolver(32, 1600.0, "fp64")
    _, _, fused = probe.step_adaptive(0.5, 0.5)
    probe.close()
    assert fused == (variant != "2")
    kw = dict(adaptive_cfl=True, dt=0.5, cfl=0.5, t_end=1.0, output_dt=0.25)
    gdir, odir = tmp_path / "gpu", tmp_path / "oracle"
    gdir.mkdir()
    odir.mkdir()
    sph.run_taylor_green(sph.RunConfig(n=32, reynolds=1600.0, output_dir=str(gdir), precision="fp64",
                                       print_every=1000, **kw))
    oracle.run_taylor_green(n=32, reynolds=1600.0, output_dir=str(odir), print_every=1000, **kw)
    hg, rg = read_csv_rows(os.path.join(str(gdir), "taylor_green_diagnostics.csv"))
    ho, ro = read_csv_rows(os.path.join(str(odir), "taylor_green_diagnostics.csv"))
    assert hg == ho and rg.shape == ro.shape
    assert np.abs(rg[:, 0] - ro[:, 0]).max() <= 1e-14
    for col in (1, 2, 3, 5, 6, 8):
        err, scale = np.abs(rg[:, col] - ro[:, col]), np.abs(ro[:, col])
        assert np.all(err <= 1e-12 * scale), (hg.split(",")[col], (err / np.maximum(scale, 1e-300)).max())


@pytest.mark.parametrize("precision,n,variant,tol", [("fp64", 32, "0", 1e-12), ("fp64", 64, "1", 1e-12),
                                                     ("fp64", 32, "3", 1e-12), ("fp32", 64, "0", 2e-6),
                                                     ("fp32", 64, "1", 2e-6), ("fp32", 128, "0", 2e-6),
                                                     ("fp32", 128, "3", 2e-6)])
def test_step_adaptive_velocity_max_equals_max_velocity(monkeypatch, precision, n, variant, tol):
    print("case step_adaptive", precision, n, "x-pass variant", variant, flush=True)
    monkeypatch.setenv("SPH_TILE_VARIANT", variant)
    a = sph.TaylorGreenSolver(n, 1600.0, precision)
    b = sph.TaylorGreenSolver(n, 1600.0, precision)
    h = 2 * math.pi / n
    for k in range(3):
        v_ref = a.max_velocity()
        vmax, dt, fused = a.step_adaptive(0.5, 0.5)
        print(f"  step {k}: fused {fused} max|u| {vmax:.15f} max_velocity() {v_ref:.15f} dt {dt:.6e}", flush=True)
        assert fused
        assert abs(vmax - v_ref) <= tol * v_ref
        assert dt == min(0.5, 0.5 * h / vmax)
        b.step_rk3(dt)
    # the same trajectory as plain steps of the same dt (the fused x pass is
    # another compilation of the same arithmetic)
    num = sum(np.linalg.norm(a.u_hat(c) - b.u_hat(c)) ** 2 for c in range(3))
    den = sum(np.linalg.norm(b.u_hat(c)) ** 2 for c in range(3))
    assert math.sqrt(num / den) <= (1e-14 if precision == "fp64" else 1e-6)


def test_fp32_adaptive_run_dt_follows_max_velocity(tmp_path, monkeypatch):
    print("case fp32 adaptive run", flush=True)
    monkeypatch.setenv("SPH_TILE_VARIANT", "0")  # the fused FP32 reduction
    dt_max, cfl, n = 0.5, 0.5, 64
    h = 2 * math.pi / n
    sph.run_taylor_green(sph.RunConfig(n=n, reynolds=1600.0, output_dir=str(tmp_path), precision="fp32",
                                       print_every=1000, adaptive_cfl=True, dt=dt_max, cfl=cfl, t_end=0.4,
                                       output_dt=0.02))
    hdr, rows = read_csv_rows(os.path.join(str(tmp_path), "taylor_green_diagnostics.csv"))
    cols = hdr.split(",")
    vcol, mcol = cols.index("max_velocity"), cols.index("main_dt")
    assert len(rows) >= 15
    # each step ends at an output time here (its dt is clipped to 0.02), so
    # row i holds the state the step of row i+1 started from
    worst = 0.0
    for i in range(len(rows) - 1):
        want = min(dt_max, cfl * h / rows[i, vcol])
        worst = max(worst, abs(rows[i + 1, mcol] - want) / want)
    print("  worst rel main_dt error", worst, flush=True)
    assert worst < 2e-6

"""GPU parity of the fused RK3 step against the CPU oracle (FP64 to 1e-12
relative L2), the reference's published end state, and the FP32 form's error.

The oracle stores the C2C state; its restriction to kx <= n/2 is the GPU's
half-spectrum state (test_gpu_fft.py covers the Re() rule on the kx = 0 and
kx = n/2 planes).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph
from spectral_solver_amd import compare

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DNS = "spectral_Re1600_512.dat"


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def state_error(gpu, orc):
    num = sum(np.linalg.norm(gpu.u_hat(c) - oracle.c2c_to_half(orc.u_hat(c))) ** 2 for c in range(3))
    den = sum(np.linalg.norm(oracle.c2c_to_half(orc.u_hat(c))) ** 2 for c in range(3))
    return np.sqrt(num / den)


@pytest.mark.parametrize("n", [16, 32, 64])
def test_initial_condition_matches_oracle(n):
    print("case ic", n, flush=True)
    g = sph.TaylorGreenSolver(n, 1600.0, "fp64")
    o = oracle.TaylorGreenOracle(n, 1600.0)
    assert state_error(g, o) < 1e-14
    dg, do = g.diagnostics(0.0), o.diagnostics(0.0)
    for k in ("kinetic_energy", "dissipation", "enstrophy", "max_velocity"):
        assert abs(getattr(dg, k) - do[k]) <= 1e-13 * abs(do[k]), k
    assert dg.max_spectral_divergence < 1e-8


def test_golden_trajectory_fixture():
    g0 = np.load(os.path.join(ROOT, "tests", "golden", "tg16_fp64.npz"))
    s = sph.TaylorGreenSolver(16, float(g0["reynolds"]), "fp64")
    for k in range(int(g0["steps"])):
        s.step_rk3(float(g0["dt"]))
        d = s.diagnostics((k + 1) * float(g0["dt"]))
        ref = g0["diagnostics"][k + 1]
        assert abs(d.kinetic_energy - ref[1]) <= 1e-13 * ref[1]
        assert abs(d.dissipation - ref[2]) <= 1e-12 * ref[2]
    for c in range(3):
        assert rel_l2(s.u_hat(c), g0[f"u{c}"]) < 1e-12


def random_state(n, seed):
    """Random solenoidal, truncated C2C state with non-Hermitian values on the
    three kept axis-Nyquist modes (which the reference evolves)."""
    rng = np.random.default_rng(seed)
    o = oracle.TaylorGreenOracle(n, 400.0, initialise=False)
    u = [oracle.fft3d(rng.standard_normal((n, n, n)) * 0.3, -1) for _ in range(3)]
    mx, my, mz, keep, _ = oracle.mode_tables(n)
    k2 = (mx ** 2 + my ** 2 + mz ** 2).astype(float)
    ok = keep & (k2 > 0)
    dot = (mx * u[0] + my * u[1] + mz * u[2]) / np.where(k2 > 0, k2, 1)
    u = [np.where(ok, u[c] - [mx, my, mz][c] * dot, 0) for c in range(3)]
    h = n // 2
    for (iz, iy, ix) in ((0, 0, h), (0, h, 0), (h, 0, 0)):
        for c in range(3):
            if [ix, iy, iz][c] == 0:  # keep it divergence free
                u[c][iz, iy, ix] += (0.2 + 0.9j) * n ** 3 * 0.01
    for c in range(3):
        o.set_u_hat(c, u[c])
    return o, u


@pytest.mark.parametrize("n,steps", [(32, 3), (48, 2), (24, 2), (80, 2), (160, 1), (192, 1)])
def test_random_state_with_nyquist_modes_matches_oracle(n, steps):
    print("case random", n, flush=True)
    o, u = random_state(n, 5 + n)
    g = sph.TaylorGreenSolver(n, 400.0, "fp64")
    for c in range(3):
        g.set_u_hat(c, oracle.c2c_to_half(u[c]))
    for _ in range(steps):
        o.step_rk3(2e-3)
        g.step_rk3(2e-3)
    err = state_error(g, o)
    assert err < 1e-12, err
    d1, d2 = g.diagnostics(0.0), o.diagnostics(0.0)
    assert abs(d1.kinetic_energy - d2["kinetic_energy"]) <= 1e-12 * d2["kinetic_energy"]
    assert abs(d1.enstrophy - d2["enstrophy"]) <= 1e-12 * d2["enstrophy"]


def test_config_64cube_fp64_200_steps_matches_oracle():
    """64^3 TG, nu = 1/1600, dt = 1e-3, 200 fixed steps, FP64: per-step E and
    dissipation, and the fields at the end, against the oracle."""
    n, dt, steps = 64, 1e-3, 200
    g = sph.TaylorGreenSolver(n, 1600.0, "fp64")
    o = oracle.TaylorGreenOracle(n, 1600.0)
    worst_e = worst_eps = 0.0
    for k in range(steps):
        g.step_rk3(dt)
        o.step_rk3(dt)
        if k % 10 == 9 or k == steps - 1:
            dg, do = g.diagnostics((k + 1) * dt), o.diagnostics((k + 1) * dt)
            worst_e = max(worst_e, abs(dg.kinetic_energy - do["kinetic_energy"]) / do["kinetic_energy"])
            worst_eps = max(worst_eps, abs(dg.dissipation - do["dissipation"]) / do["dissipation"])
    print("worst rel E", worst_e, "worst rel eps", worst_eps, flush=True)
    assert worst_e < 1e-12 and worst_eps < 1e-12
    err = state_error(g, o)
    print("field rel L2", err, flush=True)
    assert err < 1e-12


def test_published_64cube_fp64_end_state_at_t20(tmp_path):
    """Adaptive-CFL driver to t = 20 (Benchmark.md:584-591, 352 steps)."""
    cfg = sph.RunConfig(n=64, adaptive_cfl=True, dt=0.5, cfl=0.5, t_end=20.0, output_dt=1.0, reynolds=1600.0,
                        output_dir=str(tmp_path), precision="fp64", print_every=100)
    d = sph.run_taylor_green(cfg)
    assert abs(d.t - 20.0) < 1e-12
    assert abs(d.kinetic_energy - 0.01687796507292055) / 0.01687796507292055 < 1e-11
    assert abs(d.dissipation - 0.0012350569438618611) / 0.0012350569438618611 < 1e-11
    assert abs(d.enstrophy - 0.98804555508948877) / 0.98804555508948877 < 1e-11
    assert abs(d.max_velocity - 0.48555765714990651) < 1e-10
    rows = open(os.path.join(str(tmp_path), "taylor_green_diagnostics.csv")).read().strip().splitlines()
    assert len(rows) == 22  # header + t = 0..20


# Benchmark.md:748-751 (128^3, 740 steps) and :776-779 (256^3, 1527 steps),
# FP64 cuFFT runs; the FLUPS and cuFFTMp runs of 128^3 (:754-765) agree with
# them to 5e-14 (E, eps, enstrophy) and 2e-12 (max |u|).
PUBLISHED_T20 = {
    128: (0.021553419004393556, 0.0019026009263740915, 1.5220807410992732, 0.54486119854790893),
    256: (0.021598156223318035, 0.0019036857700047299, 1.522948616003784, 0.52982227104531054),
}


@pytest.mark.parametrize("n", [128, 256])
def test_published_fp64_end_state_at_t20(tmp_path, n):
    """Adaptive-CFL driver to t = 20 at 128^3 and 256^3 against the published
    finals, and the whole E / eps / enstrophy history against the 512^3
    literature curve (spectral_Re1600_512.dat) with the comparison tool (a
    coarser grid differs from that curve by its resolution error only)."""
    print("case t=20 fp64 n", n, flush=True)
    cfg = sph.RunConfig(n=n, adaptive_cfl=True, dt=0.5, cfl=0.5, t_end=20.0, output_dt=1.0, reynolds=1600.0,
                        output_dir=str(tmp_path), precision="fp64", print_every=250)
    d = sph.run_taylor_green(cfg)
    e, eps, ens, umax = PUBLISHED_T20[n]
    print("E", d.kinetic_energy, "eps", d.dissipation, "enstrophy", d.enstrophy, "max_u", d.max_velocity,
          flush=True)
    assert abs(d.t - 20.0) < 1e-12
    assert abs(d.kinetic_energy - e) / e < 1e-11
    assert abs(d.dissipation - eps) / eps < 1e-11
    assert abs(d.enstrophy - ens) / ens < 1e-11
    assert abs(d.max_velocity - umax) < 1e-10
    csv = os.path.join(str(tmp_path), "taylor_green_diagnostics.csv")
    devs = {x.quantity: x for x in compare.compare(compare.load_ours(csv, 1600.0),
                                                   compare.load_reference(os.path.join(GOLDEN, DNS)))}
    print({k: (v.max_rel, v.at_t) for k, v in devs.items()}, flush=True)
    assert devs["kinetic_energy"].points == 20  # t = 0 .. 19 (the curve ends at 19.99)
    # resolution error of the coarser grid against the 512^3 curve (128^3:
    # 1.8% in E at t = 12, 9.5% in eps at t = 14; 256^3: 0.72% in E at
    # t = 16, 5.5% in eps at t = 18)
    tol_e, tol_eps = {128: (3e-2, 0.15), 256: (1e-2, 8e-2)}[n]
    assert devs["kinetic_energy"].max_rel < tol_e
    assert devs["dissipation"].max_rel < tol_eps and devs["enstrophy"].max_rel < tol_eps


# the reference's own FP32 (cuFFT) runs of the same configurations
# (Benchmark.md:619-622 at 64^3, :633-636 at 128^3): E and eps at t = 20
PUBLISHED_T20_FP32 = {64: (0.016877809604920913, 0.0012350909170767454),
                      128: (0.021553486211075779, 0.001902701557113874)}
PUBLISHED_T20_FP64 = {64: (0.01687796507292055, 0.0012350569438618611),
                      128: PUBLISHED_T20[128][:2]}


@pytest.mark.parametrize("n", [64, 128])
def test_fp32_form_at_t20_against_published_finals(tmp_path, n):
    """The FP32 form through the adaptive-CFL driver to t = 20: its E and eps
    lie within 1e-4 of the published FP64 finals (the bound on the FP32
    form's dissipation error; the reference's own FP32 runs are 3e-6 to 5e-5
    away at these sizes) and within 1e-4 of the reference's FP32 finals."""
    print("case t=20 fp32 n", n, flush=True)
    cfg = sph.RunConfig(n=n, adaptive_cfl=True, dt=0.5, cfl=0.5, t_end=20.0, output_dt=1.0, reynolds=1600.0,
                        output_dir=str(tmp_path), precision="fp32", print_every=250)
    d = sph.run_taylor_green(cfg)
    assert abs(d.t - 20.0) < 1e-12
    for name, ref in (("FP64", PUBLISHED_T20_FP64[n]), ("reference FP32", PUBLISHED_T20_FP32[n])):
        e_err = abs(d.kinetic_energy - ref[0]) / ref[0]
        eps_err = abs(d.dissipation - ref[1]) / ref[1]
        print(f"  vs published {name}: E rel {e_err:.2e}, eps rel {eps_err:.2e}", flush=True)
        assert e_err < 1e-4 and eps_err < 1e-4


def test_fp32_form_error_against_fp64():
    n, dt, steps = 64, 1e-3, 100
    g32 = sph.TaylorGreenSolver(n, 1600.0, "fp32")
    g64 = sph.TaylorGreenSolver(n, 1600.0, "fp64")
    worst = 0.0
    for k in range(steps):
        g32.step_rk3(dt)
        g64.step_rk3(dt)
        if k % 20 == 19:
            a, b = g32.diagnostics(0.0), g64.diagnostics(0.0)
            worst = max(worst, abs(a.dissipation - b.dissipation) / b.dissipation)
    print("fp32 worst rel eps", worst, flush=True)
    assert worst < 1e-5


def test_graph_and_direct_launch_agree_bitwise():
    a = sph.TaylorGreenSolver(32, 1600.0, "fp64")
    b = sph.TaylorGreenSolver(32, 1600.0, "fp64")
    b.set_graph(False)
    for _ in range(3):
        a.step_rk3(1e-3)
        b.step_rk3(1e-3)
    for c in range(3):
        assert np.array_equal(a.u_hat(c), b.u_hat(c))


def test_step_times_are_per_step_gpu_times():
    """sph_solver_step_times: one positive GPU time per step, oldest first,
    at most the number asked for, summing to the profile's step time."""
    s = sph.TaylorGreenSolver(64, 1600.0, "fp32")
    for _ in range(5):
        s.step_rk3(1e-3)
    t = s.step_times()
    assert len(t) == 5 and np.all(t > 0)
    assert np.array_equal(s.step_times(3), t[2:])
    assert len(s.step_times(0)) == 0
    assert abs(t.sum() * 1e-3 - s.profile()["step"]) <= 1e-9 + 1e-6 * s.profile()["step"]
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_calls_between_steps_leave_the_trajectory_unchanged(precision):
    """Each stage's last kernel also computes the next stage's inverse z
    transforms into the work buffer. Diagnostics, max_velocity, get_uhat,
    set_uhat and time_passes between steps must not change the trajectory:
    a solver interleaving them ends bitwise equal to one that only steps."""
    n, dt = 32, 2e-3
    a = sph.TaylorGreenSolver(n, 1600.0, precision)
    b = sph.TaylorGreenSolver(n, 1600.0, precision)
    for k in range(6):
        a.step_rk3(dt)
        b.step_rk3(dt)
        if k == 0:
            b.diagnostics(dt)
        elif k == 1:
            b.max_velocity()
        elif k == 2:
            b.time_passes(1)
        elif k == 3:
            for c in range(3):
                b.set_u_hat(c, b.u_hat(c))  # same state, set from outside
        elif k == 4:
            b.set_graph(False)
    for c in range(3):
        assert np.array_equal(a.u_hat(c), b.u_hat(c))
    # a state set from outside after some steps is stepped from its own
    # transforms, not from those the last step left behind
    o, u = random_state(n, 5)  # oracle at Re = 400
    g = sph.TaylorGreenSolver(n, 400.0, precision)
    g.step_rk3(dt)
    for c in range(3):
        g.set_u_hat(c, oracle.c2c_to_half(u[c]))
    for _ in range(2):
        g.step_rk3(dt)
        o.step_rk3(dt)
    tol = 1e-12 if precision == "fp64" else 2e-5
    assert state_error(g, o) < tol


def test_time_passes_leaves_state_unchanged():
    s = sph.TaylorGreenSolver(32, 1600.0, "fp32")
    s.step_rk3(1e-3)
    before = [s.u_hat(c) for c in range(3)]
    t = s.time_passes(2)
    assert all(v[0] > 0 for v in t.values())
    for c in range(3):
        assert np.array_equal(before[c], s.u_hat(c))


def test_reference_cpp_tests_binary(tmp_path):
    exe = os.path.join(ROOT, "spectral_solver_amd", "spectral_ns_tests")
    r = subprocess.run([exe, str(tmp_path / "vorticity.ppm")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "all C++ spectral tests passed" in r.stdout


def test_cli_fixed_dt(tmp_path):
    exe = os.path.join(ROOT, "spectral_solver_amd", "taylor_green")
    r = subprocess.run([exe, "--n", "32", "--steps", "4", "--dt", "1e-3", "--output-dir", str(tmp_path),
                        "--precision", "fp32"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    rows = open(os.path.join(str(tmp_path), "taylor_green_diagnostics.csv")).read().strip().splitlines()
    assert len(rows) == 6 and rows[0].startswith("t,kinetic_energy,dissipation")
    assert "profile_solver" in r.stdout
    r = subprocess.run([exe, "--n", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "error: grid size must be positive" in r.stderr


@pytest.mark.parametrize("precision,n", [("fp64", 32), ("fp32", 64), ("fp64", 128)])
def test_tile_variants_agree_and_match_oracle(monkeypatch, precision, n):
    """Every autotuner variant of the y and x passes (tile widths, the
    register-direct and LDS-staged x pass, the alternative radix order of the
    register-direct x pass and the forward-z pass) gives the oracle's state (FP64:
    1e-12; FP32: the FP32 form's rounding) and the others' to rounding (the
    compiler contracts multiply-adds differently in different kernels, so the
    bits may differ). In FP32 at 64^3 the 8- and 4-line y tiles use the
    unpadded rotated LDS rows and the 8-line x tile the unpadded z-tile rows
    with the permuted first-stage write; variants 1 and 2 also run the
    forward-z pass (and the standalone inverse z) on half-width z tiles."""
    steps = 3
    o, u = random_state(n, 11)
    states = []
    alt = sph.radix_plan(n, 1) != sph.radix_plan(n, 0)  # 32 and 128: yes, 64: no
    for v in ("0", "1", "2", "3"):
        print("case tile variant", precision, n, v, flush=True)
        monkeypatch.setenv("SPH_TILE_VARIANT", v)
        g = sph.TaylorGreenSolver(n, 400.0, precision)
        k = int(v)
        tv = g.tuning()
        # y passes: three widths; x pass: variant 3 (the alternative radix
        # order) where the length has one, else the last variant (LDS-staged)
        assert [tv[p][0] for p in ("inverse_y", "x_products", "forward_y")] == \
            [min(k, 2), k if (k < 3 or alt) else 2, min(k, 2)]
        # the forward-z pass: the default width, half of it where the half
        # tile still fills whole wavefronts (64^3 FP32: 16 and 8 lines;
        # 128^3 FP64: 8 and 4; otherwise variant 1 is the default width
        # again), and the alternative radix order where there is one
        nz = 3 if alt else (2 if (precision, n) in (("fp32", 64), ("fp64", 128)) else 1)
        assert tv["forward_z_update"][0] == min(k, nz - 1)
        plans = g.radix_plans()
        want_x = sph.radix_plan(n, 1) if (k == 3 and alt) else (None if k == 2 or k == 3 else sph.radix_plan(n, 0))
        if want_x is not None:
            assert plans["x_products"] == want_x
        assert plans["forward_z_update"] == sph.radix_plan(n, 1 if (alt and k >= 2) else 0)
        for c in range(3):
            g.set_u_hat(c, oracle.c2c_to_half(u[c]))
        for _ in range(steps):
            g.step_rk3(2e-3)
        states.append([g.u_hat(c) for c in range(3)])
        g.close()
    for _ in range(steps):
        o.step_rk3(2e-3)
    agree, match = (1e-14, 1e-12) if precision == "fp64" else (1e-6, 1e-6)  # FP32 measured 8e-8
    for st in states[1:]:
        num = sum(np.linalg.norm(st[c] - states[0][c]) ** 2 for c in range(3))
        den = sum(np.linalg.norm(states[0][c]) ** 2 for c in range(3))
        assert np.sqrt(num / den) < agree
    for st in states:
        num = sum(np.linalg.norm(st[c] - oracle.c2c_to_half(o.u_hat(c))) ** 2 for c in range(3))
        den = sum(np.linalg.norm(oracle.c2c_to_half(o.u_hat(c))) ** 2 for c in range(3))
        print("  rel L2 vs oracle", np.sqrt(num / den), flush=True)
        assert np.sqrt(num / den) < match


def read_csv_rows(path):
    lines = open(path).read().strip().splitlines()
    return lines[0], np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])


@pytest.mark.parametrize("adaptive", [True, False])
def test_driver_csv_matches_oracle_driver(tmp_path, adaptive):
    """The GPU driver against the oracle's driver (impl.hpp:1310-1410) on the
    same short run, row by row: same header, same output times (the adaptive
    branch clips dt to t_end and to the next output time), diagnostics to
    1e-12."""
    print("case driver csv adaptive" if adaptive else "case driver csv fixed dt", flush=True)
    kw = (dict(adaptive_cfl=True, dt=0.5, cfl=0.5, t_end=1.0, output_dt=0.25) if adaptive
          else dict(adaptive_cfl=False, steps=20, dt=0.01, output_every=5))
    gdir, odir = tmp_path / "gpu", tmp_path / "oracle"
    gdir.mkdir()
    odir.mkdir()
    cfg = sph.RunConfig(n=32, reynolds=1600.0, output_dir=str(gdir), precision="fp64", print_every=1000, **kw)
    if not adaptive:
        cfg.t_end = kw["steps"] * kw["dt"]
    sph.run_taylor_green(cfg)
    _, _, steps = oracle.run_taylor_green(n=32, reynolds=1600.0, output_dir=str(odir), print_every=1000, **kw)
    hg, rg = read_csv_rows(os.path.join(str(gdir), "taylor_green_diagnostics.csv"))
    ho, ro = read_csv_rows(os.path.join(str(odir), "taylor_green_diagnostics.csv"))
    print("rows", len(rg), "oracle steps", steps, flush=True)
    assert hg == ho
    assert rg.shape == ro.shape and len(rg) >= 5
    assert np.abs(rg[:, 0] - ro[:, 0]).max() <= 1e-14  # output times
    # E, dissipation, enstrophy, max |u|, dt (and with them the CFL columns)
    # (the t = 0 row has dt = 0 in both)
    for col in (1, 2, 3, 5, 6):
        err, scale = np.abs(rg[:, col] - ro[:, col]), np.abs(ro[:, col])
        assert np.all(err <= 1e-12 * scale), (hg.split(",")[col], (err / np.maximum(scale, 1e-300)).max())

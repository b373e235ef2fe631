"""Pins the CPU oracle (oracle/oracle.hpp) before it is trusted as the checker.

Each oracle function is checked against something that does not retype its own
formula: brute-force DFT sums, numpy's FFT, closed forms, the reference's own
known-answer tests (cpp_tests/spectral_ns_tests.cpp:16-51), the literature
Taylor-Green data shipped with it (spectral_Re1600_512.dat row 1) and the
published 64^3 FP64 end state at t=20 (Benchmark.md:172-177, 584-591).
"""
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def brute_dft3(a, sign):
    """O(N^2) DFT along each axis with explicit DFT matrices (no FFT)."""
    out = a.astype(np.complex128)
    for axis, n in enumerate(a.shape):
        j = np.arange(n)
        W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, axis, 0), axes=(1, 0)), 0, axis)
    return out


@pytest.mark.parametrize("shape", [(4, 6, 8), (12, 12, 12), (3, 5, 7), (9, 10, 6)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_brute_force_dft(shape, sign):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = brute_dft3(a, sign)
    got = oracle.fft3d(a, sign)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_fft_matches_numpy_golden_vectors():
    data = np.load(os.path.join(GOLDEN, "fft_golden.npz"))
    for key in data.files:
        if not key.startswith("in_"):
            continue
        tag = key[3:]
        a = data[key]
        fwd = oracle.fft3d(a, -1)
        assert np.abs(fwd - data["fwd_" + tag]).max() <= 1e-12 * np.abs(fwd).max(), tag
        bwd = oracle.fft3d(a, 1)
        assert np.abs(bwd - data["bwd_" + tag]).max() <= 1e-12 * np.abs(bwd).max(), tag


def test_fft_single_mode_and_parseval():
    n = 16
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a = np.exp(2j * np.pi * (3 * x - 5 * y + 8 * z) / n)
    f = oracle.fft3d(a, -1)
    expect = np.zeros_like(f)
    expect[8 % n, (-5) % n, 3] = n ** 3
    assert np.abs(f - expect).max() < 1e-9
    rng = np.random.default_rng(3)
    b = rng.standard_normal((n, n, n))
    fb = oracle.fft3d(b, -1)
    assert np.isclose((np.abs(fb) ** 2).sum() / n ** 3, (b ** 2).sum(), rtol=1e-13)


def test_mode_tables_signed_modes_mask_and_phase():
    n = 8
    mx, my, mz, keep, ph = oracle.mode_tables(n)
    idx = np.arange(n)
    signed = np.where(idx <= n // 2, idx, idx - n)
    assert (mx[0, 0, :] == signed).all() and (my[0, :, 0] == signed).all() and (mz[:, 0, 0] == signed).all()
    assert mx[0, 0, n // 2] == n // 2  # Nyquist is +n/2
    # brute-force lattice count of the radius-n/2 ball
    count = sum(1 for a in signed for b in signed for c in signed if a * a + b * b + c * c <= (n // 2) ** 2)
    assert keep.sum() == count
    assert keep[0, 0, n // 2] and keep[0, n // 2, 0] and keep[n // 2, 0, 0]
    assert not keep[0, 1, n // 2] and not keep[1, 0, n // 2]
    assert np.abs(ph - np.exp(1j * np.pi * (mx + my + mz) / n)).max() < 1e-15


def _spec(n, comps):
    """C2C spectrum of a real field given as {(kx,ky,kz): amplitude} (x fastest)."""
    a = np.zeros((n, n, n), np.complex128)
    for (kx, ky, kz), v in comps.items():
        a[kz % n, ky % n, kx % n] += v
    return a


def test_dealiased_product_removes_single_aliases():
    # a = b = cos(p x): a*b = 1/2 + cos(2p x)/2. With p = 5 on n = 16 the
    # harmonic 2p = 10 lies outside the grid and aliases onto |k| = 6; the
    # two-grid phase shift cancels it, leaving only the mean.
    n, p = 16, 5
    N3 = n ** 3
    a = _spec(n, {(p, 0, 0): N3 / 2, (-p, 0, 0): N3 / 2})
    out = oracle.dealiased_product(a, a)
    expect = _spec(n, {(0, 0, 0): N3 / 2})
    assert np.abs(out - expect).max() < 1e-9 * N3
    # resolved product: sin(2x) * cos(3y) keeps all four modes with amplitude N3/4
    s = _spec(n, {(2, 0, 0): -0.5j * N3, (-2, 0, 0): 0.5j * N3})
    c = _spec(n, {(0, 3, 0): N3 / 2, (0, -3, 0): N3 / 2})
    out = oracle.dealiased_product(s, c)
    expect = _spec(n, {(2, 3, 0): -0.25j * N3, (2, -3, 0): -0.25j * N3,
                       (-2, 3, 0): 0.25j * N3, (-2, -3, 0): 0.25j * N3})
    assert np.abs(out - expect).max() < 1e-9 * N3


def test_taylor_green_initial_diagnostics_closed_forms():
    for n in (16, 32):
        s = oracle.TaylorGreenOracle(n, 1600.0)
        d = s.diagnostics(0.0)
        assert abs(d["kinetic_energy"] - 0.125) < 1e-12
        assert abs(d["enstrophy"] - 0.375) < 1e-12
        assert abs(d["dissipation"] - 2.0 / 1600.0 * 0.375) < 1e-15
        x = (np.arange(n) + 0.5) * 2 * np.pi / n
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        vmax = np.sqrt((np.sin(X) * np.cos(Y) * np.cos(Z)) ** 2 + (np.cos(X) * np.sin(Y) * np.cos(Z)) ** 2).max()
        assert abs(d["max_velocity"] - vmax) < 1e-13
        assert d["max_spectral_divergence"] < 1e-8
    # literature DNS row 1 (t=0): E, eps, enstrophy
    row = np.loadtxt(os.path.join(GOLDEN, "spectral_Re1600_512_head.dat"))[0]
    assert row[0] == 0.0 and abs(row[1] - 0.125) < 1e-12 and abs(row[3] - 0.375) < 1e-12
    assert abs(d["dissipation"] - row[2]) / row[2] < 2e-5


def test_reference_cpp_known_answers_on_oracle():
    # cpp_tests/spectral_ns_tests.cpp:16-51 thresholds, 12^3 grid, nu = 0.05
    n = 12
    x = (np.arange(n) + 0.5) * 2 * np.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = np.sin(X) + 0.25 * np.cos(Y)
    back = oracle.fft3d(oracle.fft3d(u, -1), 1).real / n ** 3
    assert np.abs(back - u).max() < 1e-10
    s = oracle.TaylorGreenOracle(n, 1.0 / 0.05)
    d0 = s.diagnostics(0.0)
    assert d0["max_spectral_divergence"] < 1e-8
    assert abs(d0["kinetic_energy"] - 0.125) < 1e-12
    s.step_rk3(1e-3)
    d1 = s.diagnostics(1e-3)
    assert d1["kinetic_energy"] < d0["kinetic_energy"]
    assert d1["max_spectral_divergence"] < 1e-7


@pytest.mark.parametrize("k,re,dt", [(3, 10.0, 0.1), (5, 40.0, 0.05), (1, 1600.0, 1e-3)])
def test_rk3_amplification_of_pure_shear_mode(k, re, dt):
    # u = (f(y), 0, 0): the nonlinear term vanishes identically, so one
    # Williamson step multiplies the mode by the 3-stage 3rd-order stability
    # polynomial 1 + z + z^2/2 + z^3/6 with z = -nu k^2 dt.
    n = 16
    N3 = n ** 3
    s = oracle.TaylorGreenOracle(n, re, initialise=False)
    amp = (0.3 - 0.7j) * N3
    ux = _spec(n, {(0, k, 0): amp, (0, -k, 0): np.conj(amp)})
    zero = np.zeros_like(ux)
    s.set_u_hat(0, ux)
    s.set_u_hat(1, zero)
    s.set_u_hat(2, zero)
    s.step_rk3(dt)
    z = -(1.0 / re) * k * k * dt
    g = 1 + z + z * z / 2 + z ** 3 / 6
    got = s.u_hat(0)
    assert abs(got[0, k, 0] / amp - g) < 1e-13
    assert abs(got[0, n - k, 0] / np.conj(amp) - g) < 1e-13
    mask = np.ones_like(got, bool)
    mask[0, k, 0] = mask[0, n - k, 0] = False
    assert np.abs(got[mask]).max() < 1e-10 * N3
    assert np.abs(s.u_hat(1)).max() < 1e-10 * N3 and np.abs(s.u_hat(2)).max() < 1e-10 * N3


def test_golden_trajectory_fixture_regression():
    """The oracle still produces the committed 16^3 trajectory fixture."""
    g = np.load(os.path.join(GOLDEN, "tg16_fp64.npz"))
    s = oracle.TaylorGreenOracle(16, 1600.0)
    for step in range(int(g["steps"])):
        s.step_rk3(float(g["dt"]))
    for c in range(3):
        got = oracle.c2c_to_half(s.u_hat(c))
        assert np.linalg.norm(got - g[f"u{c}"]) <= 1e-13 * np.linalg.norm(g[f"u{c}"])


@pytest.mark.slow
def test_published_64cube_fp64_end_state_at_t20(tmp_path):
    """Benchmark.md:172-177 (counts, 352 steps) and :584-591 (FFTW FP64 finals)."""
    row, cnt, steps = oracle.run_taylor_green(n=64, adaptive_cfl=True, dt=0.5, cfl=0.5, t_end=20.0,
                                              output_dt=1.0, reynolds=1600.0, output_dir=str(tmp_path))
    assert steps == 352
    assert cnt["forward_calls"] == 12675 and cnt["backward_calls"] == 7518
    assert cnt["rhs_calls"] == 1056 and cnt["diagnostics_calls"] == 21 and cnt["max_velocity_calls"] == 352
    assert abs(row["t"] - 20.0) < 1e-12
    assert abs(row["kinetic_energy"] - 0.01687796507292055) < 1e-14
    assert abs(row["dissipation"] - 0.0012350569438618611) / 0.0012350569438618611 < 1e-12
    assert abs(row["enstrophy"] - 0.98804555508948877) / 0.98804555508948877 < 1e-12
    assert abs(row["max_velocity"] - 0.48555765714990651) < 1e-11
    header = open(os.path.join(str(tmp_path), "taylor_green_diagnostics.csv")).readline().strip()
    assert header == ("t,kinetic_energy,dissipation,enstrophy,max_spectral_divergence,max_velocity,"
                      "dt,cfl,main_dt,main_cfl")

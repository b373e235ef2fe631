"""Parity at the full sizes of the benchmark configurations: 512^3 FP64 and
FP32 (dt = 5e-4) and 1024^3 FP32 (dt = 2.5e-4), and 1024^3 in FP64 (the plain
FP64 path at the largest size that fits), through properties that do not
depend on the grid. From the Taylor-Green field, every RK stage at most
doubles the largest wavenumber present, and a mode reached through s doublings
has an amplitude of order (|k| dt)^s, so after two steps (six stages) the modes
a 64^3 grid truncates hold less than ~1e-14 of the state: the solution on the
large grid equals the 64^3 one to rounding. The CPU oracle at 64^3 FP64 is the
checker: kinetic energy, dissipation and enstrophy (integrals over the box),
and the low modes of the state normalised by N^3 (forward transforms are
unnormalised). FP64 agrees to 1e-12; the FP32 form to 2e-6 (its FP32
transforms round at ~6e-8 per operation; diagnostics reduce in FP64).
The premise is checked with the oracle alone: its 32^3, 48^3 and 96^3 results
after the same steps match its 64^3 result to 6e-15 in the integrals and 4e-16
in the low modes, while one step fewer moves the integrals by 2e-6 and the
modes by 2e-4, and Re = 1000 moves the dissipation by 60%.
"""
import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu

RE = 1600.0
STEPS = 2
COARSE = 64
LOW = 8  # |mx|, |my|, |mz| <= LOW are compared


def low_modes(u_half, n):
    """Fourier coefficients of the field for 0 <= mx <= LOW and |my|, |mz| <=
    LOW, ordered (mx, mz, my), from a half spectrum [kx][kz][ky] of size n:
    the DFT over the cell-centred points x_i = (i + 1/2) 2 pi / n divided by
    n^3 and by the half-cell phase exp(i pi (mx + my + mz) / n), which is the
    only place n enters."""
    signed = np.r_[0:LOW + 1, -LOW:0]
    idx = signed % n
    mx = np.arange(LOW + 1)[:, None, None]
    mz, my = signed[None, :, None], signed[None, None, :]
    shift = np.exp(-1j * np.pi * (mx + my + mz) / n)
    return u_half[: LOW + 1][:, idx][:, :, idx] * shift / float(n) ** 3


_coarse = {}


def coarse_reference(dt):
    """Oracle 64^3 FP64 after STEPS steps: diagnostics and low modes."""
    if dt not in _coarse:
        o = oracle.TaylorGreenOracle(COARSE, RE)
        for _ in range(STEPS):
            o.step_rk3(dt)
        d = o.diagnostics(STEPS * dt)
        modes = [low_modes(oracle.c2c_to_half(o.u_hat(c)), COARSE) for c in range(3)]
        o.close()
        _coarse[dt] = (d, modes)
    return _coarse[dt]


@pytest.mark.parametrize("n,precision,dt,tol,components", [
    (512, "fp64", 5e-4, 1e-12, 3),
    (512, "fp32", 5e-4, 2e-6, 3),
    (1024, "fp32", 2.5e-4, 2e-6, 1),
    # FP64 at 1024^3: 206 GB of fields in the line layout (the transposed
    # layout's second work buffer does not fit beside them)
    (1024, "fp64", 2.5e-4, 1e-12, 1),
])
def test_full_size_step_matches_coarse_oracle(n, precision, dt, tol, components):
    print(f"case full size n={n} {precision} dt={dt} steps={STEPS}", flush=True)
    d_ref, modes_ref = coarse_reference(dt)
    s = sph.TaylorGreenSolver(n, RE, precision)
    try:
        for _ in range(STEPS):
            s.step_rk3(dt)
        d = s.diagnostics(STEPS * dt)
        for k in ("kinetic_energy", "dissipation", "enstrophy"):
            err = abs(getattr(d, k) - d_ref[k]) / abs(d_ref[k])
            print(f"  {k}: {getattr(d, k):.15e} vs {d_ref[k]:.15e} rel {err:.2e}", flush=True)
            assert err <= tol, k
        # the projection leaves the state divergence free (unnormalised units)
        assert d.max_spectral_divergence <= (1e-12 if precision == "fp64" else 1e-5) * float(n) ** 3
        num = den = 0.0
        for c in range(components):
            got = low_modes(s.u_hat(c), n)
            num += np.linalg.norm(got - modes_ref[c]) ** 2
            den += np.linalg.norm(modes_ref[c]) ** 2
        err = np.sqrt(num / den)
        print(f"  low modes (|m| <= {LOW}, {components} components): rel L2 {err:.2e}", flush=True)
        assert err <= tol
    finally:
        s.close()


def test_full_size_1024_alternative_radix_order(monkeypatch):
    """1024^3 FP32 with the z-type passes in the alternative radix order
    (8*16*8: a radix-16 middle stage, 16 values per thread), forced through
    SPH_TILE_VARIANT=3, against the same 64^3 oracle as above."""
    n, dt, tol = 1024, 2.5e-4, 2e-6
    monkeypatch.setenv("SPH_TILE_VARIANT", "3")
    print("case full size n=1024 fp32, alternative radix order", flush=True)
    d_ref, modes_ref = coarse_reference(dt)
    s = sph.TaylorGreenSolver(n, RE, "fp32")
    try:
        plans = s.radix_plans()
        assert plans["forward_z_update"] == [8, 16, 8] and plans["x_products"] == [8, 16, 8]
        for _ in range(STEPS):
            s.step_rk3(dt)
        d = s.diagnostics(STEPS * dt)
        for k in ("kinetic_energy", "dissipation", "enstrophy"):
            err = abs(getattr(d, k) - d_ref[k]) / abs(d_ref[k])
            print(f"  {k}: rel {err:.2e}", flush=True)
            assert err <= tol, k
        got = low_modes(s.u_hat(0), n)
        err = np.linalg.norm(got - modes_ref[0]) / np.linalg.norm(modes_ref[0])
        print(f"  low modes: rel L2 {err:.2e}", flush=True)
        assert err <= tol
    finally:
        s.close()

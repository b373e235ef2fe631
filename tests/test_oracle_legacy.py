"""Pins of the CPU restatement of the reference's legacy solver
(oracle/legacy.hpp <- src/spectral_solver.cpp:35-409), each chosen so that a
plausible mistake fails it:

  - on a cubic 2*pi box with the Taylor-Green field it equals the AVA-solver
    restatement (oracle.hpp, pinned separately by the published t = 20 end
    states) to rounding: the same scheme through other code (per-axis
    wavenumbers, division instead of reciprocal, RK coefficients written
    (15.0/16.0)*dt, zero_mean after projection);
  - a shear mode on a (2 pi, 4 pi, pi) box: the nonlinear term vanishes
    exactly, so one RK3 step multiplies u_hat by 1 + z + z^2/2 + z^3/6,
    z = -nu k^2 dt with k = 2 pi m / L of the mode's axis (a wrong box-length
    scaling, axis or RK coefficient fails it);
  - the single-alias closed form: for u = (sin 5x cos 5y, -cos 5x sin 5y, 0)
    on a 16x16x8 grid the products carry modes 10, which alias to 6 (at cell
    centres cos 10x = -cos 6x, sin 10x = sin 6x); without dealiasing the
    nonlinear term is exactly (1.5 sin 6x, 1.5 sin 6y, 0) in physical space,
    with phase-shift dealiasing it is exactly 0 (the single aliases cancel,
    the double ones fall outside the ellipsoidal mask).
"""
import math

import numpy as np
import pytest

import oracle

TWO_PI = 2.0 * math.pi


def cell_grid(shape, lengths):
    nx, ny, nz = shape
    x = (np.arange(nx) + 0.5) * lengths[0] / nx
    y = (np.arange(ny) + 0.5) * lengths[1] / ny
    z = (np.arange(nz) + 0.5) * lengths[2] / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_cubic_taylor_green_equals_ava_restatement():
    n, re, dt, steps = 16, 1600.0, 2e-2, 3
    diag, _, state = oracle.legacy_run((n, n, n), (TWO_PI,) * 3, 1.0 / re, True, steps, dt)
    ava = oracle.TaylorGreenOracle(n, re)
    d0 = ava.diagnostics(0.0)
    assert abs(diag[0, 1] - d0["kinetic_energy"]) <= 1e-14 * d0["kinetic_energy"]
    for _ in range(steps):
        ava.step_rk3(dt)
    ref = [ava.u_hat(c) for c in range(3)]
    err = rel(np.stack(state), np.stack(ref))
    d = ava.diagnostics(steps * dt)
    print("legacy vs AVA restatement after", steps, "steps: rel L2", err)
    assert err < 1e-13
    assert abs(diag[-1, 1] - d["kinetic_energy"]) <= 1e-13 * d["kinetic_energy"]
    assert abs(diag[-1, 2] - d["dissipation"]) <= 1e-12 * d["dissipation"]
    assert abs(diag[-1, 3] - d["enstrophy"]) <= 1e-12 * d["enstrophy"]


@pytest.mark.parametrize("axis_u,axis_k,m", [(0, 1, 1), (1, 2, 1), (2, 0, 2)])
def test_shear_mode_follows_rk3_polynomial_on_a_box(axis_u, axis_k, m):
    shape, lengths = (16, 12, 8), (TWO_PI, 2 * TWO_PI, 0.5 * TWO_PI)
    coords = cell_grid(shape, lengths)
    k = TWO_PI * m / lengths[axis_k]
    nu, dt = 1.0, 0.3 / (k * k)  # z = -0.3
    phys = [np.zeros(coords[0].shape) for _ in range(3)]
    phys[axis_u] = np.sin(k * coords[axis_k])
    _, _, s0 = oracle.legacy_run(shape, lengths, nu, True, 0, dt, phys)
    _, _, s1 = oracle.legacy_run(shape, lengths, nu, True, 1, dt, phys)
    z = -nu * k * k * dt
    amp = 1 + z + z * z / 2 + z ** 3 / 6
    print("z", z, "amplification", amp)
    assert abs(z + 0.3) < 1e-12
    assert rel(np.stack(s1), amp * np.stack(s0)) < 1e-13
    # the state is the mode +-m along axis_k of component axis_u (the other
    # components hold rounding only)
    big = np.linalg.norm(s0[axis_u])
    assert big > 0 and all(np.linalg.norm(s0[c]) < 1e-14 * big for c in range(3) if c != axis_u)


@pytest.mark.parametrize("dealias", [False, True])
def test_single_alias_closed_form(dealias):
    shape, lengths = (16, 16, 8), (TWO_PI,) * 3
    X, Y, _ = cell_grid(shape, lengths)
    phys = [np.sin(5 * X) * np.cos(5 * Y), -np.cos(5 * X) * np.sin(5 * Y), np.zeros(X.shape)]
    _, rhs0, _ = oracle.legacy_run(shape, lengths, 0.0, dealias, 0, 0.0, phys)
    n_pts = X.size
    # physical values of the spectral nonlinear term (numpy's inverse FFT, Re)
    got = [np.fft.ifftn(r).real for r in rhs0]
    assert max(float(np.max(np.abs(np.fft.ifftn(r).imag))) for r in rhs0) < 1e-12
    if dealias:
        want = [np.zeros(X.shape)] * 3
        scale = 1.0
    else:
        want = [1.5 * np.sin(6 * X), 1.5 * np.sin(6 * Y), np.zeros(X.shape)]
        scale = 1.5
    err = max(float(np.max(np.abs(g - w))) for g, w in zip(got, want))
    print("dealias", dealias, "max error", err, "points", n_pts)
    assert err < 1e-13 * scale

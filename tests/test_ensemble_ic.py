"""The batched ensemble's seeded random-phase solenoidal initial field (host
side, no GPU): the library's generator against the oracle's numpy restatement,
and the properties that fix it independently of either (divergence free,
Hermitian, dealiasing mask, energy, the E(k) shape, real physical field).

The reference has no ensemble (one TaylorGreenSolver per process,
src/ava_taylor_green_impl.hpp:658-709); the field and its normalisation are
this library's definition, so these properties are what pins it.
"""
import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph

CASES = [(8, 1), (12, 1000), (16, 7), (24, 3), (32, 1001), (48, 1015)]


def modes(n):
    H = n // 2 + 1
    sg = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    mx = np.arange(H)[:, None, None] * np.ones((1, n, n), int)
    mz = sg[None, :, None] * np.ones((H, 1, n), int)
    my = sg[None, None, :] * np.ones((H, n, 1), int)
    return mx, my, mz


@pytest.mark.parametrize("n,seed", CASES)
def test_library_field_matches_numpy_restatement(n, seed):
    a = sph.random_solenoidal_field(n, seed, 0.5)
    b = oracle.random_solenoidal_field(n, seed, 0.5)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() <= 1e-13 * np.abs(y).max()


@pytest.mark.parametrize("n,seed", CASES)
def test_field_is_solenoidal_masked_and_hermitian(n, seed):
    u = sph.random_solenoidal_field(n, seed, 0.5)
    mx, my, mz = modes(n)
    scale = max(np.abs(c).max() for c in u)
    div = mx * u[0] + my * u[1] + mz * u[2]
    assert np.abs(div).max() <= 1e-13 * scale * n
    k2 = mx ** 2 + my ** 2 + mz ** 2
    outside = (4 * k2 > n * n) | (mx == n // 2) | (my == n // 2) | (mz == n // 2) | (k2 == 0)
    for c in u:
        assert np.all(c[outside] == 0)
    # kx = 0 plane: u(0, -my, -mz) = conj u(0, my, mz)
    for c in u:
        p = c[0]
        partner = np.roll(np.flip(p, axis=(0, 1)), shift=(1, 1), axis=(0, 1))
        assert np.array_equal(p, np.conj(partner))


@pytest.mark.parametrize("n,seed", CASES)
def test_energy_and_real_physical_field(n, seed):
    energy = 0.37
    u = sph.random_solenoidal_field(n, seed, energy)
    full = [oracle.half_to_c2c(c) for c in u]
    e = 0.5 * sum(np.sum(np.abs(f) ** 2) for f in full) / n ** 6
    assert abs(e - energy) <= 1e-13 * energy
    phys = [np.fft.ifftn(f) for f in full]
    peak = max(np.abs(p.real).max() for p in phys)
    assert max(np.abs(p.imag).max() for p in phys) <= 1e-13 * peak
    # Parseval in physical space: <|u|^2>/2
    e_phys = 0.5 * sum(np.mean(p.real ** 2) for p in phys)
    assert abs(e_phys - energy) <= 1e-12 * energy


@pytest.mark.parametrize("n,seed", [(16, 2), (32, 1003)])
def test_mode_amplitudes_follow_the_spectrum(n, seed):
    """|u_hat(k)|^2 4 pi k^2 / E(k) is one constant over the generated modes,
    with E(k) = k^4 exp(-2 (k/4)^2)."""
    u = sph.random_solenoidal_field(n, seed, 0.5)
    mx, my, mz = modes(n)
    k2 = (mx ** 2 + my ** 2 + mz ** 2).astype(float)
    mag = sum(np.abs(c) ** 2 for c in u)
    live = mag > 0
    k = np.sqrt(k2[live])
    ratio = mag[live] * 4 * np.pi * k * k / (k ** 4 * np.exp(-2 * (k / 4) ** 2))
    assert ratio.max() / ratio.min() - 1 < 1e-12
    # shell spectrum peaks at k = 4 (d/dk of k^4 exp(-k^2/8) vanishes there)
    shells = np.bincount(np.rint(np.sqrt(k2[live])).astype(int), weights=mag[live])
    assert int(np.argmax(shells)) in (3, 4, 5)


def test_seeds_are_deterministic_and_distinct():
    a = sph.random_solenoidal_field(16, 1000)
    b = sph.random_solenoidal_field(16, 1000)
    c = sph.random_solenoidal_field(16, 1001)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.allclose(a[0], c[0])
    # same per-mode magnitude (fixed by k), different phases
    ma = sum(np.abs(x) ** 2 for x in a)
    mc = sum(np.abs(x) ** 2 for x in c)
    assert np.allclose(ma, mc, rtol=1e-12, atol=0)


def test_field_argument_errors():
    with pytest.raises(RuntimeError, match="even grid size"):
        sph.random_solenoidal_field(7, 1)
    with pytest.raises(RuntimeError, match="energy must be positive"):
        sph.random_solenoidal_field(16, 1, 0.0)

"""Writes the small fixtures under tests/golden/ (run in the build container).

- fft_golden.npz: numpy.fft vectors (an independent FFT implementation) for
  a few 3D shapes, forward (sign -1, unnormalised) and backward (sign +1,
  unnormalised), arrays indexed [z, y, x].
- tg16_fp64.npz: the CPU oracle's 16^3 Taylor-Green state (half spectrum
  [kx][kz][ky]) after 5 RK3 steps of dt = 1e-2 at Re = 1600, with the
  per-step diagnostics; a regression pin for the oracle and a parity vector
  for the GPU library.
- spectral_Re1600_512_head.dat: the first rows of the literature Taylor-Green
  curve shipped with the reference (spectral_Re1600_512.dat: t, E, eps,
  enstrophy), copied as data.
- spectral_Re1600_512.dat: that whole curve (t = 0 .. 19.99 every 0.01), the
  data the reference's comparison script compare_taylor_green.py plots
  against, copied as data.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def fft_vectors():
    rng = np.random.default_rng(20261203)
    out = {}
    for shape in [(16, 16, 16), (6, 8, 12), (4, 8, 48), (12, 12, 12)]:
        tag = "x".join(map(str, shape))
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        out["in_" + tag] = a
        out["fwd_" + tag] = np.fft.fftn(a)
        out["bwd_" + tag] = np.fft.ifftn(a) * a.size
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


def tg16():
    n, dt, steps = 16, 1e-2, 5
    s = oracle.TaylorGreenOracle(n, 1600.0)
    diags = [list(s.diagnostics(0.0).values())[:6]]
    for k in range(steps):
        s.step_rk3(dt)
        diags.append(list(s.diagnostics((k + 1) * dt).values())[:6])
    out = {f"u{c}": oracle.c2c_to_half(s.u_hat(c)) for c in range(3)}
    np.savez_compressed(os.path.join(HERE, "tg16_fp64.npz"), n=n, dt=dt, steps=steps, reynolds=1600.0,
                        diagnostics=np.array(diags), **out)


DNS = "/root/reference/spectral_Re1600_512.dat"


def dns_head(rows=5):
    with open(DNS) as f:
        lines = [ln for ln in f if ln.strip()][:rows]
    with open(os.path.join(HERE, "spectral_Re1600_512_head.dat"), "w") as f:
        f.writelines(lines)


def dns_full():
    with open(DNS) as f:
        lines = [ln for ln in f if ln.strip()]
    with open(os.path.join(HERE, "spectral_Re1600_512.dat"), "w") as f:
        f.writelines(lines)


if __name__ == "__main__":
    fft_vectors()
    tg16()
    if os.path.exists(DNS):
        dns_head()
        dns_full()
    print("fixtures written to", HERE)

"""Writes tests/golden/tg256_fp64_500.npz with the CPU oracle (run in the build
container; 2.5 h at low priority on 6 cores).

Taylor-Green 256^3, nu = 1/1600, dt = 1e-3, 500 fixed RK3 steps, FP64. Stored:
per-step kinetic energy and dissipation (steps 0..500), and at step 500 the
low modes |mx|,|my|,|mz| <= 4 of u_hat in the half-spectrum layout
[kx][kz][ky] (kz, ky ordered -4..4), the shell-energy spectrum
E(k) = sum over modes with round(|m|) = k of |u_hat|^2 counted over the full
C2C spectrum, and the full-field squared norm.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

N, DT, STEPS, K = 256, 1e-3, 500, 4


def low_modes(u_c2c):
    idx = [m % N for m in range(-K, K + 1)]
    half = u_c2c[:, :, : K + 1]  # [kz, ky, kx<=K]
    sub = half[np.ix_(idx, idx, range(K + 1))]  # [kz, ky, kx]
    return np.ascontiguousarray(np.transpose(sub, (2, 0, 1)))  # [kx][kz][ky]


def shells(u_list):
    m = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    kk = np.rint(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)).astype(int)
    e = sum(np.abs(u) ** 2 for u in u_list)
    return np.bincount(kk.ravel(), weights=e.ravel(), minlength=int(kk.max()) + 1)


def main():
    t0 = time.time()
    s = oracle.TaylorGreenOracle(N, 1600.0)
    E = np.zeros(STEPS + 1)
    eps = np.zeros(STEPS + 1)
    d = s.diagnostics(0.0)
    E[0], eps[0] = d["kinetic_energy"], d["dissipation"]
    for k in range(STEPS):
        s.step_rk3(DT)
        d = s.diagnostics((k + 1) * DT)
        E[k + 1], eps[k + 1] = d["kinetic_energy"], d["dissipation"]
        if k % 25 == 0:
            print(f"step {k + 1} E={E[k + 1]:.15f} eps={eps[k + 1]:.15e} {time.time() - t0:.0f}s", flush=True)
    u = [s.u_hat(c) for c in range(3)]
    np.savez_compressed(os.path.join(HERE, "tg256_fp64_500.npz"), n=N, dt=DT, steps=STEPS, reynolds=1600.0,
                        energy=E, dissipation=eps, low_k=K,
                        low0=low_modes(u[0]), low1=low_modes(u[1]), low2=low_modes(u[2]),
                        shells=shells(u), norm2=sum(float((np.abs(x) ** 2).sum()) for x in u))
    print("done", time.time() - t0, flush=True)


if __name__ == "__main__":
    main()

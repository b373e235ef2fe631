"""Writes tests/golden/tg256_fp64_500.npz with the CPU oracle (run in the build
container; about 2.5 h at low priority on 6 cores).

Taylor-Green 256^3, nu = 1/1600, dt = 1e-3, 500 fixed RK3 steps, FP64. Stored:
  - kinetic energy and dissipation after every step (steps 0..500);
  - at steps 100, 200, ..., 500: 32 seeded random linear sketches of each
    component of u_hat over the whole half spectrum [kx 0..128][kz][ky]
    (tests/sketch.py: weights regenerated by the checker from the seed);
  - at step 500: the modes |mx|,|my|,|mz| <= 16 of each component in the same
    layout ([kx 0..16][kz -16..16][ky -16..16]), the modes |m| <= 4 (older
    check), the shell-energy spectrum E(k) = sum over modes with
    round(|m|) = k of |u_hat|^2 over the full C2C spectrum, and the squared
    norm of the full C2C state.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
import sketch  # noqa: E402

N, DT, STEPS = 256, 1e-3, 500
CHECK_EVERY, SKETCHES, SEED, LOW_K, OLD_K = 100, 32, 7, 16, 4


def shells(u_list):
    m = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    kk = np.rint(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)).astype(int)
    e = sum(np.abs(u) ** 2 for u in u_list)
    return np.bincount(kk.ravel(), weights=e.ravel(), minlength=int(kk.max()) + 1)


def main(out=os.path.join(HERE, "tg256_fp64_500.npz")):
    t0 = time.time()
    s = oracle.TaylorGreenOracle(N, 1600.0)
    w = sketch.weights(N * N * (N // 2 + 1), SKETCHES, SEED)
    E = np.zeros(STEPS + 1)
    eps = np.zeros(STEPS + 1)
    d = s.diagnostics(0.0)
    E[0], eps[0] = d["kinetic_energy"], d["dissipation"]
    checks, sk = [], []
    for k in range(STEPS):
        s.step_rk3(DT)
        d = s.diagnostics((k + 1) * DT)
        E[k + 1], eps[k + 1] = d["kinetic_energy"], d["dissipation"]
        if (k + 1) % CHECK_EVERY == 0:
            half = [oracle.c2c_to_half(s.u_hat(c)) for c in range(3)]
            checks.append(k + 1)
            sk.append(np.stack([sketch.sketch(w, h) for h in half]))
        if k % 25 == 0 or (k + 1) % CHECK_EVERY == 0:
            print(f"step {k + 1} E={E[k + 1]:.15f} eps={eps[k + 1]:.15e} {time.time() - t0:.0f}s", flush=True)
    u = [s.u_hat(c) for c in range(3)]
    half = [oracle.c2c_to_half(x) for x in u]
    np.savez_compressed(out, n=N, dt=DT, steps=STEPS, reynolds=1600.0,
                        energy=E, dissipation=eps,
                        sketch_steps=np.array(checks), sketch_seed=SEED, sketches=np.array(sk),
                        low_k=OLD_K, low0=sketch.low_block(half[0], OLD_K), low1=sketch.low_block(half[1], OLD_K),
                        low2=sketch.low_block(half[2], OLD_K),
                        block_k=LOW_K, block=np.stack([sketch.low_block(h, LOW_K) for h in half]),
                        shells=shells(u), norm2=sum(float((np.abs(x) ** 2).sum()) for x in u))
    print("done", time.time() - t0, flush=True)


if __name__ == "__main__":
    main()

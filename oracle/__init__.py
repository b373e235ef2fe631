"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates the reference's AVA CPU path
(src/ava_taylor_green_impl.hpp:658-1410, FFTW backend) in C++; see oracle.hpp.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, always as the checker of the GPU library, never as its fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_fft3d.argtypes = [ctypes.c_int] * 3 + [dp, dp, ctypes.c_int, ctypes.c_int]
        L.orc_solver_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.orc_solver_destroy.argtypes = [ctypes.c_void_p]
        L.orc_solver_destroy.restype = None
        L.orc_solver_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_solver_diagnostics.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
        L.orc_solver_max_velocity.argtypes = [ctypes.c_void_p, dp]
        L.orc_solver_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.orc_solver_step_seconds.argtypes = [ctypes.c_void_p]
        L.orc_solver_step_seconds.restype = ctypes.c_double
        L.orc_solver_get_uhat.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.orc_solver_set_uhat.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.orc_run_taylor_green.argtypes = ([ctypes.c_int] * 6 + [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_char_p,
                                           dp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int)])
        L.orc_mode_tables.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3 + [
            ctypes.POINTER(ctypes.c_ubyte), dp]
        L.orc_dealiased_product.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_solver_first_stage.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
        L.orc_legacy_run.argtypes = [ctypes.POINTER(ctypes.c_int), dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, dp, dp, dp, dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


DIAG_KEYS = ("t", "kinetic_energy", "dissipation", "enstrophy", "max_spectral_divergence",
             "max_velocity", "dt", "cfl", "main_dt", "main_cfl")
COUNTER_KEYS = ("forward_calls", "backward_calls", "step_calls", "rhs_calls",
                "diagnostics_calls", "max_velocity_calls")


def fft3d(a: np.ndarray, sign: int, single: bool = False) -> np.ndarray:
    """Unnormalised C2C transform of a complex [nz, ny, nx] array."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    nz, ny, nx = a.shape
    out = np.empty_like(a)
    _check(lib().orc_fft3d(nx, ny, nz, _dptr(a.view(np.float64)), _dptr(out.view(np.float64)), sign,
                           1 if single else 0))
    return out


class TaylorGreenOracle:
    """Host restatement of TaylorGreenSolver (C2C storage, x fastest)."""

    def __init__(self, n: int, reynolds: float, single: bool = False, initialise: bool = True):
        self.n = n
        self.single = single
        h = ctypes.c_void_p()
        _check(lib().orc_solver_create(n, reynolds, 1 if single else 0, 1 if initialise else 0, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            lib().orc_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step_rk3(self, dt: float) -> None:
        _check(lib().orc_solver_step(self._h, dt))

    def diagnostics(self, t: float) -> dict:
        out = np.zeros(10)
        _check(lib().orc_solver_diagnostics(self._h, t, _dptr(out)))
        return dict(zip(DIAG_KEYS, out.tolist()))

    def max_velocity(self) -> float:
        v = ctypes.c_double()
        _check(lib().orc_solver_max_velocity(self._h, ctypes.byref(v)))
        return v.value

    def counters(self) -> dict:
        out = (ctypes.c_longlong * 6)()
        _check(lib().orc_solver_counters(self._h, out))
        return dict(zip(COUNTER_KEYS, list(out)))

    def first_stage_seconds(self, dt: float) -> float:
        """Run only RK3 stage 1 and return its wall time (CPU-baseline sample)."""
        sec = ctypes.c_double()
        _check(lib().orc_solver_first_stage(self._h, dt, ctypes.byref(sec)))
        return sec.value

    def step_seconds(self) -> float:
        return lib().orc_solver_step_seconds(self._h)

    def u_hat(self, c: int) -> np.ndarray:
        """Component c of the C2C state as complex [nz, ny, nx]."""
        n = self.n
        out = np.empty((n, n, n), np.complex128)
        _check(lib().orc_solver_get_uhat(self._h, c, _dptr(out.view(np.float64))))
        return out

    def set_u_hat(self, c: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.complex128)
        _check(lib().orc_solver_set_uhat(self._h, c, _dptr(a.view(np.float64))))


def run_taylor_green(n=64, steps=80, max_steps=10000000, output_every=1, print_every=10, adaptive_cfl=False,
                     dt=0.0025, cfl=0.4, t_end=None, output_dt=0.1, reynolds=1600.0, single=False,
                     output_dir="output_oracle"):
    """Driver restatement; returns (final_row, counters, steps_taken)."""
    if t_end is None:
        t_end = steps * dt
    row = np.zeros(10)
    cnt = (ctypes.c_longlong * 6)()
    st = ctypes.c_int()
    _check(lib().orc_run_taylor_green(n, steps, max_steps, output_every, print_every, 1 if adaptive_cfl else 0,
                                      dt, cfl, t_end, output_dt, reynolds, 1 if single else 0,
                                      output_dir.encode(), _dptr(row), cnt, ctypes.byref(st)))
    return dict(zip(DIAG_KEYS, row.tolist())), dict(zip(COUNTER_KEYS, list(cnt))), st.value


def set_threads(threads: int) -> None:
    """Threads of the oracle's OpenMP loops (1 = the reference's serial build)."""
    _check(lib().orc_set_threads(int(threads)))


def legacy_run(shape, lengths, nu: float, dealias: bool, steps: int, dt: float, phys=None):
    """Restatement of the reference's legacy SpectralNavierStokes
    (src/spectral_solver.cpp:35-409, legacy.hpp): physical velocity `phys`
    (three real [nz, ny, nx] arrays; None = its taylor_green(1, 1)) ->
    to_spectral, project, truncate, then `steps` RK3 steps of dt.

    Returns (diagnostics [(steps+1), 5] rows t, E, dissipation, enstrophy,
    max divergence; nonlinear_rhs of the initial state; final state), the
    last two as three complex C2C [nz, ny, nx] arrays each."""
    nx, ny, nz = (int(v) for v in shape)
    npts = nx * ny * nz
    shp = (ctypes.c_int * 3)(nx, ny, nz)
    lens = np.ascontiguousarray(lengths, np.float64)
    pin = None
    if phys is not None:
        pin = np.ascontiguousarray(np.stack([np.asarray(p, np.float64).reshape(nz, ny, nx) for p in phys]))
    diag = np.zeros((steps + 1, 5))
    rhs0 = np.zeros((3, nz, ny, nx), np.complex128)
    state = np.zeros((3, nz, ny, nx), np.complex128)
    _check(lib().orc_legacy_run(shp, _dptr(lens), float(nu), 1 if dealias else 0, int(steps), float(dt),
                                _dptr(pin) if pin is not None else None, _dptr(diag),
                                _dptr(rhs0.view(np.float64)), _dptr(state.view(np.float64))))
    assert npts == state[0].size
    return diag, list(rhs0), list(state)


def mode_tables(n: int):
    """(mx, my, mz, keep, phase) of the reference's metadata, shaped [nz, ny, nx]."""
    t = n * n * n
    mx, my, mz = (np.empty(t, np.int32) for _ in range(3))
    keep = np.empty(t, np.uint8)
    ph = np.empty(t, np.complex128)
    ip = ctypes.POINTER(ctypes.c_int)
    _check(lib().orc_mode_tables(n, mx.ctypes.data_as(ip), my.ctypes.data_as(ip), mz.ctypes.data_as(ip),
                                 keep.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), _dptr(ph.view(np.float64))))
    shp = (n, n, n)
    return mx.reshape(shp), my.reshape(shp), mz.reshape(shp), keep.reshape(shp).astype(bool), ph.reshape(shp)


def dealiased_product(a_hat: np.ndarray, b_hat: np.ndarray) -> np.ndarray:
    """Dealiased product spectrum (one accumulate_product combine) of two C2C spectra."""
    n = a_hat.shape[0]
    a = np.ascontiguousarray(a_hat, np.complex128)
    b = np.ascontiguousarray(b_hat, np.complex128)
    out = np.empty_like(a)
    _check(lib().orc_dealiased_product(n, _dptr(a.view(np.float64)), _dptr(b.view(np.float64)),
                                       _dptr(out.view(np.float64))))
    return out


def c2c_to_half(a: np.ndarray) -> np.ndarray:
    """Restrict a C2C [kz, ky, kx] spectrum to the half-spectrum layout
    [kx in 0..n/2][kz][ky] used by the GPU library."""
    nx = a.shape[2]
    return np.ascontiguousarray(np.transpose(a[:, :, : nx // 2 + 1], (2, 0, 1)))


def half_to_c2c(h: np.ndarray) -> np.ndarray:
    """Inverse of c2c_to_half for a Hermitian spectrum: the full C2C
    [kz, ky, kx] array whose kx > n/2 half is conj(u_hat(n - kx, -ky, -kz))."""
    nh, nz, ny = h.shape
    nx = 2 * (nh - 1)
    full = np.zeros((nz, ny, nx), np.complex128)
    full[:, :, :nh] = np.transpose(h, (1, 2, 0))
    kx = np.arange(nh, nx)
    src = full[:, :, nx - kx]  # [kz, ky, kx']
    src = np.roll(np.flip(src, axis=(0, 1)), shift=(1, 1), axis=(0, 1))  # (-kz, -ky)
    full[:, :, nh:] = np.conj(src)
    return full


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea & Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def random_solenoidal_field(n: int, seed: int, energy: float = 0.5):
    """numpy restatement of the ensemble's initial field (three complex
    [kx][kz][ky] half-spectrum arrays).

    Every kept mode (4|m|^2 <= n^2, no Nyquist index, m != 0; on the kx = 0
    plane only the half my > 0 or (my = 0, mz > 0), the other half being its
    conjugate) gets three unit phasors e^{i theta_c}, theta_c = 2 pi r / 2^53
    with r = SplitMix64(seed * 0x2545F4914F6CDD1D + 3 key + c) >> 11 and
    key = ((mx + n)(2n + 1) + my + n)(2n + 1) + mz + n; the vector is projected
    perpendicular to k and scaled to |u_hat|^2 = E(k) / (4 pi k^2) with
    E(k) = k^4 exp(-2 (k/4)^2); the whole field is then scaled to
    (1/2) sum_full |u_hat|^2 / n^6 = energy.
    """
    H = n // 2 + 1
    kx = np.arange(H)[:, None, None]
    sg = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    mz = sg[None, :, None]
    my = sg[None, None, :]
    mx, my, mz = np.broadcast_arrays(kx, my, mz)
    k2 = mx * mx + my * my + mz * mz
    gen = (k2 > 0) & (4 * k2 <= n * n) & (mx != n // 2) & (my != n // 2) & (mz != n // 2)
    gen &= (mx > 0) | (my > 0) | ((my == 0) & (mz > 0))
    key = ((mx + n).astype(np.uint64) * np.uint64(2 * n + 1) + (my + n).astype(np.uint64)) * np.uint64(2 * n + 1) \
        + (mz + n).astype(np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(0x2545F4914F6CDD1D) + key * np.uint64(3)
    v = []
    for c in range(3):
        with np.errstate(over="ignore"):
            r = _splitmix64(base + np.uint64(c))
        th = 2.0 * np.pi * (r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        v.append(np.cos(th) + 1j * np.sin(th))
    kk = [mx.astype(np.float64), my.astype(np.float64), mz.astype(np.float64)]
    k2f = np.where(gen, k2, 1).astype(np.float64)
    dot = kk[0] * v[0] + kk[1] * v[1] + kk[2] * v[2]
    v = [v[c] - kk[c] * dot / k2f for c in range(3)]
    norm = sum(np.abs(x) ** 2 for x in v)
    gen &= norm >= 1e-20
    k = np.sqrt(k2f)
    ek = k ** 4 * np.exp(-2.0 * (k / 4.0) ** 2)
    amp = np.where(gen, np.sqrt(ek / (4.0 * np.pi * k * k) / np.where(gen, norm, 1.0)), 0.0)
    u = [np.where(gen, amp * v[c], 0.0) for c in range(3)]
    # kx = 0 plane: partner (-my, -mz) is the conjugate
    for c in range(3):
        p = u[c][0]
        part = np.conj(np.roll(np.flip(p, axis=(0, 1)), shift=(1, 1), axis=(0, 1)))
        u[c][0] = np.where(gen[0], p, part)
    total = 2.0 * sum(np.sum(np.where(gen, np.abs(u[c]) ** 2, 0.0)) for c in range(3))
    scale = np.sqrt(energy * float(n) ** 6 / (0.5 * total))
    return [np.ascontiguousarray(x * scale) for x in u]

"""MI355X pseudo-spectral Navier-Stokes RK3 hot path — Python mirror.

Thin ctypes binding of libspectral_hip.so (the C ABI declared in
include/spectral_hip.h) with the reference's class shapes:

    HipFft3D            <- spectral::FftwFft3D (src/fftw_fft.hpp:13-81)
    TaylorGreenSolver   <- spectral::ava_solver::TaylorGreenSolver
                           (src/ava_taylor_green.hpp:36-53)
    Diagnostics, RunConfig, run_taylor_green
                        <- src/ava_taylor_green.hpp:8-34, 55
    Ensemble            a batch of independent TaylorGreenSolver-style solves
                        (uneven sizes) stepped in shared launches

Every compute call goes to the HIP library; if the library has not been built
(``python -m spectral_solver_amd.build``) the first use raises.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Callable, Optional

import numpy as np

__all__ = ["HipFft3D", "TaylorGreenSolver", "Diagnostics", "RunConfig", "run_taylor_green", "Ensemble",
           "random_solenoidal_field", "SpectralMode", "library_path", "load_library", "device_count",
           "copy_bandwidth", "FP32", "FP64"]

FP32 = 0
FP64 = 1
_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libspectral_hip.so")

_lib = None


class _Diag(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("t", "kinetic_energy", "dissipation", "enstrophy",
                                               "max_spectral_divergence", "max_velocity", "dt", "cfl",
                                               "main_dt", "main_cfl")]


class _Profile(ctypes.Structure):
    _fields_ = [("step", ctypes.c_double), ("diagnostics", ctypes.c_double), ("max_velocity", ctypes.c_double),
                ("step_calls", ctypes.c_longlong), ("rhs_calls", ctypes.c_longlong),
                ("diagnostics_calls", ctypes.c_longlong), ("max_velocity_calls", ctypes.c_longlong),
                ("forward_calls", ctypes.c_longlong), ("backward_calls", ctypes.c_longlong)]


class _RunConfig(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("steps", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("output_every", ctypes.c_int), ("print_every", ctypes.c_int), ("dt", ctypes.c_double),
                ("cfl", ctypes.c_double), ("t_end", ctypes.c_double), ("output_dt", ctypes.c_double),
                ("reynolds", ctypes.c_double), ("adaptive_cfl", ctypes.c_int), ("precision", ctypes.c_int),
                ("output_dir", ctypes.c_char_p)]


# (name, restype, argtypes) of every symbol include/spectral_hip.h declares
_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_I64P = ctypes.POINTER(ctypes.c_int64)
SYMBOLS = [
    ("sph_last_error", ctypes.c_char_p, []),
    ("sph_device_count", ctypes.c_int, [_IP]),
    ("sph_supported_length", ctypes.c_int, [ctypes.c_int]),
    ("sph_fft_create", ctypes.c_int, [_IP, _DP, ctypes.c_int, ctypes.POINTER(_VP)]),
    ("sph_fft_destroy", ctypes.c_int, [_VP]),
    ("sph_fft_sizes", ctypes.c_int, [_VP, _I64P, _I64P]),
    ("sph_fft_spectral_modes", ctypes.c_int, [_VP, _IP, _IP, _IP]),
    ("sph_fft_forward", ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ("sph_fft_backward", ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ("sph_fft_backward_noalloc", ctypes.c_int, [_IP, ctypes.c_int, _VP, _VP, _VP]),
    ("sph_fft_forward_host", ctypes.c_int, [_VP, _DP, _DP]),
    ("sph_fft_backward_host", ctypes.c_int, [_VP, _DP, _DP]),
    ("sph_solver_create", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_VP)]),
    ("sph_solver_destroy", ctypes.c_int, [_VP]),
    ("sph_solver_step_rk3", ctypes.c_int, [_VP, ctypes.c_double]),
    ("sph_solver_diagnostics", ctypes.c_int, [_VP, ctypes.c_double, ctypes.POINTER(_Diag)]),
    ("sph_solver_max_velocity", ctypes.c_int, [_VP, _DP]),
    ("sph_solver_sizes", ctypes.c_int, [_VP, _IP, _I64P]),
    ("sph_solver_get_uhat", ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    ("sph_solver_set_uhat", ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    ("sph_solver_profile", ctypes.c_int, [_VP, ctypes.POINTER(_Profile)]),
    ("sph_solver_step_times", ctypes.c_int, [_VP, _DP, ctypes.c_int, _IP]),
    ("sph_solver_print_profile", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    ("sph_solver_set_graph", ctypes.c_int, [_VP, ctypes.c_int]),
    ("sph_solver_synchronize", ctypes.c_int, [_VP]),
    ("sph_solver_time_passes", ctypes.c_int, [_VP, ctypes.c_int, _DP, _DP]),
    ("sph_solver_time_step_passes", ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, _DP, _DP]),
    ("sph_solver_step_adaptive", ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_double, _DP, _DP, _IP]),
    ("sph_solver_tuning", ctypes.c_int, [_VP, _IP, _IP, _DP]),
    ("sph_solver_layout", ctypes.c_int, [_VP, _IP, _DP]),
    ("sph_solver_tuned_pass", ctypes.c_int, [_VP, ctypes.c_int, _IP, _IP, _DP]),
    ("sph_solver_radix_plan", ctypes.c_int, [_VP, ctypes.c_int, _IP, _IP]),
    ("sph_radix_plan", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _IP, _IP]),
    ("sph_measure_copy_bandwidth", ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _DP]),
    ("sph_run_taylor_green", ctypes.c_int, [ctypes.POINTER(_RunConfig), ctypes.POINTER(_Diag)]),
    ("sph_ensemble_create", ctypes.c_int, [ctypes.c_int, _IP, _DP, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.c_double, ctypes.c_int, ctypes.POINTER(_VP)]),
    ("sph_ensemble_destroy", ctypes.c_int, [_VP]),
    ("sph_ensemble_step_rk3", ctypes.c_int, [_VP, _DP]),
    ("sph_ensemble_diagnostics", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_Diag)]),
    ("sph_ensemble_max_velocity", ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    ("sph_ensemble_get_uhat", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _DP]),
    ("sph_ensemble_set_uhat", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _DP]),
    ("sph_ensemble_set_graph", ctypes.c_int, [_VP, ctypes.c_int]),
    ("sph_ensemble_synchronize", ctypes.c_int, [_VP]),
    ("sph_ensemble_step_seconds", ctypes.c_int, [_VP, _DP, _I64P]),
    ("sph_ensemble_time_passes", ctypes.c_int, [_VP, ctypes.c_int, _DP, _DP]),
    ("sph_ensemble_launch_mode", ctypes.c_int, [_VP, _IP, _DP]),
    ("sph_random_solenoidal_field", ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, _DP, _DP, _DP]),
]


def load_library():
    """Load libspectral_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise RuntimeError(f"{library_path} is missing: build it with `python -m spectral_solver_amd.build`")
        lib = ctypes.CDLL(library_path)
        for name, res, args in SYMBOLS:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(load_library().sph_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


# variants per tuned pass in the autotuner's reports (kMaxVariants)
MAX_VARIANTS = 4


def radix_plan(n: int, order: int = 0) -> list:
    """Radix stages of the z-type passes for transform length n: order 0 the
    default plan, 1 the alternative order the autotuner also times ([] when n
    is not a compiled length). Needs no GPU."""
    r, c = np.zeros(16, np.int32), ctypes.c_int()
    _check(load_library().sph_radix_plan(int(n), int(order), r.ctypes.data_as(_IP), ctypes.byref(c)))
    return [int(x) for x in r[:c.value]]


def device_count() -> int:
    c = ctypes.c_int()
    _check(load_library().sph_device_count(ctypes.byref(c)))
    return c.value


def copy_bandwidth(nbytes: int = 2 << 30, reps: int = 10) -> float:
    """Measured device copy bandwidth in GB/s (read + write bytes)."""
    g = ctypes.c_double()
    _check(load_library().sph_measure_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(g)))
    return g.value


def _precision(p) -> int:
    if p in (FP32, "fp32", "float32", "single", np.float32):
        return FP32
    if p in (FP64, "fp64", "float64", "double", np.float64):
        return FP64
    raise ValueError("precision must be fp32 or fp64")


@dataclasses.dataclass
class SpectralMode:
    kx: float
    ky: float
    kz: float
    k2: float
    mx: int
    my: int
    mz: int


@dataclasses.dataclass
class Diagnostics:
    t: float = 0.0
    kinetic_energy: float = 0.0
    dissipation: float = 0.0
    enstrophy: float = 0.0
    max_spectral_divergence: float = 0.0
    max_velocity: float = 0.0
    dt: float = 0.0
    cfl: float = 0.0
    main_dt: float = 0.0
    main_cfl: float = 0.0

    @classmethod
    def _from(cls, d: _Diag) -> "Diagnostics":
        return cls(*[getattr(d, f[0]) for f in _Diag._fields_])


@dataclasses.dataclass
class RunConfig:
    n: int = 64
    steps: int = 80
    max_steps: int = 10000000
    output_every: int = 1
    print_every: int = 10
    dt: float = 0.0025
    cfl: float = 0.4
    t_end: float = 0.2
    output_dt: float = 0.1
    reynolds: float = 1600.0
    adaptive_cfl: bool = False
    output_dir: str = "output_ava"
    precision: str = "fp64"


class HipFft3D:
    """FftwFft3D-shaped backend: real [nz, ny, nx] <-> half spectrum [kx][kz][ky]."""

    def __init__(self, shape, lengths, comm=None, precision="fp64"):
        self._shape = tuple(int(s) for s in shape)
        self._lengths = tuple(float(v) for v in lengths)
        self.precision = _precision(precision)
        sh = (ctypes.c_int * 3)(*self._shape)
        ln = (ctypes.c_double * 3)(*self._lengths)
        h = _VP()
        _check(load_library().sph_fft_create(sh, ln, self.precision, ctypes.byref(h)))
        self._h = h
        npts, nm = ctypes.c_int64(), ctypes.c_int64()
        _check(load_library().sph_fft_sizes(self._h, ctypes.byref(npts), ctypes.byref(nm)))
        self._modes = nm.value
        self.mx = np.empty(nm.value, np.int32)
        self.my = np.empty(nm.value, np.int32)
        self.mz = np.empty(nm.value, np.int32)
        _check(load_library().sph_fft_spectral_modes(self._h, self.mx.ctypes.data_as(_IP),
                                                      self.my.ctypes.data_as(_IP), self.mz.ctypes.data_as(_IP)))

    def close(self):
        if getattr(self, "_h", None):
            load_library().sph_fft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self):
        return self._shape

    def lengths(self):
        return self._lengths

    def physical_size(self) -> int:
        return self._shape[0] * self._shape[1] * self._shape[2]

    def global_physical_size(self) -> int:
        return self.physical_size()

    def spectral_storage_size(self) -> int:
        return 2 * self._modes

    def comm(self):
        return None

    def rank(self) -> int:
        return 0

    def comm_size(self) -> int:
        return 1

    def for_each_physical(self, func: Callable) -> None:
        nx, ny, nz = self._shape
        for iz in range(nz):
            for iy in range(ny):
                for ix in range(nx):
                    func(ix + nx * (iy + ny * iz), ix, iy, iz)

    def for_each_spectral(self, func: Callable) -> None:
        two_pi = 2.0 * np.pi
        for i in range(self._modes):
            mx, my, mz = int(self.mx[i]), int(self.my[i]), int(self.mz[i])
            kx, ky, kz = two_pi * mx / self._lengths[0], two_pi * my / self._lengths[1], two_pi * mz / self._lengths[2]
            func(2 * i, SpectralMode(kx, ky, kz, kx * kx + ky * ky + kz * kz, mx, my, mz))

    def forward(self, physical) -> np.ndarray:
        a = np.ascontiguousarray(physical, dtype=np.float64).reshape(-1)
        if a.size != self.physical_size():
            raise RuntimeError("physical field has wrong size")
        out = np.empty(self.spectral_storage_size(), np.float64)
        _check(load_library().sph_fft_forward_host(self._h, _dptr(a), _dptr(out)))
        return out

    def backward(self, spectral) -> np.ndarray:
        a = np.ascontiguousarray(spectral, dtype=np.float64).reshape(-1)
        if a.size != self.spectral_storage_size():
            raise RuntimeError("spectral field has wrong storage size")
        out = np.empty(self.physical_size(), np.float64)
        _check(load_library().sph_fft_backward_host(self._h, _dptr(a), _dptr(out)))
        return out

    def forward_device(self, physical_ptr: int, spectral_ptr: int, stream: int = 0) -> None:
        _check(load_library().sph_fft_forward(self._h, physical_ptr, spectral_ptr, stream or None))

    def backward_device(self, spectral_ptr: int, physical_ptr: int, stream: int = 0) -> None:
        _check(load_library().sph_fft_backward(self._h, spectral_ptr, physical_ptr, stream or None))

    # numpy conveniences ------------------------------------------------------
    def half_shape(self):
        """(kx, kz, ky) extents of the spectral storage."""
        nx, ny, nz = self._shape
        return (nx // 2 + 1, nz, ny)


class TaylorGreenSolver:
    """spectral::ava_solver::TaylorGreenSolver on the GPU."""

    def __init__(self, n: int, reynolds: float, precision="fp64"):
        self.n = int(n)
        self.reynolds = float(reynolds)
        self.precision = _precision(precision)
        h = _VP()
        _check(load_library().sph_solver_create(self.n, self.reynolds, self.precision, ctypes.byref(h)))
        self._h = h
        nn, nm = ctypes.c_int(), ctypes.c_int64()
        _check(load_library().sph_solver_sizes(self._h, ctypes.byref(nn), ctypes.byref(nm)))
        self.spectral_modes = nm.value

    def close(self):
        if getattr(self, "_h", None):
            load_library().sph_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step_rk3(self, dt: float) -> None:
        _check(load_library().sph_solver_step_rk3(self._h, float(dt)))

    def diagnostics(self, t: float) -> Diagnostics:
        d = _Diag()
        _check(load_library().sph_solver_diagnostics(self._h, float(t), ctypes.byref(d)))
        return Diagnostics._from(d)

    def max_velocity(self) -> float:
        v = ctypes.c_double()
        _check(load_library().sph_solver_max_velocity(self._h, ctypes.byref(v)))
        return v.value

    def print_profile(self, rank: int, init_time: float, run_time: float) -> None:
        _check(load_library().sph_solver_print_profile(self._h, rank, init_time, run_time))

    def synchronize(self) -> None:
        _check(load_library().sph_solver_synchronize(self._h))

    def set_graph(self, enabled: bool) -> None:
        _check(load_library().sph_solver_set_graph(self._h, 1 if enabled else 0))

    def profile(self) -> dict:
        p = _Profile()
        _check(load_library().sph_solver_profile(self._h, ctypes.byref(p)))
        return {f[0]: getattr(p, f[0]) for f in _Profile._fields_}

    def step_times(self, last: int = 4096) -> np.ndarray:
        """GPU time (ms, HIP events around each step) of up to `last` of
        the most recent RK3 steps, oldest first."""
        ms = np.zeros(max(int(last), 0))
        n = ctypes.c_int()
        _check(load_library().sph_solver_step_times(self._h, _dptr(ms), len(ms), ctypes.byref(n)))
        return ms[:n.value].copy()

    def time_passes(self, reps: int = 5):
        ms = np.zeros(len(_PASS_NAMES))
        by = np.zeros(len(_PASS_NAMES))
        _check(load_library().sph_solver_time_passes(self._h, int(reps), _dptr(ms), _dptr(by)))
        return {k: (float(m), float(b)) for k, m, b in zip(_PASS_NAMES, ms, by)}

    def time_step_passes(self, dt: float, steps: int = 3):
        """Pass kernels timed inside `steps` real steps (direct launches, HIP
        events between them; the state advances): {pass: (mean ms, bytes)}
        for the four in-step passes, plus "inverse_z" (0 unless the velocity
        pairs had to be formed first) and "step_ms"."""
        ms = np.zeros(len(_PASS_NAMES) + 1)
        by = np.zeros(len(_PASS_NAMES))
        _check(load_library().sph_solver_time_step_passes(self._h, float(dt), int(steps), _dptr(ms), _dptr(by)))
        out = {k: (float(m), float(b)) for k, m, b in zip(_PASS_NAMES, ms, by)}
        out["step_ms"] = float(ms[-1])
        return out

    def step_adaptive(self, dt_max: float, cfl: float):
        """One adaptive step (dt from the CFL condition): returns
        (max |u| before the step, dt taken, whether max |u| came from the
        step's own x pass)."""
        v, d, f = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(load_library().sph_solver_step_adaptive(self._h, float(dt_max), float(cfl), ctypes.byref(v),
                                                       ctypes.byref(d), ctypes.byref(f)))
        return v.value, d.value, bool(f.value)

    def layout(self):
        """(work layout kept: 0 line, 1 transposed, 2 ky-blocked, [ms of
        passes 1-3 in the line layout, in the transposed layout, ms of passes
        1-4 in the ky-blocked layout, with the faster of the other two])."""
        t, ms = ctypes.c_int(), np.zeros(4)
        _check(load_library().sph_solver_layout(self._h, ctypes.byref(t), _dptr(ms)))
        return int(t.value), [float(x) for x in ms]

    def tuning(self) -> dict:
        """Autotuner choice per tuned pass: (variant, lines per tile, ms of
        each variant)."""
        var = np.zeros(3, np.int32)
        lines = np.zeros(3, np.int32)
        ms = np.zeros(3 * MAX_VARIANTS)
        _check(load_library().sph_solver_tuning(self._h, var.ctypes.data_as(_IP), lines.ctypes.data_as(_IP),
                                                _dptr(ms)))
        names = ("inverse_y", "x_products", "forward_y")
        V = MAX_VARIANTS
        out = {k: (int(var[p]), int(lines[p]), [float(x) for x in ms[V * p:V * p + V]]) for p, k in enumerate(names)}
        v, l, m4 = ctypes.c_int(), ctypes.c_int(), np.zeros(V)
        _check(load_library().sph_solver_tuned_pass(self._h, 3, ctypes.byref(v), ctypes.byref(l), _dptr(m4)))
        out["forward_z_update"] = (v.value, l.value, [float(x) for x in m4])
        return out

    def radix_plans(self) -> dict:
        """Radix stages of the transforms in the kept variant of each tuned
        pass (first stage first)."""
        out = {}
        for p, k in enumerate(("inverse_y", "x_products", "forward_y", "forward_z_update")):
            r, c = np.zeros(16, np.int32), ctypes.c_int()
            _check(load_library().sph_solver_radix_plan(self._h, p, r.ctypes.data_as(_IP), ctypes.byref(c)))
            out[k] = [int(x) for x in r[:c.value]]
        return out

    def u_hat(self, c: int) -> np.ndarray:
        """Component c as complex [kx in 0..n/2][kz][ky]."""
        n = self.n
        out = np.empty((n // 2 + 1, n, n), np.complex128)
        _check(load_library().sph_solver_get_uhat(self._h, int(c), _dptr(out.view(np.float64))))
        return out

    def set_u_hat(self, c: int, a: np.ndarray) -> None:
        n = self.n
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != (n // 2 + 1, n, n):
            raise ValueError("u_hat component must have shape (n/2+1, n, n)")
        _check(load_library().sph_solver_set_uhat(self._h, int(c), _dptr(a.view(np.float64))))


_PASS_NAMES = ("inverse_z", "inverse_y", "x_products", "forward_y", "forward_z_update")


def random_solenoidal_field(n: int, seed: int, energy: float = 0.5):
    """The ensemble's seeded initial field: three complex [kx][kz][ky] arrays.

    Random phases, |u_hat|^2 shaped to E(k) ~ k^4 exp(-2 (k/4)^2), divergence
    free, kept modes only, kinetic energy `energy`. Host-side (no GPU).
    """
    n = int(n)
    u = [np.empty((n // 2 + 1, n, n), np.complex128) for _ in range(3)]
    _check(load_library().sph_random_solenoidal_field(n, ctypes.c_uint64(int(seed)), float(energy),
                                                      *[_dptr(a.view(np.float64)) for a in u]))
    return u


class Ensemble:
    """A batch of independent periodic solves stepped together.

    Members may differ in grid size; members of one size share every pass
    launch and the whole step is one hipGraph. Each member takes the same RK3
    step as TaylorGreenSolver (impl.hpp:976-1039) from a seeded random-phase
    solenoidal initial field.
    """

    def __init__(self, sizes, nu, seeds=None, energy: float = 0.5, precision="fp32"):
        sizes = [int(s) for s in sizes]
        count = len(sizes)
        nu = np.broadcast_to(np.asarray(nu, np.float64), (count,)).copy()
        seeds = list(range(count)) if seeds is None else [int(s) for s in seeds]
        if len(seeds) != count:
            raise ValueError("one seed per member")
        self.sizes = sizes
        self.nu = nu
        self.precision = _precision(precision)
        n_arr = (ctypes.c_int * count)(*sizes)
        s_arr = (ctypes.c_uint64 * count)(*seeds)
        h = _VP()
        _check(load_library().sph_ensemble_create(count, n_arr, _dptr(nu), s_arr, float(energy), self.precision,
                                                  ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().sph_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.sizes)

    def step_rk3(self, dt) -> None:
        d = np.broadcast_to(np.asarray(dt, np.float64), (len(self.sizes),)).copy()
        _check(load_library().sph_ensemble_step_rk3(self._h, _dptr(d)))

    def diagnostics(self, member: int, t: float = 0.0) -> Diagnostics:
        d = _Diag()
        _check(load_library().sph_ensemble_diagnostics(self._h, int(member), float(t), ctypes.byref(d)))
        return Diagnostics._from(d)

    def max_velocity(self, member: int) -> float:
        v = ctypes.c_double()
        _check(load_library().sph_ensemble_max_velocity(self._h, int(member), ctypes.byref(v)))
        return v.value

    def u_hat(self, member: int, c: int) -> np.ndarray:
        n = self.sizes[member]
        out = np.empty((n // 2 + 1, n, n), np.complex128)
        _check(load_library().sph_ensemble_get_uhat(self._h, int(member), int(c), _dptr(out.view(np.float64))))
        return out

    def set_u_hat(self, member: int, c: int, a: np.ndarray) -> None:
        n = self.sizes[member]
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != (n // 2 + 1, n, n):
            raise ValueError("u_hat component must have shape (n/2+1, n, n)")
        _check(load_library().sph_ensemble_set_uhat(self._h, int(member), int(c), _dptr(a.view(np.float64))))

    def set_graph(self, enabled: bool) -> None:
        _check(load_library().sph_ensemble_set_graph(self._h, 1 if enabled else 0))

    def synchronize(self) -> None:
        _check(load_library().sph_ensemble_synchronize(self._h))

    def step_seconds(self):
        s, k = ctypes.c_double(), ctypes.c_int64()
        _check(load_library().sph_ensemble_step_seconds(self._h, ctypes.byref(s), ctypes.byref(k)))
        return s.value, k.value

    def time_passes(self, reps: int = 5):
        ms = np.zeros(len(_PASS_NAMES))
        by = np.zeros(len(_PASS_NAMES))
        _check(load_library().sph_ensemble_time_passes(self._h, int(reps), _dptr(ms), _dptr(by)))
        return {k: (float(m), float(b)) for k, m, b in zip(_PASS_NAMES, ms, by)}

    def launch_mode(self):
        """(launch structure in use: 0 per-size chains, 1 one launch per
        pass, 2 one launch per stage; [step ms with each, 0 where not
        timed])."""
        f, ms = ctypes.c_int(), np.zeros(3)
        _check(load_library().sph_ensemble_launch_mode(self._h, ctypes.byref(f), _dptr(ms)))
        return int(f.value), [float(x) for x in ms]


def run_taylor_green(config: RunConfig) -> Diagnostics:
    """run_taylor_green (src/ava_taylor_green_impl.hpp:1310-1410) on the GPU."""
    c = _RunConfig(config.n, config.steps, config.max_steps, config.output_every, config.print_every, config.dt,
                   config.cfl, config.t_end, config.output_dt, config.reynolds, 1 if config.adaptive_cfl else 0,
                   _precision(config.precision), config.output_dir.encode())
    d = _Diag()
    _check(load_library().sph_run_taylor_green(ctypes.byref(c), ctypes.byref(d)))
    return Diagnostics._from(d)

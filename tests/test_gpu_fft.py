"""GPU parity of the HIP real<->complex 3D transforms against the oracle.

The oracle's C2C transform (pinned in test_oracle.py) restricted to
kx <= nx/2 is the expected half spectrum. For the backward transform the
expectation is Re(C2C backward)/N of the conjugate-symmetric completion of the
half spectrum, which keeps the reference's treatment of the kx = 0 and
kx = nx/2 planes (their values enter through Re, impl.hpp:555-561).
"""
import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (12, 12, 12), (32, 16, 8), (8, 24, 48), (64, 64, 64), (96, 32, 16), (128, 8, 12),
          (16, 8, 96), (24, 12, 24),
          # the x and z passes' long plans (1024 = 8*4*4*8, 256 = 8*4*8) along x and z
          (1024, 8, 12), (12, 8, 1024), (256, 16, 256),
          # lengths with a factor 5, and 2^a 3 beyond 96 (radix-4 z/x plans
          # with a radix 2 or 6 and the 3s / 5s: 20 or 12 values per thread)
          (80, 16, 160), (320, 8, 192), (16, 384, 80), (640, 12, 8), (12, 8, 768), (768, 12, 640)]


def half_of(c2c):
    return oracle.c2c_to_half(c2c)


def completion(half, shape):
    """C2C [z, y, x] spectrum from a half spectrum [kx][kz][ky]."""
    nx, ny, nz = shape
    full = np.zeros((nz, ny, nx), np.complex128)
    full[:, :, : nx // 2 + 1] = np.transpose(half, (1, 2, 0))
    for kx in range(nx // 2 + 1, nx):
        src = np.transpose(half[nx - kx], (0, 1))  # [kz][ky]
        full[:, :, kx] = np.conj(np.roll(np.flip(src, (0, 1)), 1, axis=(0, 1)))
    return full


@pytest.mark.parametrize("shape", SHAPES)
def test_forward_matches_oracle_fp64(shape):
    print("case forward fp64", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(11)
    u = rng.standard_normal((nz, ny, nx))
    fft = sph.HipFft3D(shape, (2 * np.pi,) * 3, precision="fp64")
    got = fft.forward(u.reshape(-1)).view(np.complex128).reshape(fft.half_shape())
    ref = half_of(oracle.fft3d(u, -1))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("shape", SHAPES)
def test_backward_matches_oracle_fp64(shape):
    print("case backward fp64", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(12)
    H = nx // 2 + 1
    half = rng.standard_normal((H, nz, ny)) + 1j * rng.standard_normal((H, nz, ny))
    fft = sph.HipFft3D(shape, (1.0, 2.0, 3.0), precision="fp64")
    got = fft.backward(half.view(np.float64).reshape(-1)).reshape(nz, ny, nx)
    ref = oracle.fft3d(completion(half, shape), 1).real / (nx * ny * nz)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(16, 16, 16), (64, 64, 64), (256, 256, 256)])
def test_forward_fp32_error(shape):
    print("case forward fp32", shape, flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(13)
    u = rng.standard_normal((nz, ny, nx))
    f32 = sph.HipFft3D(shape, (1.0, 1.0, 1.0), precision="fp32")
    f64 = sph.HipFft3D(shape, (1.0, 1.0, 1.0), precision="fp64")
    a = f32.forward(u.reshape(-1).astype(np.float32).astype(np.float64))
    b = f64.forward(u.reshape(-1).astype(np.float32).astype(np.float64))
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel < 2e-6, rel


def test_reference_roundtrip_known_answer():
    # cpp_tests/spectral_ns_tests.cpp:16-32 on the 12^3 grid
    n = 12
    x = (np.arange(n) + 0.5) * 2 * np.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = (np.sin(X) + 0.25 * np.cos(Y)).reshape(-1)
    fft = sph.HipFft3D((n, n, n), (2 * np.pi,) * 3)
    assert np.abs(fft.backward(fft.forward(u)) - u).max() < 1e-10


def test_golden_numpy_vectors_forward():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
    for key in g.files:
        if not key.startswith("in_"):
            continue
        a = g[key].real  # real input; its forward spectrum is fft of the real part
        nz, ny, nx = a.shape
        if not all(sph.load_library().sph_supported_length(v) for v in (nx, ny, nz)):
            continue
        fft = sph.HipFft3D((nx, ny, nz), (1.0, 1.0, 1.0))
        got = fft.forward(a.reshape(-1)).view(np.complex128).reshape(fft.half_shape())
        ref = half_of(np.fft.fftn(a))
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), key


def test_spectral_mode_table():
    fft = sph.HipFft3D((8, 12, 16), (1.0, 1.0, 1.0))
    assert fft.spectral_storage_size() == 2 * 5 * 12 * 16
    assert fft.mx.max() == 4 and fft.my.max() == 6 and fft.my.min() == -5 and fft.mz.max() == 8
    seen = []
    fft.for_each_spectral(lambda i, m: seen.append((i, m.mx, m.my, m.mz)) if len(seen) < 3 else None)
    assert seen[0] == (0, 0, 0, 0) and seen[1] == (2, 0, 1, 0)


def _hip_runtime():
    """The HIP runtime libspectral_hip.so itself links (device memory for the
    device-pointer entry points)."""
    import ctypes
    sph.load_library()
    lib = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


def test_device_pointer_transforms():
    import ctypes
    hip = _hip_runtime()
    H2D, D2H = 1, 2
    n = 32
    fft = sph.HipFft3D((n, n, n), (1.0, 1.0, 1.0), precision="fp64")
    u = np.random.default_rng(5).standard_normal(n ** 3)
    nspec = (n // 2 + 1) * n * n * 2
    bufs = [ctypes.c_void_p() for _ in range(3)]
    for b, sz in zip(bufs, (u.nbytes, nspec * 8, u.nbytes)):
        assert hip.hipMalloc(ctypes.byref(b), sz) == 0
    try:
        assert hip.hipMemcpy(bufs[0], u.ctypes.data, u.nbytes, H2D) == 0
        fft.forward_device(bufs[0].value, bufs[1].value)
        fft.backward_device(bufs[1].value, bufs[2].value)
        spec = np.empty(nspec)
        back = np.empty_like(u)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(spec.ctypes.data, bufs[1], spec.nbytes, D2H) == 0
        assert hip.hipMemcpy(back.ctypes.data, bufs[2], back.nbytes, D2H) == 0
    finally:
        for b in bufs:
            hip.hipFree(b)
    ref = fft.forward(u)
    assert np.abs(spec - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(back - u).max() < 1e-12

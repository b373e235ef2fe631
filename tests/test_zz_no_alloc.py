"""sph_fft_backward_noalloc (no device allocation: its intermediate field and
twiddles live in the library's static scratch array) against
sph_fft_backward (the handle's own scratch) through the C ABI: two calls in
a row on the same device, the second with a smaller input, in FP64 and FP32;
the outputs must be bitwise equal (the same kernels and twiddle values). A
transform whose intermediate does not fit the scratch is rejected before
anything runs.
"""
import ctypes

import numpy as np
import pytest

import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


def hip_runtime():
    sph.load_library()
    lib = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


def backward_both(hip, shape, precision, seed):
    """(noalloc output, handle output) of one random half spectrum."""
    lib = sph.load_library()
    nx, ny, nz = shape
    real = np.float32 if precision == "fp32" else np.float64
    cplx = np.complex64 if precision == "fp32" else np.complex128
    rng = np.random.default_rng(seed)
    H = nx // 2 + 1
    spec = (rng.standard_normal((H, nz, ny)) + 1j * rng.standard_normal((H, nz, ny))).astype(cplx)
    npts = nx * ny * nz
    bufs = [ctypes.c_void_p() for _ in range(3)]
    sizes = (spec.nbytes, npts * np.dtype(real).itemsize, npts * np.dtype(real).itemsize)
    for b, sz in zip(bufs, sizes):
        assert hip.hipMalloc(ctypes.byref(b), sz) == 0
    try:
        assert hip.hipMemcpy(bufs[0], spec.ctypes.data, spec.nbytes, H2D) == 0
        shp = (ctypes.c_int * 3)(*shape)
        prec = sph._precision(precision)
        rc = lib.sph_fft_backward_noalloc(shp, prec, bufs[0], bufs[1], None)
        assert rc == 0, lib.sph_last_error().decode()
        fft = sph.HipFft3D(shape, (1.0, 1.0, 1.0), precision=precision)
        fft.backward_device(bufs[0].value, bufs[2].value)
        assert hip.hipDeviceSynchronize() == 0
        a = np.empty(npts, real)
        b = np.empty(npts, real)
        assert hip.hipMemcpy(a.ctypes.data, bufs[1], a.nbytes, D2H) == 0
        assert hip.hipMemcpy(b.ctypes.data, bufs[2], b.nbytes, D2H) == 0
        fft.close()
    finally:
        for b_ in bufs:
            hip.hipFree(b_)
    return a, b


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_backward_equals_handle_backward(precision):
    hip = hip_runtime()
    for shape in [(128, 96, 64), (32, 24, 16)]:  # the second call smaller
        print("case noalloc backward", precision, shape, flush=True)
        a, b = backward_both(hip, shape, precision, seed=sum(shape))
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b)


def test_noalloc_rejects_too_large_and_invalid_input():
    lib = sph.load_library()
    print("case noalloc errors", flush=True)
    dummy = ctypes.c_void_p(16)  # never dereferenced: the calls fail before any launch
    shp = (ctypes.c_int * 3)(512, 512, 512)
    assert lib.sph_fft_backward_noalloc(shp, sph._precision("fp32"), dummy, dummy, None) == 1
    assert "bytes of scratch" in lib.sph_last_error().decode()
    shp = (ctypes.c_int * 3)(64, 1, 64)
    assert lib.sph_fft_backward_noalloc(shp, sph._precision("fp64"), dummy, dummy, None) == 1
    assert "greater than one" in lib.sph_last_error().decode()
    shp = (ctypes.c_int * 3)(14, 64, 64)
    assert lib.sph_fft_backward_noalloc(shp, sph._precision("fp64"), dummy, dummy, None) == 1
    assert "unsupported transform length" in lib.sph_last_error().decode()

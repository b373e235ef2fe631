// sph_fft_backward_noalloc: the backward 3D transform of sph_fft_backward for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc). It takes no handle:
// the transform's intermediate field and its three twiddle tables live in
// `scratch`, a static device array of fixed size that every GPU the library
// runs on holds until the process ends.
//
// Layout of scratch (bytes): three twiddle slots of kTwiddleSlot bytes (the
// x, y and z tables, kMaxLength complex values of the precision each), then
// the intermediate half-spectrum field, at most kFieldBytes.
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address is taken once per device with hipGetSymbolAddress (on the
// entry point's first call on that device) and only that address is handed
// to kernels.
#include "spectral_hip.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

namespace sph {
namespace {

constexpr int kMaxLength = 1024;
constexpr std::size_t kTwiddleSlot = kMaxLength * sizeof(cx<double>);  // 16 KiB
constexpr std::size_t kFieldOffset = 4 * kTwiddleSlot;                 // 64 KiB
constexpr std::size_t kFieldBytes = std::size_t(80) << 20;             // 80 MiB
constexpr std::size_t kScratchBytes = kFieldOffset + kFieldBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Twiddle entries travel as kernel arguments (no host buffer, no copy from
// pageable memory), kBlock values per launch.
constexpr int kBlock = 128;
template <typename T>
struct TwiddleBlock {
    cx<T> v[kBlock];
    int count;
};

template <typename T>
__global__ void k_put_twiddles(cx<T>* dst, TwiddleBlock<T> b) {
    const int i = threadIdx.x;
    if (i < b.count) dst[i] = b.v[i];
}

// What each device's scratch holds: its address and the precision and shape
// whose twiddles are in it (-1: none yet).
struct DeviceScratch {
    unsigned char* ptr = nullptr;
    int precision = -1;
    int shape[3] = {0, 0, 0};
};
std::mutex g_mu;
std::vector<DeviceScratch> g_devices;

DeviceScratch& device_scratch() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (static_cast<int>(g_devices.size()) <= dev) g_devices.resize(static_cast<std::size_t>(dev) + 1);
    DeviceScratch& d = g_devices[static_cast<std::size_t>(dev)];
    if (!d.ptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned no address");
        d.ptr = static_cast<unsigned char*>(p);
    }
    return d;
}

template <typename T>
void put_twiddles(cx<T>* dst, int n, hipStream_t s) {
    for (int m0 = 0; m0 < n; m0 += kBlock) {
        TwiddleBlock<T> b{};
        b.count = n - m0 < kBlock ? n - m0 : kBlock;
        for (int i = 0; i < b.count; ++i) b.v[i] = twiddle_entry<T>(m0 + i, n);
        k_put_twiddles<T><<<1, kBlock, 0, s>>>(dst + m0, b);
    }
}

// The kernels of Fft3<T>::backward, with the intermediate field and the
// twiddles in scratch.
template <typename T>
void backward(const int shape[3], int precision, const void* spec, void* phys, hipStream_t s) {
    const int nx = shape[0], ny = shape[1], nz = shape[2], H = nx / 2 + 1;
    const PassTable<T>* pt[3];
    for (int d = 0; d < 3; ++d) pt[d] = table_or_throw<T>(shape[d]);
    const std::size_t need = static_cast<std::size_t>(H) * ny * nz * sizeof(cx<T>);
    if (need > kFieldBytes)
        throw std::runtime_error("sph_fft_backward_noalloc: the transform needs " + std::to_string(need) +
                                 " bytes of scratch, the library holds " + std::to_string(kFieldBytes));
    std::lock_guard<std::mutex> lock(g_mu);
    DeviceScratch& d = device_scratch();
    cx<T>* tw[3];
    for (int a = 0; a < 3; ++a) tw[a] = reinterpret_cast<cx<T>*>(d.ptr + a * kTwiddleSlot);
    cx<T>* work = reinterpret_cast<cx<T>*>(d.ptr + kFieldOffset);
    if (d.precision != precision || d.shape[0] != nx || d.shape[1] != ny || d.shape[2] != nz) {
        for (int a = 0; a < 3; ++a) put_twiddles<T>(tw[a], shape[a], s);
        d.precision = precision;
        for (int a = 0; a < 3; ++a) d.shape[a] = shape[a];
    }
    InvZArgs<T> iz{};
    iz.in[0] = static_cast<const cx<T>*>(spec);
    iz.out[0] = work;
    iz.nfields = 1;
    iz.ny = ny;
    iz.tw = tw[2];
    pt[2]->inv_z(iz, ny, H, s);
    pt[1]->lines(work, static_cast<long long>(H) * nz, true, tw[1], s);
    InvXArgs<T> ix{};
    ix.a[0] = work;
    ix.b[0] = nullptr;
    ix.oa[0] = static_cast<T*>(phys);
    ix.ob[0] = nullptr;
    ix.npairs = 1;
    ix.ny = ny;
    ix.nz = nz;
    ix.scale = T(1) / static_cast<T>(static_cast<std::int64_t>(nx) * ny * nz);
    ix.tw = tw[0];
    pt[0]->inv_x(ix, ny, nz, s);
    hip_check(hipGetLastError(), "sph_fft_backward_noalloc launch");
}

}  // namespace
}  // namespace sph

extern "C" int sph_fft_backward_noalloc(const int shape[3], int precision, const void* spectral, void* physical,
                                        void* stream) {
    try {
        if (!shape || !spectral || !physical) throw std::runtime_error("null pointer argument");
        if (precision != SPH_FP32 && precision != SPH_FP64)
            throw std::runtime_error("precision must be SPH_FP32 or SPH_FP64");
        for (int d = 0; d < 3; ++d)
            if (shape[d] <= 1) throw std::runtime_error("all grid sizes must be greater than one");
        sph::require_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (precision == SPH_FP32)
            sph::backward<float>(shape, precision, spectral, physical, s);
        else
            sph::backward<double>(shape, precision, spectral, physical, s);
        sph::g_last_error.clear();
        return 0;
    } catch (const std::exception& e) {
        sph::g_last_error = e.what();
        return 1;
    }
}

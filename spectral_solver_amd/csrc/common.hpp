// Host-side helpers shared by the FFT handle, the solvers and the C ABI:
// error checks, RAII device memory and streams, twiddle and phase tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pass_table.hpp"

namespace sph {

constexpr double kPi = 3.141592653589793238462643383279502884;

// Text of the last failure on this thread (returned by sph_last_error).
inline thread_local std::string g_last_error;

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void require_device() {
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) throw std::runtime_error("no HIP device available");
}

template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) { resize(n); }
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void resize(std::size_t n) {
        release();
        if (n > 0) hip_check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    std::size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// Pinned (page-locked) host memory from hipHostMalloc, for asynchronous
// device-to-host copies.
template <typename T>
class PinnedHost {
public:
    explicit PinnedHost(std::size_t n) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    }
    ~PinnedHost() {
        if (p_) (void)hipHostFree(p_);
    }
    PinnedHost(const PinnedHost&) = delete;
    PinnedHost& operator=(const PinnedHost&) = delete;
    T* get() const { return p_; }

private:
    T* p_ = nullptr;
};

// An event without timing, for host waits on one point of a stream.
class Event {
public:
    Event() { hip_check(hipEventCreateWithFlags(&e_, hipEventDisableTiming), "hipEventCreate"); }
    ~Event() {
        if (e_) (void)hipEventDestroy(e_);
    }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    hipEvent_t get() const { return e_; }

private:
    hipEvent_t e_ = nullptr;
};

// A non-blocking stream at the default priority.
class Stream {
public:
    Stream() { hip_check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate"); }
    ~Stream() {
        if (s_) (void)hipStreamDestroy(s_);
    }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    hipStream_t get() const { return s_; }

private:
    hipStream_t s_ = nullptr;
};

// e^{-2 pi i m / n} from long double, rounded once.
template <typename T>
cx<T> twiddle_entry(int m, int n) {
    const long double ang = -2.0L * static_cast<long double>(kPi) * m / n;
    return {static_cast<T>(std::cos(ang)), static_cast<T>(std::sin(ang))};
}

// twiddle_entry(m, n) for m in [0, n)
template <typename T>
std::vector<cx<T>> twiddle_table(int n) {
    std::vector<cx<T>> tw(n);
    for (int m = 0; m < n; ++m) tw[m] = twiddle_entry<T>(m, n);
    return tw;
}

// entries of the half-cell phase table: s = mx+my+mz in [-n, 3n/2]
inline int phase_table_size(int n) { return 5 * n / 2 + 1; }

// Half-cell phase e^{i pi s / n} stored at s + n, evaluated as the reference
// does: angle = (double(s) * pi) / double(n) in double, cast to T
// (src/ava_taylor_green_impl.hpp:779-784).
template <typename T>
std::vector<cx<T>> phase_table(int n) {
    std::vector<cx<T>> ph(phase_table_size(n));
    for (int s = -n; s <= 3 * n / 2; ++s) {
        const double angle = (static_cast<double>(s) * kPi) / static_cast<double>(n);
        ph[s + n] = {static_cast<T>(std::cos(angle)), static_cast<T>(std::sin(angle))};
    }
    return ph;
}

template <typename T>
const PassTable<T>* table_or_throw(int n) {
    const PassTable<T>* t = pass_table<T>(n);
    if (!t) throw std::runtime_error("unsupported transform length " + std::to_string(n));
    return t;
}

template <typename T>
void upload(DeviceArray<T>& dst, const std::vector<T>& src) {
    dst.resize(src.size());
    hip_check(hipMemcpy(dst.get(), src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

}  // namespace sph

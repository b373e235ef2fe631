// Host side of the MI355X spectral library: FFT backend handle, Taylor-Green
// RK3 solver, driver loop and the extern "C" boundary declared in
// include/spectral_hip.h.
#include "spectral_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pass_table.hpp"

namespace sph {
namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

thread_local std::string g_last_error;

// entries of the half-cell phase table: s = mx+my+mz in [-n, 3n/2]
int phase_table_size(int n) { return 5 * n / 2 + 1; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void require_device() {
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) throw std::runtime_error("no HIP device available");
}

// RAII device allocation
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

// e^{-2 pi i m / n}, m in [0, n), from long double and rounded once.
template <typename T>
std::vector<cx<T>> twiddle_table(int n) {
    std::vector<cx<T>> tw(n);
    for (int m = 0; m < n; ++m) {
        const long double ang = -2.0L * static_cast<long double>(kPi) * m / n;
        tw[m] = {static_cast<T>(std::cos(ang)), static_cast<T>(std::sin(ang))};
    }
    return tw;
}

// Half-cell phase e^{i pi s / n} for s = mx+my+mz in [-n, 3n/2], stored at
// s + n, evaluated as the reference does: angle = (double(s) * pi) / double(n)
// in double, cast to T (src/ava_taylor_green_impl.hpp:779-784).
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

__global__ void k_set_coef_f(float* c, float a0, float b0, float a1, float b1, float a2, float b2) {
    c[0] = a0;
    c[1] = b0;
    c[2] = a1;
    c[3] = b1;
    c[4] = a2;
    c[5] = b2;
}
__global__ void k_set_coef_d(double* c, double a0, double b0, double a1, double b1, double a2, double b2) {
    c[0] = a0;
    c[1] = b0;
    c[2] = a1;
    c[3] = b1;
    c[4] = a2;
    c[5] = b2;
}
void set_coef(float* c, const float* v, hipStream_t s) {
    k_set_coef_f<<<1, 1, 0, s>>>(c, v[0], v[1], v[2], v[3], v[4], v[5]);
}
void set_coef(double* c, const double* v, hipStream_t s) {
    k_set_coef_d<<<1, 1, 0, s>>>(c, v[0], v[1], v[2], v[3], v[4], v[5]);
}

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

}  // namespace

// ---------------------------------------------------------------------------
// FFT backend handle (FftwFft3D-shaped): real [nz][ny][nx] <-> half spectrum
// ---------------------------------------------------------------------------
class FftBase {
public:
    virtual ~FftBase() = default;
    int shape[3] = {0, 0, 0};
    double lengths[3] = {0, 0, 0};
    int precision = SPH_FP64;
    std::int64_t physical_size() const { return static_cast<std::int64_t>(shape[0]) * shape[1] * shape[2]; }
    std::int64_t spectral_modes() const {
        return static_cast<std::int64_t>(shape[0] / 2 + 1) * shape[1] * shape[2];
    }
    virtual void forward(const void* phys, void* spec, hipStream_t s) = 0;
    virtual void backward(const void* spec, void* phys, hipStream_t s) = 0;
    virtual void forward_host(const double* phys, double* spec) = 0;
    virtual void backward_host(const double* spec, double* phys) = 0;
};

template <typename T>
class Fft3 final : public FftBase {
public:
    Fft3(const int shp[3], const double len[3]) {
        for (int d = 0; d < 3; ++d) {
            shape[d] = shp[d];
            lengths[d] = len[d];
        }
        precision = sizeof(T) == 4 ? SPH_FP32 : SPH_FP64;
        for (int d = 0; d < 3; ++d) {
            pt_[d] = table_or_throw<T>(shape[d]);
            const auto tw = twiddle_table<T>(shape[d]);
            tw_[d].resize(tw.size());
            hip_check(hipMemcpy(tw_[d].get(), tw.data(), tw.size() * sizeof(cx<T>), hipMemcpyHostToDevice),
                      "hipMemcpy twiddles");
        }
        scratch_.resize(static_cast<std::size_t>(spectral_modes()));
    }

    void forward(const void* phys, void* spec, hipStream_t s) override {
        if (!s) s = stream_.get();
        const int nx = shape[0], ny = shape[1], nz = shape[2], H = nx / 2 + 1;
        auto* out = static_cast<cx<T>*>(spec);
        FwdXArgs<T> fx{};
        fx.in[0] = static_cast<const T*>(phys);
        fx.out[0] = out;
        fx.nin = 1;
        fx.ny = ny;
        fx.nz = nz;
        fx.tw = tw_[0].get();
        pt_[0]->fwd_x(fx, FX_PLAIN, ny, nz, s);
        pt_[1]->lines(out, static_cast<long long>(H) * nz, false, tw_[1].get(), s);
        FwdZArgs<T> fz{};
        fz.in[0] = out;
        fz.out[0] = out;
        fz.nfields = 1;
        fz.ny = ny;
        fz.n = nz;
        fz.tw = tw_[2].get();
        pt_[2]->fwd_z(fz, FZ_PLAIN, ny, H, s);
        hip_check(hipGetLastError(), "sph_fft_forward launch");
    }

    void backward(const void* spec, void* phys, hipStream_t s) override {
        if (!s) s = stream_.get();
        const int nx = shape[0], ny = shape[1], nz = shape[2], H = nx / 2 + 1;
        InvZArgs<T> iz{};
        iz.in[0] = static_cast<const cx<T>*>(spec);
        iz.out[0] = scratch_.get();
        iz.nfields = 1;
        iz.ny = ny;
        iz.n = nz;
        iz.tw = tw_[2].get();
        pt_[2]->inv_z(iz, IZ_PLAIN, ny, H, s);
        pt_[1]->lines(scratch_.get(), static_cast<long long>(H) * nz, true, tw_[1].get(), s);
        InvXArgs<T> ix{};
        ix.a[0] = scratch_.get();
        ix.b[0] = nullptr;
        ix.oa[0] = static_cast<T*>(phys);
        ix.ob[0] = nullptr;
        ix.npairs = 1;
        ix.ny = ny;
        ix.nz = nz;
        ix.scale = T(1) / static_cast<T>(physical_size());
        ix.tw = tw_[0].get();
        pt_[0]->inv_x(ix, ny, nz, s);
        hip_check(hipGetLastError(), "sph_fft_backward launch");
    }

    void forward_host(const double* phys, double* spec) override {
        ensure_host_buffers();
        const std::size_t np = static_cast<std::size_t>(physical_size());
        const std::size_t nm = static_cast<std::size_t>(spectral_modes());
        std::vector<T> tmp(np);
        for (std::size_t i = 0; i < np; ++i) tmp[i] = static_cast<T>(phys[i]);
        hip_check(hipMemcpyAsync(dphys_.get(), tmp.data(), np * sizeof(T), hipMemcpyHostToDevice, stream_.get()),
                  "hipMemcpy");
        forward(dphys_.get(), dspec_.get(), stream_.get());
        std::vector<cx<T>> out(nm);
        hip_check(hipMemcpyAsync(out.data(), dspec_.get(), nm * sizeof(cx<T>), hipMemcpyDeviceToHost,
                                 stream_.get()),
                  "hipMemcpy");
        hip_check(hipStreamSynchronize(stream_.get()), "hipStreamSynchronize");
        for (std::size_t i = 0; i < nm; ++i) {
            spec[2 * i] = out[i].x;
            spec[2 * i + 1] = out[i].y;
        }
    }

    void backward_host(const double* spec, double* phys) override {
        ensure_host_buffers();
        const std::size_t np = static_cast<std::size_t>(physical_size());
        const std::size_t nm = static_cast<std::size_t>(spectral_modes());
        std::vector<cx<T>> in(nm);
        for (std::size_t i = 0; i < nm; ++i) in[i] = {static_cast<T>(spec[2 * i]), static_cast<T>(spec[2 * i + 1])};
        hip_check(hipMemcpyAsync(dspec_.get(), in.data(), nm * sizeof(cx<T>), hipMemcpyHostToDevice, stream_.get()),
                  "hipMemcpy");
        backward(dspec_.get(), dphys_.get(), stream_.get());
        std::vector<T> out(np);
        hip_check(hipMemcpyAsync(out.data(), dphys_.get(), np * sizeof(T), hipMemcpyDeviceToHost, stream_.get()),
                  "hipMemcpy");
        hip_check(hipStreamSynchronize(stream_.get()), "hipStreamSynchronize");
        for (std::size_t i = 0; i < np; ++i) phys[i] = out[i];
    }

private:
    void ensure_host_buffers() {
        if (!dphys_.get()) dphys_.resize(static_cast<std::size_t>(physical_size()));
        if (!dspec_.get()) dspec_.resize(static_cast<std::size_t>(spectral_modes()));
    }

    Stream stream_;
    const PassTable<T>* pt_[3] = {nullptr, nullptr, nullptr};
    DeviceArray<cx<T>> tw_[3];
    DeviceArray<cx<T>> scratch_;
    DeviceArray<T> dphys_;
    DeviceArray<cx<T>> dspec_;
};

// ---------------------------------------------------------------------------
// Taylor-Green solver
// ---------------------------------------------------------------------------
class SolverBase {
public:
    virtual ~SolverBase() = default;
    virtual void step(double dt) = 0;
    virtual sph_diagnostics diagnostics(double t) = 0;
    virtual double max_velocity() = 0;
    virtual void get_uhat(int c, double* host) = 0;
    virtual void set_uhat(int c, const double* host) = 0;
    virtual void profile(sph_profile* out) = 0;
    virtual void set_graph(bool on) = 0;
    virtual void synchronize() = 0;
    virtual void time_passes(int reps, double* ms, double* bytes) = 0;
    int n = 0;
    std::int64_t modes = 0;
};

template <typename T>
class Solver final : public SolverBase {
public:
    Solver(int n_in, double reynolds) {
        if (n_in <= 0) throw std::runtime_error("grid size must be positive");
        if (reynolds <= 0.0) throw std::runtime_error("Reynolds number must be positive");
        require_device();
        n = n_in;
        nu_ = static_cast<T>(1.0 / reynolds);
        pt_ = table_or_throw<T>(n);
        H_ = n / 2 + 1;
        modes = static_cast<std::int64_t>(H_) * n * n;
        npts_ = static_cast<std::int64_t>(n) * n * n;
        U_.resize(3 * modes);
        G_.resize(3 * modes);
        W_.resize(12 * modes);
        P_.resize(6 * npts_);
        coef_.resize(8);
        partial_.resize(3 * kReduceBlocks + 8);
        result_.resize(8);
        const auto tw = twiddle_table<T>(n);
        tw_.resize(tw.size());
        hip_check(hipMemcpy(tw_.get(), tw.data(), tw.size() * sizeof(cx<T>), hipMemcpyHostToDevice), "hipMemcpy");
        const auto ph = phase_table<T>(n);
        phase_.resize(ph.size());
        hip_check(hipMemcpy(phase_.get(), ph.data(), ph.size() * sizeof(cx<T>), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemset(G_.get(), 0, 3 * modes * sizeof(cx<T>)), "hipMemset");
        initial_condition();
        hip_check(hipStreamSynchronize(stream_.get()), "TaylorGreenSolver initialisation");
    }

    ~Solver() override {
        (void)hipStreamSynchronize(stream_.get());
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        for (auto& p : pending_) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }

    void step(double dt) override {
        const hipStream_t s = stream_.get();
        if (!(dt == coef_dt_)) {
            const T rdt = static_cast<T>(dt);
            // impl.hpp:998, 1011, 1014, 1027, 1030: coefficients in Real as written
            const T c[6] = {T(0), rdt / T(3), -T(5) / T(9), T(15) * rdt / T(16), -T(153) / T(128), T(8) * rdt / T(15)};
            set_coef(coef_.get(), c, s);
            coef_dt_ = dt;
        }
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
        if (use_graph_) {
            if (!exec_) capture();
            hip_check(hipGraphLaunch(exec_, s), "hipGraphLaunch");
        } else {
            for (int st = 0; st < 3; ++st) launch_stage(st);
            hip_check(hipGetLastError(), "TaylorGreenSolver::step");
        }
        hip_check(hipEventRecord(e1, s), "hipEventRecord");
        pending_.emplace_back(e0, e1);
        if (pending_.size() > 64) resolve_one();
        prof_.step_calls += 1;
        prof_.rhs_calls += 3;
        prof_.forward_calls += 36;
        prof_.backward_calls += 18;
    }

    sph_diagnostics diagnostics(double t) override {
        const auto t0 = std::chrono::steady_clock::now();
        const hipStream_t s = stream_.get();
        inverse(PZ_DIAG);
        k_reduce_phys<T><<<kReduceBlocks, kReduceThreads, 0, s>>>(P_.get(), P_.get() + npts_, P_.get() + 2 * npts_,
                                                                   P_.get() + 3 * npts_, P_.get() + 4 * npts_,
                                                                   P_.get() + 5 * npts_, npts_, partial_.get());
        k_reduce_final<<<1, kReduceThreads, 0, s>>>(partial_.get(), kReduceBlocks, 3, 0b010, result_.get());
        k_reduce_div3<T><<<kReduceBlocks, kReduceThreads, 0, s>>>(U_.get(), n, partial_.get());
        k_reduce_final<<<1, kReduceThreads, 0, s>>>(partial_.get(), kReduceBlocks, 1, 0b1, result_.get() + 3);
        double r[4];
        hip_check(hipMemcpyAsync(r, result_.get(), sizeof(r), hipMemcpyDeviceToHost, s), "hipMemcpy diagnostics");
        hip_check(hipStreamSynchronize(s), "TaylorGreenSolver::diagnostics");
        const double inv = 1.0 / static_cast<double>(npts_);
        sph_diagnostics d{};
        d.t = t;
        d.kinetic_energy = r[0] * inv;
        d.enstrophy = r[2] * inv;
        d.dissipation = 2.0 * static_cast<double>(nu_) * d.enstrophy;
        d.max_spectral_divergence = r[3];
        d.max_velocity = std::sqrt(r[1]);
        prof_.diagnostics += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        prof_.diagnostics_calls += 1;
        prof_.backward_calls += 6;
        return d;
    }

    double max_velocity() override {
        const auto t0 = std::chrono::steady_clock::now();
        const hipStream_t s = stream_.get();
        inverse(PZ_VEL);
        k_reduce_phys<T><<<kReduceBlocks, kReduceThreads, 0, s>>>(P_.get(), P_.get() + npts_, P_.get() + 2 * npts_,
                                                                   nullptr, nullptr, nullptr, npts_, partial_.get());
        k_reduce_final<<<1, kReduceThreads, 0, s>>>(partial_.get(), kReduceBlocks, 3, 0b010, result_.get());
        double r[3];
        hip_check(hipMemcpyAsync(r, result_.get(), sizeof(r), hipMemcpyDeviceToHost, s), "hipMemcpy max_velocity");
        hip_check(hipStreamSynchronize(s), "TaylorGreenSolver::max_velocity");
        prof_.max_velocity += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        prof_.max_velocity_calls += 1;
        prof_.backward_calls += 3;
        return std::sqrt(r[1]);
    }

    void get_uhat(int c, double* host) override {
        if (c < 0 || c > 2) throw std::runtime_error("component out of range");
        std::vector<cx<T>> tmp(static_cast<std::size_t>(modes));
        // component c of the component-interleaved state: a strided 2D copy
        hip_check(hipMemcpy2DAsync(tmp.data(), sizeof(cx<T>), U_.get() + c, 3 * sizeof(cx<T>), sizeof(cx<T>), modes,
                                   hipMemcpyDeviceToHost, stream_.get()),
                  "hipMemcpy get_uhat");
        hip_check(hipStreamSynchronize(stream_.get()), "get_uhat");
        for (std::int64_t i = 0; i < modes; ++i) {
            host[2 * i] = tmp[i].x;
            host[2 * i + 1] = tmp[i].y;
        }
    }

    void set_uhat(int c, const double* host) override {
        if (c < 0 || c > 2) throw std::runtime_error("component out of range");
        std::vector<cx<T>> tmp(static_cast<std::size_t>(modes));
        for (std::int64_t i = 0; i < modes; ++i) tmp[i] = {static_cast<T>(host[2 * i]), static_cast<T>(host[2 * i + 1])};
        hip_check(hipMemcpy2DAsync(U_.get() + c, 3 * sizeof(cx<T>), tmp.data(), sizeof(cx<T>), sizeof(cx<T>), modes,
                                   hipMemcpyHostToDevice, stream_.get()),
                  "hipMemcpy set_uhat");
        hip_check(hipStreamSynchronize(stream_.get()), "set_uhat");
    }

    void profile(sph_profile* out) override {
        while (!pending_.empty()) resolve_one();
        *out = prof_;
    }

    void set_graph(bool on) override { use_graph_ = on; }

    void synchronize() override { hip_check(hipStreamSynchronize(stream_.get()), "hipStreamSynchronize"); }

    // Mean duration of each pass kernel over `reps` launches; the RK pass
    // updates scratch copies of the state so the state is unchanged.
    void time_passes(int reps, double* ms, double* bytes) override {
        if (reps < 1) reps = 1;
        const hipStream_t s = stream_.get();
        synchronize();
        DeviceArray<cx<T>> u2(3 * modes), g2(3 * modes);
        hip_check(hipMemcpyAsync(u2.get(), U_.get(), 3 * modes * sizeof(cx<T>), hipMemcpyDeviceToDevice, s), "copy");
        hip_check(hipMemcpyAsync(g2.get(), G_.get(), 3 * modes * sizeof(cx<T>), hipMemcpyDeviceToDevice, s), "copy");
        // fill W and P with a genuine stage's intermediates first
        if (!(coef_dt_ == coef_dt_)) {
            const T c[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
            set_coef(coef_.get(), c, s);
        }
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        for (int k = 0; k < 6; ++k) {
            // warm-up + make inputs of pass k valid by running the passes before it
            for (int j = 0; j <= k; ++j) launch_pass(j, 1, u2.get(), g2.get());
            hip_check(hipEventRecord(a, s), "hipEventRecord");
            for (int r = 0; r < reps; ++r) launch_pass(k, 1, u2.get(), g2.get());
            hip_check(hipEventRecord(b, s), "hipEventRecord");
            hip_check(hipEventSynchronize(b), "hipEventSynchronize");
            float e = 0.f;
            hip_check(hipEventElapsedTime(&e, a, b), "hipEventElapsedTime");
            ms[k] = static_cast<double>(e) / reps;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        hip_check(hipGetLastError(), "time_passes");
        const double C = static_cast<double>(modes) * 2.0 * sizeof(T);
        const double R = static_cast<double>(npts_) * sizeof(T);
        const double b6[6] = {9 * C, 12 * C, 6 * C + 6 * R, 6 * R + 12 * C, 24 * C, 24 * C};
        for (int k = 0; k < 6; ++k) bytes[k] = b6[k];
        synchronize();
    }

private:
    void resolve_one() {
        auto p = pending_.front();
        pending_.pop_front();
        hip_check(hipEventSynchronize(p.second), "hipEventSynchronize");
        float e = 0.f;
        hip_check(hipEventElapsedTime(&e, p.first, p.second), "hipEventElapsedTime");
        prof_.step += 1e-3 * e;
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }

    void capture() {
        const hipStream_t s = stream_.get();
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        for (int st = 0; st < 3; ++st) launch_stage(st);
        hip_check(hipStreamEndCapture(s, &graph_), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    }

    // pass k of stage `stage` (0 inv_z, 1 inv_y, 2 inv_x, 3 fwd_x, 4 fwd_y, 5 fwd_z+update)
    // on the pair-interleaved work buffer W (pair field p at W + 2*p*modes) and
    // the component-interleaved state u, g
    void launch_pass(int k, int stage, cx<T>* u, cx<T>* g) {
        const hipStream_t s = stream_.get();
        const std::int64_t M2 = 2 * modes;
        switch (k) {
            case 0: {
                PairInvZArgs<T> a{};
                a.u3 = u;
                for (int c = 0; c < 3; ++c) a.out[c] = W_.get() + c * M2;
                a.ny = n;
                a.n = n;
                a.phase = phase_.get();
                a.tw = tw_.get();
                pt_->pinv_z(a, PZ_RHS, n, H_, s);
                break;
            }
            case 1:
                pt_->plines(W_.get(), 3LL * H_ * n, true, tw_.get(), s);
                break;
            case 2: {
                PairInvXArgs<T> a{};
                for (int c = 0; c < 3; ++c) {
                    a.ab[c] = W_.get() + c * M2;
                    a.oa[c] = P_.get() + c * npts_;
                    a.ob[c] = P_.get() + (3 + c) * npts_;
                }
                a.npairs = 3;
                a.ny = n;
                a.nz = n;
                a.scale = T(1) / static_cast<T>(npts_);
                a.tw = tw_.get();
                pt_->pinv_x(a, n, n, s);
                break;
            }
            case 3: {
                PairFwdXArgs<T> a{};
                for (int f = 0; f < 6; ++f) a.in[f] = P_.get() + f * npts_;
                for (int q = 0; q < 6; ++q) a.out[q] = W_.get() + q * M2;
                a.ny = n;
                a.nz = n;
                a.tw = tw_.get();
                pt_->pfwd_x(a, n, n, s);
                break;
            }
            case 4:
                pt_->plines(W_.get(), 6LL * H_ * n, false, tw_.get(), s);
                break;
            default: {
                PairFwdZArgs<T> a{};
                for (int q = 0; q < 6; ++q) a.in[q] = W_.get() + q * M2;
                a.u3 = u;
                a.g3 = g;
                a.ny = n;
                a.n = n;
                a.nu = nu_;
                a.coef = coef_.get();
                a.stage = stage;
                a.phase = phase_.get();
                a.tw = tw_.get();
                pt_->pfwd_z_rk(a, n, H_, s);
                break;
            }
        }
    }

    void launch_stage(int stage) {
        for (int k = 0; k < 6; ++k) launch_pass(k, stage, U_.get(), G_.get());
    }

    // u_hat -> physical: PZ_DIAG gives u in P[0..2] and omega in P[3..5];
    // PZ_VEL gives u in P[0..2] only.
    void inverse(int mode) {
        const hipStream_t s = stream_.get();
        const std::int64_t M2 = 2 * modes;
        PairInvZArgs<T> a{};
        a.u3 = U_.get();
        for (int c = 0; c < 3; ++c) a.out[c] = W_.get() + c * M2;
        a.ny = n;
        a.n = n;
        a.phase = phase_.get();
        a.tw = tw_.get();
        PairInvXArgs<T> x{};
        x.ny = n;
        x.nz = n;
        x.scale = T(1) / static_cast<T>(npts_);
        x.tw = tw_.get();
        if (mode == PZ_DIAG) {
            pt_->pinv_z(a, PZ_DIAG, n, H_, s);
            pt_->plines(W_.get(), 3LL * H_ * n, true, tw_.get(), s);
            for (int c = 0; c < 3; ++c) {
                x.ab[c] = W_.get() + c * M2;
                x.oa[c] = P_.get() + c * npts_;
                x.ob[c] = P_.get() + (3 + c) * npts_;
            }
            x.npairs = 3;
        } else {
            pt_->pinv_z(a, PZ_VEL, n, H_, s);
            pt_->plines(W_.get(), 2LL * H_ * n, true, tw_.get(), s);
            x.ab[0] = W_.get();
            x.oa[0] = P_.get();
            x.ob[0] = P_.get() + npts_;
            x.ab[1] = W_.get() + M2;
            x.oa[1] = P_.get() + 2 * npts_;
            x.ob[1] = nullptr;
            x.npairs = 2;
        }
        pt_->pinv_x(x, n, n, s);
        hip_check(hipGetLastError(), "inverse transform launch");
    }

    // impl.hpp:796-817: TG field at cell centres, 3 forward transforms, project_truncate
    void initial_condition() {
        const hipStream_t s = stream_.get();
        const std::int64_t M = modes;
        k_taylor_green<T><<<2048, 256, 0, s>>>(P_.get(), P_.get() + npts_, P_.get() + 2 * npts_, n);
        FwdXArgs<T> fx{};
        for (int c = 0; c < 3; ++c) {
            fx.in[c] = P_.get() + c * npts_;
            fx.out[c] = W_.get() + c * M;
        }
        fx.nin = 3;
        fx.ny = n;
        fx.nz = n;
        fx.tw = tw_.get();
        pt_->fwd_x(fx, FX_PLAIN, n, n, s);
        pt_->lines(W_.get(), 3LL * H_ * n, false, tw_.get(), s);
        FwdZArgs<T> fz{};
        for (int c = 0; c < 3; ++c) fz.in[c] = W_.get() + c * M;
        fz.u3 = U_.get();
        fz.ny = n;
        fz.n = n;
        fz.tw = tw_.get();
        pt_->fwd_z(fz, FZ_PROJECT, n, H_, s);
        hip_check(hipGetLastError(), "initial condition launch");
        prof_.forward_calls += 3;
    }

    Stream stream_;
    const PassTable<T>* pt_ = nullptr;
    T nu_ = T(0);
    int H_ = 0;
    std::int64_t npts_ = 0;
    DeviceArray<cx<T>> U_, G_, W_, tw_, phase_;
    DeviceArray<T> P_, coef_;
    DeviceArray<double> partial_, result_;
    double coef_dt_ = std::nan("");
    bool use_graph_ = true;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    std::deque<std::pair<hipEvent_t, hipEvent_t>> pending_;
    sph_profile prof_{};
};

}  // namespace sph

// ---------------------------------------------------------------------------
// extern "C" boundary
// ---------------------------------------------------------------------------
struct sph_fft_s {
    std::unique_ptr<sph::FftBase> impl;
};
struct sph_solver_s {
    std::unique_ptr<sph::SolverBase> impl;
    double init_seconds = 0.0;
};

namespace {
template <class F>
int guarded(F&& f) {
    try {
        f();
        sph::g_last_error.clear();
        return 0;
    } catch (const std::exception& e) {
        sph::g_last_error = e.what();
        return 1;
    } catch (...) {
        sph::g_last_error = "unknown error";
        return 1;
    }
}
void require_handle(const void* h) {
    if (!h) throw std::runtime_error("null handle");
}
void require_precision(int p) {
    if (p != SPH_FP32 && p != SPH_FP64) throw std::runtime_error("precision must be SPH_FP32 or SPH_FP64");
}
}  // namespace

extern "C" {

const char* sph_last_error(void) { return sph::g_last_error.c_str(); }

int sph_device_count(int* count) {
    return guarded([&] {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *count = c;
    });
}

int sph_supported_length(int n) { return sph::pass_table<double>(n) != nullptr ? 1 : 0; }

int sph_fft_create(const int shape[3], const double lengths[3], int precision, sph_fft_t* out) {
    return guarded([&] {
        require_precision(precision);
        for (int d = 0; d < 3; ++d)
            if (shape[d] <= 1) throw std::runtime_error("all grid sizes must be greater than one");
        for (int d = 0; d < 3; ++d)
            if (!(lengths[d] > 0.0)) throw std::runtime_error("all domain lengths must be positive");
        for (int d = 0; d < 3; ++d)
            if (!sph_supported_length(shape[d]))
                throw std::runtime_error("unsupported transform length " + std::to_string(shape[d]));
        sph::require_device();
        auto h = std::make_unique<sph_fft_s>();
        if (precision == SPH_FP32)
            h->impl = std::make_unique<sph::Fft3<float>>(shape, lengths);
        else
            h->impl = std::make_unique<sph::Fft3<double>>(shape, lengths);
        *out = h.release();
    });
}

int sph_fft_destroy(sph_fft_t fft) {
    return guarded([&] { delete fft; });
}

int sph_fft_sizes(sph_fft_t fft, int64_t* physical_size, int64_t* spectral_modes) {
    return guarded([&] {
        require_handle(fft);
        *physical_size = fft->impl->physical_size();
        *spectral_modes = fft->impl->spectral_modes();
    });
}

int sph_fft_spectral_modes(sph_fft_t fft, int* mx, int* my, int* mz) {
    return guarded([&] {
        require_handle(fft);
        const int nx = fft->impl->shape[0], ny = fft->impl->shape[1], nz = fft->impl->shape[2];
        std::int64_t id = 0;
        for (int kx = 0; kx <= nx / 2; ++kx)
            for (int kz = 0; kz < nz; ++kz)
                for (int ky = 0; ky < ny; ++ky, ++id) {
                    mx[id] = kx;
                    my[id] = ky <= ny / 2 ? ky : ky - ny;
                    mz[id] = kz <= nz / 2 ? kz : kz - nz;
                }
    });
}

int sph_fft_forward(sph_fft_t fft, const void* physical, void* spectral, void* stream) {
    return guarded([&] {
        require_handle(fft);
        fft->impl->forward(physical, spectral, static_cast<hipStream_t>(stream));
    });
}

int sph_fft_backward(sph_fft_t fft, const void* spectral, void* physical, void* stream) {
    return guarded([&] {
        require_handle(fft);
        fft->impl->backward(spectral, physical, static_cast<hipStream_t>(stream));
    });
}

int sph_fft_forward_host(sph_fft_t fft, const double* physical, double* spectral) {
    return guarded([&] {
        require_handle(fft);
        fft->impl->forward_host(physical, spectral);
    });
}

int sph_fft_backward_host(sph_fft_t fft, const double* spectral, double* physical) {
    return guarded([&] {
        require_handle(fft);
        fft->impl->backward_host(spectral, physical);
    });
}

int sph_solver_create(int n, double reynolds, int precision, sph_solver_t* out) {
    return guarded([&] {
        require_precision(precision);
        if (n <= 0) throw std::runtime_error("grid size must be positive");
        if (reynolds <= 0.0) throw std::runtime_error("Reynolds number must be positive");
        if (!sph_supported_length(n)) throw std::runtime_error("unsupported transform length " + std::to_string(n));
        sph::require_device();
        const auto t0 = std::chrono::steady_clock::now();
        auto h = std::make_unique<sph_solver_s>();
        if (precision == SPH_FP32)
            h->impl = std::make_unique<sph::Solver<float>>(n, reynolds);
        else
            h->impl = std::make_unique<sph::Solver<double>>(n, reynolds);
        h->init_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = h.release();
    });
}

int sph_solver_destroy(sph_solver_t solver) {
    return guarded([&] { delete solver; });
}

int sph_solver_step_rk3(sph_solver_t solver, double dt) {
    return guarded([&] {
        require_handle(solver);
        solver->impl->step(dt);
    });
}

int sph_solver_diagnostics(sph_solver_t solver, double t, sph_diagnostics* out) {
    return guarded([&] {
        require_handle(solver);
        *out = solver->impl->diagnostics(t);
    });
}

int sph_solver_max_velocity(sph_solver_t solver, double* out) {
    return guarded([&] {
        require_handle(solver);
        *out = solver->impl->max_velocity();
    });
}

int sph_solver_sizes(sph_solver_t solver, int* n, int64_t* spectral_modes) {
    return guarded([&] {
        require_handle(solver);
        *n = solver->impl->n;
        *spectral_modes = solver->impl->modes;
    });
}

int sph_solver_get_uhat(sph_solver_t solver, int component, double* host) {
    return guarded([&] {
        require_handle(solver);
        solver->impl->get_uhat(component, host);
    });
}

int sph_solver_set_uhat(sph_solver_t solver, int component, const double* host) {
    return guarded([&] {
        require_handle(solver);
        solver->impl->set_uhat(component, host);
    });
}

int sph_solver_profile(sph_solver_t solver, sph_profile* out) {
    return guarded([&] {
        require_handle(solver);
        solver->impl->profile(out);
    });
}

int sph_solver_print_profile(sph_solver_t solver, int rank, double init_time, double run_time) {
    return guarded([&] {
        require_handle(solver);
        sph_profile p{};
        solver->impl->profile(&p);
        if (rank != 0) return;
        std::cout << std::setprecision(6) << "profile_solver seconds"
                  << " init=" << init_time << " run=" << run_time << " step=" << p.step << " rhs=" << p.step
                  << " diagnostics=" << p.diagnostics << " max_velocity=" << p.max_velocity
                  << " step_calls=" << p.step_calls << " rhs_calls=" << p.rhs_calls
                  << " diagnostics_calls=" << p.diagnostics_calls << " max_velocity_calls=" << p.max_velocity_calls
                  << "\n";
        std::cout << std::setprecision(6) << "profile_fft seconds"
                  << " forward=" << 0.0 << " backward=" << 0.0 << " forward_calls=" << p.forward_calls
                  << " backward_calls=" << p.backward_calls << "\n";
    });
}

int sph_solver_set_graph(sph_solver_t solver, int enabled) {
    return guarded([&] {
        require_handle(solver);
        solver->impl->set_graph(enabled != 0);
    });
}

int sph_solver_synchronize(sph_solver_t solver) {
    return guarded([&] {
        require_handle(solver);
        solver->impl->synchronize();
    });
}

int sph_solver_time_passes(sph_solver_t solver, int reps, double* ms, double* bytes) {
    return guarded([&] {
        require_handle(solver);
        solver->impl->time_passes(reps, ms, bytes);
    });
}

// run_taylor_green (src/ava_taylor_green_impl.hpp:1310-1410)
int sph_run_taylor_green(const sph_run_config* cfg_in, sph_diagnostics* final_row) {
    return guarded([&] {
        if (!cfg_in) throw std::runtime_error("null config");
        sph_run_config cfg = *cfg_in;
        if (cfg.dt <= 0.0 || cfg.cfl <= 0.0 || cfg.output_dt <= 0.0)
            throw std::runtime_error("--dt, --cfl, and --output-dt must be positive");
        if (cfg.output_every < 1) cfg.output_every = 1;
        if (cfg.print_every < 1) cfg.print_every = 1;
        const std::string dir = cfg.output_dir ? cfg.output_dir : "output_ava";
        std::filesystem::create_directories(dir);
        const std::string csv_path = dir + "/taylor_green_diagnostics.csv";
        std::ofstream csv(csv_path);
        csv << std::setprecision(17);
        csv << "t,kinetic_energy,dissipation,enstrophy,max_spectral_divergence,max_velocity,dt,cfl,main_dt,main_cfl\n";

        const auto clock0 = std::chrono::steady_clock::now();
        sph_solver_t solver = nullptr;
        if (sph_solver_create(cfg.n, cfg.reynolds, cfg.precision, &solver) != 0)
            throw std::runtime_error(sph::g_last_error);
        std::unique_ptr<sph_solver_s> guard(solver);
        const double init_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - clock0).count();
        const auto run0 = std::chrono::steady_clock::now();
        auto& S = *solver->impl;
        const double h = 2.0 * sph::kPi / static_cast<double>(cfg.n);
        double t = 0.0;
        int step = 0;
        int outputs = 0;
        sph_diagnostics last{};
        auto emit = [&](double time, bool force, double dt_, double cfl_, double mdt, double mcfl) {
            sph_diagnostics row = S.diagnostics(time);
            row.dt = dt_;
            row.cfl = cfl_;
            row.main_dt = mdt;
            row.main_cfl = mcfl;
            csv << row.t << ',' << row.kinetic_energy << ',' << row.dissipation << ',' << row.enstrophy << ','
                << row.max_spectral_divergence << ',' << row.max_velocity << ',' << dt_ << ',' << cfl_ << ',' << mdt
                << ',' << mcfl << '\n';
            csv.flush();
            if (force || outputs % cfg.print_every == 0) {
                std::cout << std::setprecision(10) << "step " << step << " t=" << row.t << " E=" << row.kinetic_energy
                          << " eps=" << row.dissipation << " max_u=" << row.max_velocity << " dt=" << dt_
                          << " cfl=" << cfl_ << " main_dt=" << mdt << " main_cfl=" << mcfl << "\n";
            }
            ++outputs;
            last = row;
        };
        if (cfg.adaptive_cfl) {
            double next_output = 0.0;
            emit(t, true, 0.0, 0.0, 0.0, 0.0);
            next_output += cfg.output_dt;
            while (t < cfg.t_end - 1.0e-14) {
                if (step >= cfg.max_steps)
                    throw std::runtime_error("adaptive CFL run exceeded --max-steps before reaching --t-end");
                const double vmax = S.max_velocity();
                double main_dt = cfg.dt;
                if (vmax > 1.0e-14) main_dt = std::min(main_dt, cfg.cfl * h / vmax);
                const double main_cfl = vmax * main_dt / h;
                double dt_step = std::min(main_dt, cfg.t_end - t);
                if (next_output <= cfg.t_end + 1.0e-14) dt_step = std::min(dt_step, next_output - t);
                if (dt_step <= 1.0e-14) {
                    next_output += cfg.output_dt;
                    continue;
                }
                const double actual_cfl = vmax * dt_step / h;
                S.step(dt_step);
                t += dt_step;
                ++step;
                if (t >= next_output - 1.0e-12 || t >= cfg.t_end - 1.0e-12) {
                    emit(t, t >= cfg.t_end - 1.0e-12, dt_step, actual_cfl, main_dt, main_cfl);
                    while (next_output <= t + 1.0e-12) next_output += cfg.output_dt;
                }
            }
        } else {
            for (step = 0; step <= cfg.steps; ++step) {
                if (step % cfg.output_every == 0 || step == cfg.steps) {
                    const double fixed_cfl = S.max_velocity() * cfg.dt / h;
                    emit(t, step == cfg.steps, step == 0 ? 0.0 : cfg.dt, step == 0 ? 0.0 : fixed_cfl,
                         step == 0 ? 0.0 : cfg.dt, step == 0 ? 0.0 : fixed_cfl);
                }
                if (step != cfg.steps) {
                    S.step(cfg.dt);
                    t += cfg.dt;
                }
            }
        }
        S.synchronize();
        const double run_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - run0).count();
        if (sph_solver_print_profile(solver, 0, init_time, run_time) != 0) throw std::runtime_error(sph::g_last_error);
        std::cout << "saved diagnostics: " << csv_path << "\n";
        if (final_row) *final_row = last;
    });
}

}  // extern "C"

// Host side of the MI355X spectral library: the FFT backend handle, the
// driver loop and the extern "C" boundary declared in include/spectral_hip.h.
// The solvers themselves are in solver.hpp, shared helpers in common.hpp.
#include "spectral_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"
#include "solver.hpp"

namespace sph {
namespace {

__global__ void __launch_bounds__(256) k_copy16(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        b[i] = a[i];
}

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
            upload(tw_[d], twiddle_table<T>(shape[d]));
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
        pt_[0]->fwd_x(fx, ny, nz, s);
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
        iz.tw = tw_[2].get();
        pt_[2]->inv_z(iz, ny, H, s);
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
struct sph_ensemble_s {
    std::unique_ptr<sph::EnsembleBase> impl;
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
void require_pointer(const void* p) {
    if (!p) throw std::runtime_error("null pointer argument");
}
void require_precision(int p) {
    if (p != SPH_FP32 && p != SPH_FP64) throw std::runtime_error("precision must be SPH_FP32 or SPH_FP64");
}
}  // namespace

extern "C" {

const char* sph_last_error(void) { return sph::g_last_error.c_str(); }

int sph_device_count(int* count) {
    return guarded([&] {
        require_pointer(count);
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *count = c;
    });
}

int sph_supported_length(int n) { return sph::pass_table<double>(n) != nullptr ? 1 : 0; }

int sph_fft_create(const int shape[3], const double lengths[3], int precision, sph_fft_t* out) {
    return guarded([&] {
        require_pointer(shape);
        require_pointer(lengths);
        require_pointer(out);
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
        require_pointer(out);
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

int sph_solver_step_times(sph_solver_t solver, double* ms, int max, int* count) {
    return guarded([&] {
        require_handle(solver);
        require_pointer(count);
        if (max > 0) require_pointer(ms);
        *count = solver->impl->step_times(ms, max);
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

int sph_solver_time_step_passes(sph_solver_t solver, double dt, int steps, double* ms, double* bytes) {
    return guarded([&] {
        require_handle(solver);
        require_pointer(ms);
        require_pointer(bytes);
        if (!(dt > 0.0)) throw std::runtime_error("dt must be positive");
        solver->impl->time_step_passes(dt, steps, ms, bytes);
    });
}

int sph_solver_step_adaptive(sph_solver_t solver, double dt_max, double cfl, double* max_velocity, double* dt,
                             int* fused) {
    return guarded([&] {
        require_handle(solver);
        require_pointer(max_velocity);
        require_pointer(dt);
        if (!(dt_max > 0.0) || !(cfl > 0.0)) throw std::runtime_error("--dt, --cfl, and --output-dt must be positive");
        solver->impl->step_adaptive(dt_max, cfl, max_velocity, dt);
        if (fused) *fused = solver->impl->fused_velocity_max() ? 1 : 0;
    });
}

int sph_solver_tuning(sph_solver_t solver, int* variant, int* lines, double* ms) {
    return guarded([&] {
        require_handle(solver);
        require_pointer(variant);
        require_pointer(lines);
        require_pointer(ms);
        solver->impl->tuning(variant, lines, ms);
    });
}

int sph_solver_tuned_pass(sph_solver_t solver, int pass, int* variant, int* lines, double* ms) {
    return guarded([&] {
        require_handle(solver);
        require_pointer(variant);
        require_pointer(lines);
        require_pointer(ms);
        solver->impl->tuned_pass(pass, variant, lines, ms);
    });
}

int sph_solver_radix_plan(sph_solver_t solver, int pass, int* radices, int* count) {
    return guarded([&] {
        require_handle(solver);
        require_pointer(radices);
        require_pointer(count);
        const sph::RadixPlan p = solver->impl->radix_plan_of(pass);
        for (int s = 0; s < p.count; ++s) radices[s] = p.r[s];
        *count = p.count;
    });
}

int sph_radix_plan(int n, int order, int* radices, int* count) {
    return guarded([&] {
        require_pointer(radices);
        require_pointer(count);
        if (order != 0 && order != 1) throw std::runtime_error("radix order must be 0 or 1");
        *count = 0;
        if (!sph::pass_table<float>(n)) return;
        const sph::RadixPlan p = sph::z_radix_plan(n, order);
        for (int s = 0; s < p.count; ++s) radices[s] = p.r[s];
        *count = p.count;
    });
}

int sph_solver_layout(sph_solver_t solver, int* layout, double* ms) {
    return guarded([&] {
        require_handle(solver);
        require_pointer(layout);
        require_pointer(ms);
        solver->impl->layout(layout, ms);
    });
}

int sph_measure_copy_bandwidth(int64_t bytes, int reps, double* gbs) {
    return guarded([&] {
        require_pointer(gbs);
        sph::require_device();
        if (bytes < 16 || reps < 1) throw std::runtime_error("copy bandwidth needs bytes >= 16 and reps >= 1");
        const long long n = bytes / 16;
        sph::DeviceArray<float4> a(static_cast<std::size_t>(n)), b(static_cast<std::size_t>(n));
        sph::hip_check(hipMemset(a.get(), 0, n * 16), "hipMemset");
        hipEvent_t e0, e1;
        sph::hip_check(hipEventCreate(&e0), "hipEventCreate");
        sph::hip_check(hipEventCreate(&e1), "hipEventCreate");
        sph::k_copy16<<<1024, 256>>>(a.get(), b.get(), n);
        sph::hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
        for (int r = 0; r < reps; ++r) sph::k_copy16<<<1024, 256>>>(a.get(), b.get(), n);
        sph::hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
        sph::hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        sph::hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbs = 2.0 * static_cast<double>(n) * 16.0 * reps / (ms * 1e-3) / 1e9;
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
                // dt from max |u| of the current state (the step's own x pass
                // reduces it: TaylorGreenSolver::step_cfl)
                double main_dt = cfg.dt, main_cfl = 0.0, dt_step = 0.0, actual_cfl = 0.0;
                S.step_cfl([&](double vmax) {
                    main_dt = cfg.dt;
                    if (vmax > 1.0e-14) main_dt = std::min(main_dt, cfg.cfl * h / vmax);
                    main_cfl = vmax * main_dt / h;
                    dt_step = std::min(main_dt, cfg.t_end - t);
                    if (next_output <= cfg.t_end + 1.0e-14) dt_step = std::min(dt_step, next_output - t);
                    actual_cfl = vmax * dt_step / h;
                    return dt_step <= 1.0e-14 ? 0.0 : dt_step;
                });
                if (dt_step <= 1.0e-14) {
                    next_output += cfg.output_dt;
                    continue;
                }
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

// ---------------------------------------------------------------------------
// Batched ensemble
// ---------------------------------------------------------------------------
int sph_ensemble_create(int count, const int* n, const double* nu, const uint64_t* seeds, double energy,
                        int precision, sph_ensemble_t* out) {
    return guarded([&] {
        require_pointer(out);
        require_pointer(n);
        require_pointer(nu);
        require_pointer(seeds);
        require_precision(precision);
        if (count < 1) throw std::runtime_error("ensemble needs at least one member");
        if (!(energy > 0.0)) throw std::runtime_error("initial energy must be positive");
        for (int j = 0; j < count; ++j) {
            if (n[j] <= 0) throw std::runtime_error("grid size must be positive");
            if (!sph_supported_length(n[j]))
                throw std::runtime_error("unsupported transform length " + std::to_string(n[j]));
            if (!(nu[j] > 0.0)) throw std::runtime_error("viscosity must be positive");
        }
        sph::require_device();
        auto h = std::make_unique<sph_ensemble_s>();
        if (precision == SPH_FP32)
            h->impl = std::make_unique<sph::Ensemble<float>>(count, n, nu, seeds, energy);
        else
            h->impl = std::make_unique<sph::Ensemble<double>>(count, n, nu, seeds, energy);
        *out = h.release();
    });
}

int sph_ensemble_destroy(sph_ensemble_t ens) {
    return guarded([&] { delete ens; });
}

int sph_ensemble_step_rk3(sph_ensemble_t ens, const double* dt) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(dt);
        ens->impl->step(dt);
    });
}

int sph_ensemble_diagnostics(sph_ensemble_t ens, int member, double t, sph_diagnostics* out) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(out);
        *out = ens->impl->diagnostics(member, t);
    });
}

int sph_ensemble_max_velocity(sph_ensemble_t ens, int member, double* out) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(out);
        *out = ens->impl->max_velocity(member);
    });
}

int sph_ensemble_get_uhat(sph_ensemble_t ens, int member, int component, double* host) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(host);
        ens->impl->get_uhat(member, component, host);
    });
}

int sph_ensemble_set_uhat(sph_ensemble_t ens, int member, int component, const double* host) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(host);
        ens->impl->set_uhat(member, component, host);
    });
}

int sph_ensemble_set_graph(sph_ensemble_t ens, int enabled) {
    return guarded([&] {
        require_handle(ens);
        ens->impl->set_graph(enabled != 0);
    });
}

int sph_ensemble_synchronize(sph_ensemble_t ens) {
    return guarded([&] {
        require_handle(ens);
        ens->impl->synchronize();
    });
}

int sph_ensemble_step_seconds(sph_ensemble_t ens, double* seconds, int64_t* steps) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(seconds);
        long long k = 0;
        *seconds = ens->impl->step_seconds(&k);
        if (steps) *steps = k;
    });
}

int sph_ensemble_time_passes(sph_ensemble_t ens, int reps, double* ms, double* bytes) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(ms);
        require_pointer(bytes);
        ens->impl->time_passes(reps, ms, bytes);
    });
}

int sph_ensemble_launch_mode(sph_ensemble_t ens, int* mode, double* step_ms) {
    return guarded([&] {
        require_handle(ens);
        require_pointer(mode);
        require_pointer(step_ms);
        ens->impl->launch_mode(mode, step_ms);
    });
}

int sph_random_solenoidal_field(int n, uint64_t seed, double energy, double* u0, double* u1, double* u2) {
    return guarded([&] {
        require_pointer(u0);
        require_pointer(u1);
        require_pointer(u2);
        sph::random_solenoidal_field(n, seed, energy, u0, u1, u2);
    });
}

}  // extern "C"

// spectral_hip.hpp — C++ classes with the reference's shapes, implemented
// over the C ABI in spectral_hip.h. Header-only; link libspectral_hip.so.
//
//   spectral::HipFft3D                   <- spectral::FftwFft3D  (src/fftw_fft.hpp:13-81)
//   spectral::SpectralMode               <- src/flups_fft.hpp:18-26
//   spectral::SpectralNavierStokes       <- src/spectral_solver.hpp:48-91
//   spectral::ava_solver::TaylorGreenSolver, Diagnostics, RunConfig,
//   run_taylor_green                     <- src/ava_taylor_green.hpp:8-55
//
// Errors surface as std::runtime_error carrying the library's message, as the
// reference's constructors and drivers throw (src/fftw_fft.cpp:12-24,
// src/ava_taylor_green_impl.hpp:704-709, 1311-1313).
#pragma once

#include "spectral_hip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

#if defined(SPECTRAL_HIP_WITH_MPI) && !defined(MPI_VERSION)
#include <mpi.h>
#endif

namespace spectral {

inline void sph_check(int rc) {
    if (rc != 0) throw std::runtime_error(sph_last_error());
}

// Communicator. The reference's backends take and return an MPI_Comm
// (src/fftw_fft.hpp:15,26) and its solver hands fft.comm() to MPI_Allreduce
// (src/spectral_solver.cpp:317,334). With MPI (<mpi.h> included before this
// header, or SPECTRAL_HIP_WITH_MPI defined and <mpi.h> on the include path),
// Comm is MPI_Comm for any MPI implementation (an int under MPICH, a pointer
// under Open MPI) and kCommSelf is MPI_COMM_SELF; the backend runs on one rank
// like the serial FFTW backend. Without MPI, Comm is a one-rank stand-in.
#ifdef MPI_VERSION
using Comm = MPI_Comm;
inline const Comm kCommSelf = MPI_COMM_SELF;
inline int communicator_size(Comm comm) {
    int initialized = 0;
    MPI_Initialized(&initialized);
    int size = 1;
    if (initialized) MPI_Comm_size(comm, &size);
    return size;
}
inline double allreduce_sum(double v, Comm comm) {
    int initialized = 0;
    MPI_Initialized(&initialized);
    if (!initialized) return v;
    double out = 0.0;
    MPI_Allreduce(&v, &out, 1, MPI_DOUBLE, MPI_SUM, comm);
    return out;
}
inline double allreduce_max(double v, Comm comm) {
    int initialized = 0;
    MPI_Initialized(&initialized);
    if (!initialized) return v;
    double out = 0.0;
    MPI_Allreduce(&v, &out, 1, MPI_DOUBLE, MPI_MAX, comm);
    return out;
}
#else
using Comm = int;
constexpr Comm kCommSelf = 0;
inline int communicator_size(Comm) { return 1; }
inline double allreduce_sum(double v, Comm) { return v; }
inline double allreduce_max(double v, Comm) { return v; }
#endif

struct SpectralMode {
    double kx = 0.0;
    double ky = 0.0;
    double kz = 0.0;
    double k2 = 0.0;
    int mx = 0;
    int my = 0;
    int mz = 0;
};

class HipFft3D {
public:
    HipFft3D(std::array<int, 3> shape, std::array<double, 3> lengths, Comm comm = kCommSelf,
             int precision = SPH_FP64)
        : shape_(shape), lengths_(lengths) {
        // one rank, as the serial FFTW backend requires (src/fftw_fft.cpp:10-14)
        if (communicator_size(comm) != 1) throw std::runtime_error("The HIP backend supports one MPI rank");
        sph_check(sph_fft_create(shape_.data(), lengths_.data(), precision, &h_));
        int64_t np = 0, nm = 0;
        sph_check(sph_fft_sizes(h_, &np, &nm));
        modes_ = static_cast<std::size_t>(nm);
        mx_.resize(modes_);
        my_.resize(modes_);
        mz_.resize(modes_);
        sph_check(sph_fft_spectral_modes(h_, mx_.data(), my_.data(), mz_.data()));
    }
    ~HipFft3D() { sph_fft_destroy(h_); }
    HipFft3D(const HipFft3D&) = delete;
    HipFft3D& operator=(const HipFft3D&) = delete;

    const std::array<int, 3>& shape() const { return shape_; }
    const std::array<double, 3>& lengths() const { return lengths_; }
    int physical_size() const { return shape_[0] * shape_[1] * shape_[2]; }
    int global_physical_size() const { return physical_size(); }
    // interleaved doubles of the half spectrum
    std::size_t spectral_storage_size() const { return 2 * modes_; }
    Comm comm() const { return kCommSelf; }
    int rank() const { return 0; }
    int comm_size() const { return 1; }
    sph_fft_t handle() const { return h_; }

    template <class Func>
    void for_each_physical(Func&& func) const {
        for (int iz = 0; iz < shape_[2]; ++iz)
            for (int iy = 0; iy < shape_[1]; ++iy)
                for (int ix = 0; ix < shape_[0]; ++ix)
                    func(static_cast<std::size_t>(ix + shape_[0] * (iy + shape_[1] * iz)), ix, iy, iz);
    }

    // f(id, mode) with id the index of the real part in the interleaved storage
    template <class Func>
    void for_each_spectral(Func&& func) const {
        constexpr double two_pi = 6.283185307179586476925286766559005768;
        for (std::size_t i = 0; i < modes_; ++i) {
            SpectralMode m;
            m.mx = mx_[i];
            m.my = my_[i];
            m.mz = mz_[i];
            m.kx = two_pi * m.mx / lengths_[0];
            m.ky = two_pi * m.my / lengths_[1];
            m.kz = two_pi * m.mz / lengths_[2];
            m.k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
            func(2 * i, m);
        }
    }

    std::vector<double> forward(const std::vector<double>& physical) {
        if (physical.size() != static_cast<std::size_t>(physical_size()))
            throw std::runtime_error("physical field has wrong size");
        std::vector<double> out(spectral_storage_size());
        sph_check(sph_fft_forward_host(h_, physical.data(), out.data()));
        return out;
    }

    std::vector<double> backward(const std::vector<double>& spectral) {
        if (spectral.size() != spectral_storage_size())
            throw std::runtime_error("spectral field has wrong storage size");
        std::vector<double> out(static_cast<std::size_t>(physical_size()));
        sph_check(sph_fft_backward_host(h_, spectral.data(), out.data()));
        return out;
    }

    // device-pointer transforms (FlupsFft3D::forward_device/backward_device)
    void forward_device(const void* physical, void* spectral, void* stream = nullptr) {
        sph_check(sph_fft_forward(h_, physical, spectral, stream));
    }
    void backward_device(const void* spectral, void* physical, void* stream = nullptr) {
        sph_check(sph_fft_backward(h_, spectral, physical, stream));
    }

private:
    std::array<int, 3> shape_{};
    std::array<double, 3> lengths_{};
    std::size_t modes_ = 0;
    std::vector<int> mx_, my_, mz_;
    sph_fft_t h_ = nullptr;
};

using FftBackend = HipFft3D;

using ScalarPhysical = std::vector<double>;
using ScalarSpectral = std::vector<double>;
using VectorPhysical = std::array<ScalarPhysical, 3>;
using VectorSpectral = std::array<ScalarSpectral, 3>;

struct Diagnostics {
    double t = 0.0;
    double energy = 0.0;
    double dissipation = 0.0;
    double enstrophy = 0.0;
    double max_spectral_divergence = 0.0;
};

// Value-semantics solver on host vectors with the transforms on the GPU
// (src/spectral_solver.cpp:35-419): arbitrary box lengths, per-axis sizes and
// switchable phase-shift dealiasing.
class SpectralNavierStokes {
public:
    SpectralNavierStokes(std::array<int, 3> shape, std::array<double, 3> lengths, double viscosity,
                         Comm comm = kCommSelf, bool phase_shift_dealiasing = true)
        : fft_(shape, lengths, comm), nu_(viscosity), dealias_(phase_shift_dealiasing) {
        if (nu_ < 0.0) throw std::runtime_error("viscosity must be non-negative");
        const std::size_t nm = fft_.spectral_storage_size() / 2;
        shift_.resize(nm);
        keep_.resize(nm);
        const auto& L = fft_.lengths();
        const auto& n = fft_.shape();
        const double hx = 0.5 * L[0] / n[0], hy = 0.5 * L[1] / n[1], hz = 0.5 * L[2] / n[2];
        fft_.for_each_spectral([&](std::size_t id, const SpectralMode& m) {
            shift_[id / 2] = std::exp(std::complex<double>(0.0, m.kx * hx + m.ky * hy + m.kz * hz));
            const double ex = static_cast<double>(m.mx * m.mx) / std::pow(n[0] / 2.0, 2.0);
            const double ey = static_cast<double>(m.my * m.my) / std::pow(n[1] / 2.0, 2.0);
            const double ez = static_cast<double>(m.mz * m.mz) / std::pow(n[2] / 2.0, 2.0);
            keep_[id / 2] = ex + ey + ez <= 1.0;
        });
    }

    FftBackend& fft() { return fft_; }
    const FftBackend& fft() const { return fft_; }

    VectorPhysical taylor_green(double amplitude = 1.0, int mode = 1) const {
        constexpr double pi = 3.141592653589793238462643383279502884;
        const auto& n = fft_.shape();
        const auto& L = fft_.lengths();
        const std::size_t np = static_cast<std::size_t>(fft_.physical_size());
        VectorPhysical u = {ScalarPhysical(np, 0.0), ScalarPhysical(np, 0.0), ScalarPhysical(np, 0.0)};
        fft_.for_each_physical([&](std::size_t id, int ix, int iy, int iz) {
            const double ax = 2.0 * pi * mode * ((ix + 0.5) * L[0] / n[0]) / L[0];
            const double ay = 2.0 * pi * mode * ((iy + 0.5) * L[1] / n[1]) / L[1];
            const double az = 2.0 * pi * mode * ((iz + 0.5) * L[2] / n[2]) / L[2];
            u[0][id] = amplitude * std::sin(ax) * std::cos(ay) * std::cos(az);
            u[1][id] = -amplitude * std::cos(ax) * std::sin(ay) * std::cos(az);
        });
        return u;
    }

    VectorSpectral to_spectral(const VectorPhysical& u) {
        VectorSpectral out;
        for (int c = 0; c < 3; ++c) out[c] = fft_.forward(u[c]);
        return out;
    }
    VectorPhysical to_physical(const VectorSpectral& u_hat) {
        VectorPhysical out;
        for (int c = 0; c < 3; ++c) out[c] = fft_.backward(u_hat[c]);
        return out;
    }

    // Leray projection; k = 0 is then zeroed (spectral_solver.cpp:94-109)
    void project(VectorSpectral& u) const {
        fft_.for_each_spectral([&](std::size_t id, const SpectralMode& m) {
            if (m.k2 == 0.0) {
                for (auto& c : u) c[id] = c[id + 1] = 0.0;
                return;
            }
            const auto ux = at(u[0], id), uy = at(u[1], id), uz = at(u[2], id);
            const auto f = (m.kx * ux + m.ky * uy + m.kz * uz) / m.k2;
            put(u[0], id, ux - m.kx * f);
            put(u[1], id, uy - m.ky * f);
            put(u[2], id, uz - m.kz * f);
        });
    }

    void truncate(VectorSpectral& u) const {
        fft_.for_each_spectral([&](std::size_t id, const SpectralMode&) {
            if (!keep_[id / 2])
                for (auto& c : u) c[id] = c[id + 1] = 0.0;
        });
    }

    // dealiased spectrum of a*b (spectral_solver.cpp:122-158)
    ScalarSpectral product_dealiased(const ScalarSpectral& a_hat, const ScalarSpectral& b_hat) {
        const auto a = fft_.backward(a_hat), b = fft_.backward(b_hat);
        ScalarPhysical p(a.size());
        for (std::size_t i = 0; i < p.size(); ++i) p[i] = a[i] * b[i];
        auto out = fft_.forward(p);
        auto as = a_hat, bs = b_hat;
        apply_shift(as);
        apply_shift(bs);
        const auto a2 = fft_.backward(as), b2 = fft_.backward(bs);
        for (std::size_t i = 0; i < p.size(); ++i) p[i] = a2[i] * b2[i];
        const auto q = fft_.forward(p);
        combine(out, q);
        return out;
    }

    VectorSpectral nonlinear_rhs(const VectorSpectral& u_hat) {
        const auto u = to_physical(u_hat);
        VectorPhysical us;
        if (dealias_) {
            auto sh = u_hat;
            for (auto& c : sh) apply_shift(c);
            us = to_physical(sh);
        }
        std::array<std::array<ScalarSpectral, 3>, 3> P;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) {
                ScalarPhysical p(u[0].size());
                for (std::size_t i = 0; i < p.size(); ++i) p[i] = u[a][i] * u[b][i];
                auto ph = fft_.forward(p);
                if (dealias_) {
                    for (std::size_t i = 0; i < p.size(); ++i) p[i] = us[a][i] * us[b][i];
                    combine(ph, fft_.forward(p));
                } else {
                    mask(ph);
                }
                P[a][b] = ph;
                if (a != b) P[b][a] = ph;
            }
        VectorSpectral rhs;
        for (int i = 0; i < 3; ++i) {
            rhs[i].assign(fft_.spectral_storage_size(), 0.0);
            fft_.for_each_spectral([&](std::size_t id, const SpectralMode& m) {
                const auto s = m.kx * at(P[0][i], id) + m.ky * at(P[1][i], id) + m.kz * at(P[2][i], id);
                put(rhs[i], id, std::complex<double>(0.0, -1.0) * s);
            });
        }
        truncate(rhs);
        return rhs;
    }

    VectorSpectral rhs(const VectorSpectral& u_hat) {
        auto out = nonlinear_rhs(u_hat);
        fft_.for_each_spectral([&](std::size_t id, const SpectralMode& m) {
            for (int c = 0; c < 3; ++c) put(out[c], id, at(out[c], id) - nu_ * m.k2 * at(u_hat[c], id));
        });
        project(out);
        truncate(out);
        return out;
    }

    // Williamson low-storage RK3 (spectral_solver.cpp:262-309)
    VectorSpectral step_rk3(const VectorSpectral& u_hat, double t, double dt) {
        (void)t;
        auto u = u_hat;
        auto g = rhs(u);
        axpy(u, dt / 3.0, g);
        project(u);
        truncate(u);
        auto h = rhs(u);
        for (int c = 0; c < 3; ++c)
            for (std::size_t i = 0; i < u[c].size(); ++i) g[c][i] = -(5.0 / 9.0) * g[c][i] + h[c][i];
        axpy(u, (15.0 / 16.0) * dt, g);
        project(u);
        truncate(u);
        h = rhs(u);
        for (int c = 0; c < 3; ++c)
            for (std::size_t i = 0; i < u[c].size(); ++i) g[c][i] = -(153.0 / 128.0) * g[c][i] + h[c][i];
        axpy(u, (8.0 / 15.0) * dt, g);
        project(u);
        truncate(u);
        return u;
    }

    // sums and maxima over the ranks of fft().comm() (spectral_solver.cpp:311-336)
    double kinetic_energy(const VectorPhysical& u) const {
        double s = 0.0;
        for (std::size_t i = 0; i < u[0].size(); ++i) s += 0.5 * (u[0][i] * u[0][i] + u[1][i] * u[1][i] + u[2][i] * u[2][i]);
        return allreduce_sum(s, fft_.comm()) / static_cast<double>(fft_.global_physical_size());
    }
    double kinetic_energy_spectral(const VectorSpectral& u_hat) { return kinetic_energy(to_physical(u_hat)); }
    double max_velocity(const VectorSpectral& u_hat) {
        const auto u = to_physical(u_hat);
        double m = 0.0;
        for (std::size_t i = 0; i < u[0].size(); ++i)
            m = std::max(m, std::sqrt(u[0][i] * u[0][i] + u[1][i] * u[1][i] + u[2][i] * u[2][i]));
        return allreduce_max(m, fft_.comm());
    }
    VectorSpectral vorticity_hat(const VectorSpectral& u_hat) const {
        VectorSpectral w;
        for (auto& c : w) c.assign(fft_.spectral_storage_size(), 0.0);
        const std::complex<double> I(0.0, 1.0);
        fft_.for_each_spectral([&](std::size_t id, const SpectralMode& m) {
            const auto ux = at(u_hat[0], id), uy = at(u_hat[1], id), uz = at(u_hat[2], id);
            put(w[0], id, I * (m.ky * uz - m.kz * uy));
            put(w[1], id, I * (m.kz * ux - m.kx * uz));
            put(w[2], id, I * (m.kx * uy - m.ky * ux));
        });
        return w;
    }
    double enstrophy(const VectorSpectral& u_hat) {
        const auto w = to_physical(vorticity_hat(u_hat));
        double s = 0.0;
        for (std::size_t i = 0; i < w[0].size(); ++i) s += w[0][i] * w[0][i] + w[1][i] * w[1][i] + w[2][i] * w[2][i];
        return 0.5 * allreduce_sum(s, fft_.comm()) / static_cast<double>(fft_.global_physical_size());
    }
    double dissipation(const VectorSpectral& u_hat) { return 2.0 * nu_ * enstrophy(u_hat); }
    double max_spectral_divergence(const VectorSpectral& u_hat) const {
        double m = 0.0;
        fft_.for_each_spectral([&](std::size_t id, const SpectralMode& md) {
            m = std::max(m, std::abs(md.kx * at(u_hat[0], id) + md.ky * at(u_hat[1], id) + md.kz * at(u_hat[2], id)));
        });
        return allreduce_max(m, fft_.comm());
    }
    Diagnostics diagnostics(const VectorSpectral& u_hat, double t) {
        const double ens = enstrophy(u_hat);
        return Diagnostics{t, kinetic_energy_spectral(u_hat), 2.0 * nu_ * ens, ens, max_spectral_divergence(u_hat)};
    }

private:
    static std::complex<double> at(const std::vector<double>& a, std::size_t id) { return {a[id], a[id + 1]}; }
    static void put(std::vector<double>& a, std::size_t id, std::complex<double> z) {
        a[id] = z.real();
        a[id + 1] = z.imag();
    }
    void apply_shift(ScalarSpectral& a) const {
        for (std::size_t i = 0; i < shift_.size(); ++i) put(a, 2 * i, at(a, 2 * i) * shift_[i]);
    }
    void mask(ScalarSpectral& a) const {
        for (std::size_t i = 0; i < keep_.size(); ++i)
            if (!keep_[i]) a[2 * i] = a[2 * i + 1] = 0.0;
    }
    // 0.5*(P + Q*conj(shift)), then the truncation mask
    void combine(ScalarSpectral& p, const ScalarSpectral& q) const {
        for (std::size_t i = 0; i < shift_.size(); ++i) {
            auto v = 0.5 * (at(p, 2 * i) + at(q, 2 * i) * std::conj(shift_[i]));
            if (!keep_[i]) v = 0.0;
            put(p, 2 * i, v);
        }
    }
    static void axpy(VectorSpectral& u, double a, const VectorSpectral& g) {
        for (int c = 0; c < 3; ++c)
            for (std::size_t i = 0; i < u[c].size(); ++i) u[c][i] += a * g[c][i];
    }

    FftBackend fft_;
    double nu_ = 0.0;
    bool dealias_ = true;
    std::vector<std::complex<double>> shift_;
    std::vector<unsigned char> keep_;
};

inline void write_diagnostics_csv(const std::string& path, const std::vector<Diagnostics>& history) {
    std::ofstream f(path);
    f << "t,kinetic_energy,dissipation,enstrophy,max_spectral_divergence\n";
    f << std::setprecision(17);
    for (const auto& r : history)
        f << r.t << ',' << r.energy << ',' << r.dissipation << ',' << r.enstrophy << ',' << r.max_spectral_divergence
          << '\n';
}

// Binary PPM (P6) image of |ω| on the mid z plane iz = n_z/2, scaled by the
// plane's maximum (1 when the plane is zero), rows from iy = n_y-1 down to 0;
// colour of q = |ω|/max: r = 255·√q, g = 255·q², b = 255·(1−q)/4, each
// truncated to a byte (src/spectral_solver.cpp:421-486). One rank holds the
// whole plane, so nothing is gathered.
inline void write_vorticity_ppm(const std::string& path, SpectralNavierStokes& solver, const VectorSpectral& u_hat) {
    const auto w = solver.to_physical(solver.vorticity_hat(u_hat));
    const auto& n = solver.fft().shape();
    const int iz = n[2] / 2;
    const std::size_t plane = static_cast<std::size_t>(n[0]) * n[1];
    std::vector<double> mag(plane);
    double peak = 0.0;
    for (std::size_t i = 0; i < plane; ++i) {
        const std::size_t id = static_cast<std::size_t>(iz) * plane + i;   // x fastest, then y, then z
        mag[i] = std::sqrt(w[0][id] * w[0][id] + w[1][id] * w[1][id] + w[2][id] * w[2][id]);
        peak = std::max(peak, mag[i]);
    }
    if (peak == 0.0) peak = 1.0;
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    f << "P6\n" << n[0] << " " << n[1] << "\n255\n";
    std::vector<unsigned char> row(3 * static_cast<std::size_t>(n[0]));
    for (int iy = n[1] - 1; iy >= 0; --iy) {
        for (int ix = 0; ix < n[0]; ++ix) {
            const double q = std::clamp(mag[static_cast<std::size_t>(iy) * n[0] + ix] / peak, 0.0, 1.0);
            row[3 * ix + 0] = static_cast<unsigned char>(255.0 * std::sqrt(q));
            row[3 * ix + 1] = static_cast<unsigned char>(255.0 * q * q);
            row[3 * ix + 2] = static_cast<unsigned char>(255.0 * (1.0 - q) * 0.25);
        }
        f.write(reinterpret_cast<const char*>(row.data()), static_cast<std::streamsize>(row.size()));
    }
}

namespace ava_solver {

struct Diagnostics {
    double t = 0.0;
    double kinetic_energy = 0.0;
    double dissipation = 0.0;
    double enstrophy = 0.0;
    double max_spectral_divergence = 0.0;
    double max_velocity = 0.0;
    double dt = 0.0;
    double cfl = 0.0;
    double main_dt = 0.0;
    double main_cfl = 0.0;
};

struct RunConfig {
    int n = 64;
    int steps = 80;
    int max_steps = 10000000;
    int output_every = 1;
    int print_every = 10;
    double dt = 0.0025;
    double cfl = 0.4;
    double t_end = 0.2;
    double output_dt = 0.1;
    double reynolds = 1600.0;
    bool adaptive_cfl = false;
    std::string output_dir = "output_ava";
    int precision = SPH_FP64;  // the reference's SPECTRAL_PRECISION, chosen at run time
};

class TaylorGreenSolver {
public:
    TaylorGreenSolver(int n, double reynolds, int precision = SPH_FP64) {
        sph_check(sph_solver_create(n, reynolds, precision, &h_));
    }
    ~TaylorGreenSolver() { sph_solver_destroy(h_); }
    TaylorGreenSolver(const TaylorGreenSolver&) = delete;
    TaylorGreenSolver& operator=(const TaylorGreenSolver&) = delete;

    void step_rk3(double dt) { sph_check(sph_solver_step_rk3(h_, dt)); }
    Diagnostics diagnostics(double t) {
        sph_diagnostics d{};
        sph_check(sph_solver_diagnostics(h_, t, &d));
        return Diagnostics{d.t, d.kinetic_energy, d.dissipation, d.enstrophy, d.max_spectral_divergence,
                           d.max_velocity, d.dt, d.cfl, d.main_dt, d.main_cfl};
    }
    double max_velocity() {
        double v = 0.0;
        sph_check(sph_solver_max_velocity(h_, &v));
        return v;
    }
    void print_profile(int rank, double init_time, double run_time) const {
        sph_check(sph_solver_print_profile(h_, rank, init_time, run_time));
    }
    sph_solver_t handle() const { return h_; }

private:
    sph_solver_t h_ = nullptr;
};

inline void run_taylor_green(const RunConfig& c) {
    sph_run_config r{};
    r.n = c.n;
    r.steps = c.steps;
    r.max_steps = c.max_steps;
    r.output_every = c.output_every;
    r.print_every = c.print_every;
    r.dt = c.dt;
    r.cfl = c.cfl;
    r.t_end = c.t_end;
    r.output_dt = c.output_dt;
    r.reynolds = c.reynolds;
    r.adaptive_cfl = c.adaptive_cfl ? 1 : 0;
    r.precision = c.precision;
    r.output_dir = c.output_dir.c_str();
    sph_check(sph_run_taylor_green(&r, nullptr));
}

}  // namespace ava_solver
}  // namespace spectral

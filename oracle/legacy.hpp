// CPU oracle for the reference's legacy value-semantics solver.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ through liboracle.so as the
// checker of the GPU library's SpectralNavierStokes shim, never by the product.
//
// Restates spectral::SpectralNavierStokes (src/spectral_solver.cpp:35-409)
// over its serial FFTW backend (src/fftw_fft.hpp:13-81, src/fftw_fft.cpp:53-85):
// every field of nx*ny*nz points is stored C2C as interleaved doubles, x
// fastest; forward transforms are unnormalised (sign e^{-i}), backward
// transforms take the real part and multiply by 1/(nx*ny*nz). What differs
// from the AVA solver restated in oracle.hpp is kept as the reference writes
// it: per-axis sizes and box lengths (wavenumbers 2*pi*m/L), the per-axis
// ellipsoidal keep mask (:392-398), the phase shift exp(i k.h/2) from physical
// wavenumbers (:44-54), projection that skips k2 == 0 and then zeroes the mean
// (:94-109, 400-409), the nonlinear term -i k_j P_ji summed over j after all
// six products (:160-240), the dealiasing switch (:169, 193), and the RK
// coefficients written (dt/3.0), (15.0/16.0)*dt, (8.0/15.0)*dt (:262-309).
// The transforms are oracle.hpp's Grid3d (FFTW 3.3.10 is absent here).
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <vector>

#include "oracle.hpp"

namespace oracle {

class LegacyNS {
public:
    using Vec = std::vector<double>;
    using Vec3 = std::array<Vec, 3>;
    using C = std::complex<double>;

    struct Mode {
        double kx, ky, kz, k2;
        int mx, my, mz;
    };

    struct Row {
        double t, energy, dissipation, enstrophy, max_spectral_divergence;
    };

    // src/spectral_solver.cpp:35-55 and src/fftw_fft.cpp:8-36
    LegacyNS(std::array<int, 3> shape, std::array<double, 3> lengths, double viscosity, bool dealias)
        : n_(shape), L_(lengths), nu_(viscosity), dealias_(dealias),
          grid_(shape[0] > 1 ? shape[0] : 2, shape[1] > 1 ? shape[1] : 2, shape[2] > 1 ? shape[2] : 2) {
        for (int d : n_)
            if (d <= 1) throw std::runtime_error("all grid sizes must be greater than one");
        for (double l : L_)
            if (l <= 0.0) throw std::runtime_error("all domain lengths must be positive");
        if (nu_ < 0.0) throw std::runtime_error("viscosity must be non-negative");
        np_ = static_cast<std::size_t>(n_[0]) * n_[1] * n_[2];
        phase_.resize(np_);
        keep_.resize(np_);
        const double dx = 0.5 * L_[0] / n_[0];
        const double dy = 0.5 * L_[1] / n_[1];
        const double dz = 0.5 * L_[2] / n_[2];
        for_each_spectral([&](std::size_t id, const Mode& m) {
            phase_[id / 2] = std::exp(C(0.0, m.kx * dx + m.ky * dy + m.kz * dz));
            const double ex = static_cast<double>(m.mx * m.mx) / std::pow(n_[0] / 2.0, 2.0);
            const double ey = static_cast<double>(m.my * m.my) / std::pow(n_[1] / 2.0, 2.0);
            const double ez = static_cast<double>(m.mz * m.mz) / std::pow(n_[2] / 2.0, 2.0);
            keep_[id / 2] = ex + ey + ez <= 1.0 ? 1 : 0;
        });
    }

    std::size_t physical_size() const { return np_; }
    std::size_t spectral_storage_size() const { return 2 * np_; }

    // src/fftw_fft.hpp:45-71: iz, iy, ix loops, id = 2*(ix + nx*(iy + ny*iz))
    template <class F>
    void for_each_spectral(F&& f) const {
        constexpr double two_pi = 6.283185307179586476925286766559005768;
        for (int iz = 0; iz < n_[2]; ++iz) {
            const int mz = signed_mode(iz, 2);
            for (int iy = 0; iy < n_[1]; ++iy) {
                const int my = signed_mode(iy, 1);
                for (int ix = 0; ix < n_[0]; ++ix) {
                    Mode m;
                    m.mx = signed_mode(ix, 0);
                    m.my = my;
                    m.mz = mz;
                    m.kx = two_pi * m.mx / L_[0];
                    m.ky = two_pi * m.my / L_[1];
                    m.kz = two_pi * m.mz / L_[2];
                    m.k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
                    f(2 * (static_cast<std::size_t>(ix) + n_[0] * (static_cast<std::size_t>(iy) + n_[1] * iz)), m);
                }
            }
        }
    }

    // src/fftw_fft.cpp:53-68
    Vec forward(const Vec& phys) const {
        if (phys.size() != np_) throw std::runtime_error("physical field has wrong size");
        std::vector<Cx<double>> a(np_);
        for (std::size_t i = 0; i < np_; ++i) a[i] = {phys[i], 0.0};
        grid_.exec(a.data(), -1);
        Vec out(2 * np_);
        for (std::size_t i = 0; i < np_; ++i) {
            out[2 * i] = a[i].x;
            out[2 * i + 1] = a[i].y;
        }
        return out;
    }

    // src/fftw_fft.cpp:70-85
    Vec backward(const Vec& spec) const {
        if (spec.size() != 2 * np_) throw std::runtime_error("spectral field has wrong storage size");
        std::vector<Cx<double>> a(np_);
        for (std::size_t i = 0; i < np_; ++i) a[i] = {spec[2 * i], spec[2 * i + 1]};
        grid_.exec(a.data(), +1);
        const double scale = 1.0 / static_cast<double>(np_);
        Vec out(np_);
        for (std::size_t i = 0; i < np_; ++i) out[i] = scale * a[i].x;
        return out;
    }

    // src/spectral_solver.cpp:57-76
    Vec3 taylor_green(double amplitude = 1.0, int mode = 1) const {
        constexpr double pi = 3.141592653589793238462643383279502884;
        Vec3 u = {Vec(np_, 0.0), Vec(np_, 0.0), Vec(np_, 0.0)};
        for (int iz = 0; iz < n_[2]; ++iz)
            for (int iy = 0; iy < n_[1]; ++iy)
                for (int ix = 0; ix < n_[0]; ++ix) {
                    const std::size_t id = static_cast<std::size_t>(ix) + n_[0] * (static_cast<std::size_t>(iy) + n_[1] * iz);
                    const double x = (ix + 0.5) * L_[0] / n_[0];
                    const double y = (iy + 0.5) * L_[1] / n_[1];
                    const double z = (iz + 0.5) * L_[2] / n_[2];
                    const double ax = 2.0 * pi * mode * x / L_[0];
                    const double ay = 2.0 * pi * mode * y / L_[1];
                    const double az = 2.0 * pi * mode * z / L_[2];
                    u[0][id] = amplitude * std::sin(ax) * std::cos(ay) * std::cos(az);
                    u[1][id] = -amplitude * std::cos(ax) * std::sin(ay) * std::cos(az);
                }
        return u;
    }

    Vec3 to_spectral(const Vec3& u) const { return {forward(u[0]), forward(u[1]), forward(u[2])}; }
    Vec3 to_physical(const Vec3& u) const { return {backward(u[0]), backward(u[1]), backward(u[2])}; }

    // src/spectral_solver.cpp:94-109
    void project(Vec3& u) const {
        for_each_spectral([&](std::size_t id, const Mode& m) {
            if (m.k2 == 0.0) return;
            const C ux = at(u[0], id), uy = at(u[1], id), uz = at(u[2], id);
            const C dot = m.kx * ux + m.ky * uy + m.kz * uz;
            const C factor = dot / m.k2;
            put(u[0], id, ux - m.kx * factor);
            put(u[1], id, uy - m.ky * factor);
            put(u[2], id, uz - m.kz * factor);
        });
        zero_mean(u);
    }

    // src/spectral_solver.cpp:111-120
    void truncate(Vec3& u) const {
        for_each_spectral([&](std::size_t id, const Mode&) {
            if (!keep_[id / 2])
                for (auto& c : u) c[id] = c[id + 1] = 0.0;
        });
    }

    // src/spectral_solver.cpp:160-240
    Vec3 nonlinear_rhs(const Vec3& u_hat) const {
        const Vec3 u = to_physical(u_hat);
        Vec3 us;
        if (dealias_) {
            Vec3 sh = u_hat;
            for_each_spectral([&](std::size_t id, const Mode&) {
                const C ph = phase_[id / 2];
                for (int c = 0; c < 3; ++c) put(sh[c], id, at(sh[c], id) * ph);
            });
            us = to_physical(sh);
        }
        std::array<std::array<Vec, 3>, 3> P;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) {
                Vec c0(np_), cs;
                if (dealias_) cs.assign(np_, 0.0);
                for (std::size_t p = 0; p < np_; ++p) {
                    c0[p] = u[a][p] * u[b][p];
                    if (dealias_) cs[p] = us[a][p] * us[b][p];
                }
                const Vec c0_hat = forward(c0);
                Vec cs_hat;
                if (dealias_) cs_hat = forward(cs);
                Vec out(2 * np_, 0.0);
                for_each_spectral([&](std::size_t id, const Mode&) {
                    C value = at(c0_hat, id);
                    if (dealias_) value = 0.5 * (value + at(cs_hat, id) * std::conj(phase_[id / 2]));
                    if (!keep_[id / 2]) value = 0.0;
                    put(out, id, value);
                });
                P[a][b] = out;
                if (a != b) P[b][a] = out;
            }
        Vec3 rhs = {Vec(2 * np_, 0.0), Vec(2 * np_, 0.0), Vec(2 * np_, 0.0)};
        const C imag(0.0, 1.0);
        for (int i = 0; i < 3; ++i)
            for_each_spectral([&](std::size_t id, const Mode& m) {
                put(rhs[i], id, -imag * (m.kx * at(P[0][i], id) + m.ky * at(P[1][i], id) + m.kz * at(P[2][i], id)));
            });
        truncate(rhs);
        return rhs;
    }

    // src/spectral_solver.cpp:242-260
    Vec3 rhs(const Vec3& u_hat) const {
        Vec3 out = nonlinear_rhs(u_hat);
        for_each_spectral([&](std::size_t id, const Mode& m) {
            for (int c = 0; c < 3; ++c) put(out[c], id, at(out[c], id) - nu_ * m.k2 * at(u_hat[c], id));
        });
        project(out);
        truncate(out);
        return out;
    }

    // src/spectral_solver.cpp:262-309
    Vec3 step_rk3(const Vec3& u_hat, double dt) const {
        Vec3 u = u_hat;
        Vec3 g = rhs(u);
        for (int c = 0; c < 3; ++c)
            for (std::size_t i = 0; i < u[c].size(); ++i) u[c][i] += (dt / 3.0) * g[c][i];
        project(u);
        truncate(u);
        Vec3 h = rhs(u);
        for (int c = 0; c < 3; ++c)
            for (std::size_t i = 0; i < u[c].size(); ++i) {
                g[c][i] = -(5.0 / 9.0) * g[c][i] + h[c][i];
                u[c][i] += (15.0 / 16.0) * dt * g[c][i];
            }
        project(u);
        truncate(u);
        h = rhs(u);
        for (int c = 0; c < 3; ++c)
            for (std::size_t i = 0; i < u[c].size(); ++i) {
                g[c][i] = -(153.0 / 128.0) * g[c][i] + h[c][i];
                u[c][i] += (8.0 / 15.0) * dt * g[c][i];
            }
        project(u);
        truncate(u);
        return u;
    }

    // src/spectral_solver.cpp:311-386 (one rank: the all-reduces are identities)
    double kinetic_energy(const Vec3& u) const {
        double s = 0.0;
        for (std::size_t i = 0; i < np_; ++i) s += 0.5 * (u[0][i] * u[0][i] + u[1][i] * u[1][i] + u[2][i] * u[2][i]);
        return s / static_cast<double>(np_);
    }
    double enstrophy(const Vec3& u_hat) const {
        Vec3 w = {Vec(2 * np_, 0.0), Vec(2 * np_, 0.0), Vec(2 * np_, 0.0)};
        const C imag(0.0, 1.0);
        for_each_spectral([&](std::size_t id, const Mode& m) {
            const C ux = at(u_hat[0], id), uy = at(u_hat[1], id), uz = at(u_hat[2], id);
            put(w[0], id, imag * (m.ky * uz - m.kz * uy));
            put(w[1], id, imag * (m.kz * ux - m.kx * uz));
            put(w[2], id, imag * (m.kx * uy - m.ky * ux));
        });
        const Vec3 om = to_physical(w);
        double s = 0.0;
        for (std::size_t i = 0; i < np_; ++i) s += om[0][i] * om[0][i] + om[1][i] * om[1][i] + om[2][i] * om[2][i];
        return 0.5 * s / static_cast<double>(np_);
    }
    double max_spectral_divergence(const Vec3& u_hat) const {
        double m = 0.0;
        for_each_spectral([&](std::size_t id, const Mode& md) {
            const C d = md.kx * at(u_hat[0], id) + md.ky * at(u_hat[1], id) + md.kz * at(u_hat[2], id);
            m = std::max(m, std::abs(d));
        });
        return m;
    }
    Row diagnostics(const Vec3& u_hat, double t) const {
        const double ens = enstrophy(u_hat);
        return Row{t, kinetic_energy(to_physical(u_hat)), 2.0 * nu_ * ens, ens, max_spectral_divergence(u_hat)};
    }

private:
    int signed_mode(int idx, int dim) const { return idx <= n_[dim] / 2 ? idx : idx - n_[dim]; }
    static C at(const Vec& a, std::size_t id) { return {a[id], a[id + 1]}; }
    static void put(Vec& a, std::size_t id, C z) {
        a[id] = z.real();
        a[id + 1] = z.imag();
    }
    // src/spectral_solver.cpp:400-409
    void zero_mean(Vec3& u) const {
        for_each_spectral([&](std::size_t id, const Mode& m) {
            if (m.mx == 0 && m.my == 0 && m.mz == 0)
                for (auto& c : u) c[id] = c[id + 1] = 0.0;
        });
    }

    std::array<int, 3> n_;
    std::array<double, 3> L_;
    double nu_;
    bool dealias_;
    Grid3d<double> grid_;
    std::size_t np_ = 0;
    std::vector<C> phase_;
    std::vector<unsigned char> keep_;
};

}  // namespace oracle

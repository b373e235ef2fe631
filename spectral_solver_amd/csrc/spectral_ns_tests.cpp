// Known-answer checks of cpp_tests/spectral_ns_tests.cpp:16-51 run against the
// GPU backend: 12^3 grid, L = 2*pi, nu = 0.05, through the legacy
// SpectralNavierStokes API with HipFft3D as FftBackend.
#include "spectral_hip.hpp"

#include <cmath>
#include <iostream>
#include <stdexcept>

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

void require(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(what);
}

// max |backward(forward(u)) - u| < 1e-10 (spectral_ns_tests.cpp:16-32)
void fft_roundtrip(spectral::SpectralNavierStokes& s) {
    auto& fft = s.fft();
    std::vector<double> u(static_cast<std::size_t>(fft.physical_size()));
    const auto& n = fft.shape();
    fft.for_each_physical([&](std::size_t id, int ix, int iy, int) {
        u[id] = std::sin(2.0 * kPi * (ix + 0.5) / n[0]) + 0.25 * std::cos(2.0 * kPi * (iy + 0.5) / n[1]);
    });
    const auto back = fft.backward(fft.forward(u));
    double err = 0.0;
    for (std::size_t i = 0; i < u.size(); ++i) err = std::max(err, std::abs(u[i] - back[i]));
    std::cout << "roundtrip_max_error " << err << "\n";
    require(err < 1.0e-10, "FFT roundtrip error too large");
}

// divergence < 1e-8 after project+truncate, E = 0.125 to 1e-12 (:34-41)
void projection_and_energy(spectral::SpectralNavierStokes& s) {
    const auto u = s.taylor_green();
    auto u_hat = s.to_spectral(u);
    s.project(u_hat);
    s.truncate(u_hat);
    const double div = s.max_spectral_divergence(u_hat);
    const double e = s.kinetic_energy(u);
    std::cout << "projected_divergence " << div << " energy " << std::setprecision(17) << e << "\n";
    require(div < 1.0e-8, "projection did not remove divergence");
    require(std::abs(e - 0.125) < 1.0e-12, "Taylor-Green energy mismatch");
}

// one RK3 step with dt = 1e-3 lowers E and keeps divergence < 1e-7 (:43-51)
void viscous_decay(spectral::SpectralNavierStokes& s) {
    auto u_hat = s.to_spectral(s.taylor_green());
    s.project(u_hat);
    const double e0 = s.kinetic_energy_spectral(u_hat);
    const auto u1 = s.step_rk3(u_hat, 0.0, 1.0e-3);
    const double e1 = s.kinetic_energy_spectral(u1);
    const double div = s.max_spectral_divergence(u1);
    std::cout << "energy_before " << std::setprecision(17) << e0 << " after " << e1 << " divergence " << div << "\n";
    require(e1 < e0, "viscous step did not reduce kinetic energy");
    require(div < 1.0e-7, "RK step left nonzero divergence");
}

}  // namespace

int main() {
    try {
        spectral::SpectralNavierStokes solver({12, 12, 12}, {2.0 * kPi, 2.0 * kPi, 2.0 * kPi}, 0.05);
        fft_roundtrip(solver);
        projection_and_energy(solver);
        viscous_decay(solver);
        std::cout << "all C++ spectral tests passed\n";
    } catch (const std::exception& e) {
        std::cerr << "test failure: " << e.what() << "\n";
        return 1;
    }
    return 0;
}

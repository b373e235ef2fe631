// Known-answer checks of cpp_tests/spectral_ns_tests.cpp:16-51 run against the
// GPU backend: 12^3 grid, L = 2*pi, nu = 0.05, through the legacy
// SpectralNavierStokes API with HipFft3D as FftBackend; then the vorticity
// image writer (src/spectral_solver.cpp:421-486) on a 16x12x8 grid. The image
// goes to argv[1] (default: vorticity_mid_plane.ppm in the working directory).
#include "spectral_hip.hpp"

#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

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

// write_vorticity_ppm on a non-cubic grid: header, size, and every pixel
// within one grey level of the image of the closed-form Taylor-Green
// vorticity ω = (−cos x sin y sin z, −sin x cos y sin z, 2 sin x sin y cos z)
// on the plane iz = n_z/2.
void vorticity_image(const std::string& path) {
    const std::array<int, 3> n = {16, 12, 8};
    spectral::SpectralNavierStokes s(n, {2.0 * kPi, 2.0 * kPi, 2.0 * kPi}, 0.05);
    auto u_hat = s.to_spectral(s.taylor_green());
    spectral::write_vorticity_ppm(path, s, u_hat);

    std::ifstream f(path, std::ios::binary);
    std::string magic;
    int w = 0, h = 0, maxval = 0;
    f >> magic >> w >> h >> maxval;
    f.get();
    require(magic == "P6" && w == n[0] && h == n[1] && maxval == 255, "PPM header mismatch");
    std::vector<unsigned char> px(3 * static_cast<std::size_t>(w) * h);
    f.read(reinterpret_cast<char*>(px.data()), static_cast<std::streamsize>(px.size()));
    require(f.gcount() == static_cast<std::streamsize>(px.size()) && f.peek() == EOF, "PPM size mismatch");

    const double z = 2.0 * kPi * (n[2] / 2 + 0.5) / n[2];
    std::vector<double> mag(static_cast<std::size_t>(w) * h);
    double peak = 0.0;
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            const double x = 2.0 * kPi * (ix + 0.5) / n[0], y = 2.0 * kPi * (iy + 0.5) / n[1];
            const double a = -std::cos(x) * std::sin(y) * std::sin(z), b = -std::sin(x) * std::cos(y) * std::sin(z),
                         c = 2.0 * std::sin(x) * std::sin(y) * std::cos(z);
            mag[static_cast<std::size_t>(iy) * w + ix] = std::sqrt(a * a + b * b + c * c);
            peak = std::max(peak, mag[static_cast<std::size_t>(iy) * w + ix]);
        }
    int worst = 0;
    for (int r = 0; r < h; ++r)
        for (int ix = 0; ix < w; ++ix) {
            const double q = mag[static_cast<std::size_t>(h - 1 - r) * w + ix] / peak;
            const double expect[3] = {255.0 * std::sqrt(q), 255.0 * q * q, 255.0 * (1.0 - q) * 0.25};
            for (int k = 0; k < 3; ++k) {
                const int got = px[3 * (static_cast<std::size_t>(r) * w + ix) + k];
                worst = std::max(worst, std::abs(got - static_cast<int>(expect[k])));
            }
        }
    std::cout << "vorticity_ppm " << w << "x" << h << " max_level_difference " << worst << "\n";
    require(worst <= 1, "vorticity image differs from the closed form");
}

}  // namespace

int main(int argc, char** argv) {
    try {
        spectral::SpectralNavierStokes solver({12, 12, 12}, {2.0 * kPi, 2.0 * kPi, 2.0 * kPi}, 0.05);
        fft_roundtrip(solver);
        projection_and_energy(solver);
        viscous_decay(solver);
        vorticity_image(argc > 1 ? argv[1] : "vorticity_mid_plane.ppm");
        std::cout << "all C++ spectral tests passed\n";
    } catch (const std::exception& e) {
        std::cerr << "test failure: " << e.what() << "\n";
        return 1;
    }
    return 0;
}

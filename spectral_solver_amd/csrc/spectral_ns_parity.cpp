// Drives the legacy value-semantics API (spectral::SpectralNavierStokes over
// HipFft3D, the reference's src/spectral_solver.hpp:48-91 with the GPU
// transforms) on a given physical velocity and writes what it computes, for
// comparison with a CPU restatement of src/spectral_solver.cpp:
//
//   spectral_ns_parity nx ny nz Lx Ly Lz nu dealias steps dt in.bin out.bin
//
// in.bin : 3 * nx*ny*nz doubles, the physical velocity (x fastest).
// out.bin: (steps+1) diagnostics rows (t, E, dissipation, enstrophy, max
//          divergence) after to_spectral, project, truncate and each
//          step_rk3; then nonlinear_rhs of the initial state and the final
//          state, each 3 * spectral_storage_size() doubles in the backend's
//          storage order (for_each_spectral ids).
#include "spectral_hip.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    try {
        if (argc != 13) throw std::runtime_error("usage: spectral_ns_parity nx ny nz Lx Ly Lz nu dealias steps dt in out");
        const std::array<int, 3> shape = {std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3])};
        const std::array<double, 3> lengths = {std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6])};
        const double nu = std::atof(argv[7]);
        const bool dealias = std::atoi(argv[8]) != 0;
        const int steps = std::atoi(argv[9]);
        const double dt = std::atof(argv[10]);
        spectral::SpectralNavierStokes s(shape, lengths, nu, spectral::kCommSelf, dealias);
        const std::size_t np = static_cast<std::size_t>(s.fft().physical_size());
        spectral::VectorPhysical u;
        std::ifstream in(argv[11], std::ios::binary);
        for (auto& c : u) {
            c.resize(np);
            in.read(reinterpret_cast<char*>(c.data()), static_cast<std::streamsize>(np * sizeof(double)));
        }
        if (!in) throw std::runtime_error(std::string("cannot read ") + argv[11]);
        auto u_hat = s.to_spectral(u);
        s.project(u_hat);
        s.truncate(u_hat);
        std::vector<double> rows;
        auto record = [&](double t) {
            const auto d = s.diagnostics(u_hat, t);
            rows.insert(rows.end(), {d.t, d.energy, d.dissipation, d.enstrophy, d.max_spectral_divergence});
        };
        record(0.0);
        const auto rhs0 = s.nonlinear_rhs(u_hat);
        double t = 0.0;
        for (int k = 0; k < steps; ++k) {
            u_hat = s.step_rk3(u_hat, t, dt);
            t += dt;
            record(t);
        }
        std::ofstream out(argv[12], std::ios::binary);
        out.write(reinterpret_cast<const char*>(rows.data()), static_cast<std::streamsize>(rows.size() * sizeof(double)));
        const spectral::VectorSpectral* fields[2] = {&rhs0, &u_hat};
        for (const auto* f : fields)
            for (const auto& c : *f)
                out.write(reinterpret_cast<const char*>(c.data()), static_cast<std::streamsize>(c.size() * sizeof(double)));
        if (!out) throw std::runtime_error(std::string("cannot write ") + argv[12]);
        std::cout << "spectral_ns_parity " << shape[0] << "x" << shape[1] << "x" << shape[2] << " dealias=" << dealias
                  << " steps=" << steps << " E=" << rows[rows.size() - 4] << "\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}

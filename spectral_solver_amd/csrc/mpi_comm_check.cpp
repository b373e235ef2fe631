// The FFT backend class with a real MPI communicator, as the reference's code
// uses it: <mpi.h> comes first (src/fftw_fft.hpp:8-9), the solver is
// constructed with an MPI_Comm (src/spectral_solver.hpp:50-54) and fft().comm()
// goes to MPI_Allreduce (src/spectral_solver.cpp:317). Built only where the
// build finds <mpi.h>; prints the Taylor-Green energy reduced over
// fft().comm() and checks it against 1/8.
#include <mpi.h>

#include "spectral_hip.hpp"

#include <cmath>
#include <cstdio>
#include <iostream>
#include <type_traits>

static_assert(std::is_same_v<spectral::Comm, MPI_Comm>, "with <mpi.h> the backend's communicator is MPI_Comm");
static_assert(std::is_same_v<decltype(std::declval<spectral::HipFft3D&>().comm()), MPI_Comm>,
              "comm() returns MPI_Comm");

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rc = 0;
    try {
        constexpr double two_pi = 6.283185307179586476925286766559005768;
        spectral::SpectralNavierStokes s({16, 12, 8}, {two_pi, two_pi, two_pi}, 0.05, MPI_COMM_SELF);
        const auto u = s.taylor_green();
        double local = 0.0;
        for (std::size_t i = 0; i < u[0].size(); ++i)
            local += 0.5 * (u[0][i] * u[0][i] + u[1][i] * u[1][i] + u[2][i] * u[2][i]);
        double global = 0.0;
        MPI_Allreduce(&local, &global, 1, MPI_DOUBLE, MPI_SUM, s.fft().comm());
        const double e = global / s.fft().global_physical_size();
        const double e_solver = s.kinetic_energy(u);
        int size = 0;
        MPI_Comm_size(s.fft().comm(), &size);
        std::printf("mpi_comm_check comm_size=%d energy=%.17g solver_energy=%.17g\n", size, e, e_solver);
        if (size != 1 || std::abs(e - 0.125) > 1e-14 || std::abs(e_solver - 0.125) > 1e-14) rc = 1;
    } catch (const std::exception& ex) {
        std::cerr << "error: " << ex.what() << "\n";
        rc = 1;
    }
    MPI_Finalize();
    return rc;
}

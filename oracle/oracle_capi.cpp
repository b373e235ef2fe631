// C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY; see oracle.hpp).
// Loaded with ctypes by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg. Fields cross the boundary as interleaved doubles in the
// reference's C2C layout (id = ix + n*(iy + n*iz), x fastest).
#include "oracle.hpp"
#include "legacy.hpp"

#include <cstring>
#include <omp.h>
#include <memory>
#include <string>

namespace {
thread_local std::string g_error;

struct Handle {
    int single = 0;
    std::unique_ptr<oracle::TaylorGreen<double>> d;
    std::unique_ptr<oracle::TaylorGreen<float>> f;
};

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

void fill_row(const oracle::Diagnostics& d, double* out) {
    const double v[10] = {d.t, d.kinetic_energy, d.dissipation, d.enstrophy, d.max_spectral_divergence,
                          d.max_velocity, d.dt, d.cfl, d.main_dt, d.main_cfl};
    std::memcpy(out, v, sizeof(v));
}

void fill_counters(const oracle::Counters& c, long long* out) {
    out[0] = c.forward_calls;
    out[1] = c.backward_calls;
    out[2] = c.step_calls;
    out[3] = c.rhs_calls;
    out[4] = c.diagnostics_calls;
    out[5] = c.max_velocity_calls;
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_error.c_str(); }

// Unnormalised 3D C2C transform of an interleaved complex [nz][ny][nx] array.
// sign = -1 forward, +1 backward. `single` != 0 runs it in float.
int orc_fft3d(int nx, int ny, int nz, const double* in, double* out, int sign, int single) {
    return guarded([&] {
        const std::size_t total = static_cast<std::size_t>(nx) * ny * nz;
        if (single) {
            std::vector<oracle::Cx<float>> a(total);
            for (std::size_t i = 0; i < total; ++i) a[i] = {static_cast<float>(in[2 * i]), static_cast<float>(in[2 * i + 1])};
            oracle::Grid3d<float>(nx, ny, nz).exec(a.data(), sign);
            for (std::size_t i = 0; i < total; ++i) {
                out[2 * i] = a[i].x;
                out[2 * i + 1] = a[i].y;
            }
        } else {
            std::vector<oracle::Cx<double>> a(total);
            std::memcpy(a.data(), in, total * sizeof(a[0]));
            oracle::Grid3d<double>(nx, ny, nz).exec(a.data(), sign);
            std::memcpy(out, a.data(), total * sizeof(a[0]));
        }
    });
}

int orc_solver_create(int n, double reynolds, int single, int initialise, void** out) {
    return guarded([&] {
        auto h = std::make_unique<Handle>();
        h->single = single;
        if (single)
            h->f = std::make_unique<oracle::TaylorGreen<float>>(n, reynolds, initialise != 0);
        else
            h->d = std::make_unique<oracle::TaylorGreen<double>>(n, reynolds, initialise != 0);
        *out = h.release();
    });
}

void orc_solver_destroy(void* h) { delete static_cast<Handle*>(h); }

int orc_solver_step(void* hp, double dt) {
    auto* h = static_cast<Handle*>(hp);
    return guarded([&] {
        if (h->single)
            h->f->step(dt);
        else
            h->d->step(dt);
    });
}

int orc_solver_diagnostics(void* hp, double t, double* out10) {
    auto* h = static_cast<Handle*>(hp);
    return guarded([&] { fill_row(h->single ? h->f->diagnostics(t) : h->d->diagnostics(t), out10); });
}

int orc_solver_max_velocity(void* hp, double* out) {
    auto* h = static_cast<Handle*>(hp);
    return guarded([&] { *out = h->single ? h->f->max_velocity() : h->d->max_velocity(); });
}

int orc_solver_counters(void* hp, long long* out6) {
    auto* h = static_cast<Handle*>(hp);
    return guarded([&] { fill_counters(h->single ? h->f->counters() : h->d->counters(), out6); });
}

// Runs only the first RK3 stage (RHS + update + projection) and returns its
// wall time in seconds through *seconds (a bounded CPU-baseline sample).
int orc_solver_first_stage(void* hp, double dt, double* seconds) {
    auto* h = static_cast<Handle*>(hp);
    return guarded([&] {
        const auto t0 = std::chrono::steady_clock::now();
        if (h->single)
            h->f->stage(dt, 0);
        else
            h->d->stage(dt, 0);
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

double orc_solver_step_seconds(void* hp) {
    auto* h = static_cast<Handle*>(hp);
    return h->single ? h->f->step_seconds() : h->d->step_seconds();
}

// Copy component c of u_hat (interleaved doubles, n^3 complex) out / in.
int orc_solver_get_uhat(void* hp, int c, double* out) {
    auto* h = static_cast<Handle*>(hp);
    return guarded([&] {
        if (c < 0 || c > 2) throw std::runtime_error("component out of range");
        if (h->single) {
            const auto& u = h->f->u_hat(c);
            for (std::size_t i = 0; i < u.size(); ++i) {
                out[2 * i] = u[i].x;
                out[2 * i + 1] = u[i].y;
            }
        } else {
            const auto& u = h->d->u_hat(c);
            std::memcpy(out, u.data(), u.size() * sizeof(u[0]));
        }
    });
}

int orc_solver_set_uhat(void* hp, int c, const double* in) {
    auto* h = static_cast<Handle*>(hp);
    return guarded([&] {
        if (c < 0 || c > 2) throw std::runtime_error("component out of range");
        if (h->single) {
            auto& u = h->f->u_hat(c);
            for (std::size_t i = 0; i < u.size(); ++i) u[i] = {static_cast<float>(in[2 * i]), static_cast<float>(in[2 * i + 1])};
        } else {
            auto& u = h->d->u_hat(c);
            std::memcpy(u.data(), in, u.size() * sizeof(u[0]));
        }
    });
}

// Metadata tables of an n^3 solver (impl.hpp:767-794): signed modes, keep
// mask and half-cell phase (interleaved doubles), C2C layout.
int orc_mode_tables(int n, int* mx, int* my, int* mz, unsigned char* keep, double* phase) {
    return guarded([&] {
        oracle::TaylorGreen<double> s(n, 1.0, false);
        const std::size_t total = s.total();
        std::memcpy(mx, s.mode_x().data(), total * sizeof(int));
        std::memcpy(my, s.mode_y().data(), total * sizeof(int));
        std::memcpy(mz, s.mode_z().data(), total * sizeof(int));
        std::memcpy(keep, s.keep().data(), total);
        std::memcpy(phase, s.phase().data(), total * 2 * sizeof(double));
    });
}

// Dealiased product spectrum of two C2C spectra on an n^3 grid (FP64).
int orc_dealiased_product(int n, const double* a_hat, const double* b_hat, double* out) {
    return guarded([&] {
        oracle::TaylorGreen<double> s(n, 1.0, false);
        const std::size_t total = s.total();
        std::vector<oracle::Cx<double>> a(total), b(total), o;
        std::memcpy(a.data(), a_hat, total * sizeof(a[0]));
        std::memcpy(b.data(), b_hat, total * sizeof(b[0]));
        s.dealiased_product(a, b, o);
        std::memcpy(out, o.data(), total * sizeof(o[0]));
    });
}

// Driver restatement; writes <output_dir>/taylor_green_diagnostics.csv.
int orc_run_taylor_green(int n, int steps, int max_steps, int output_every, int print_every, int adaptive,
                         double dt, double cfl, double t_end, double output_dt, double reynolds, int single,
                         const char* output_dir, double* final_row10, long long* counters6, int* steps_taken) {
    return guarded([&] {
        oracle::RunConfig cfg;
        cfg.n = n;
        cfg.steps = steps;
        cfg.max_steps = max_steps;
        cfg.output_every = output_every < 1 ? 1 : output_every;
        cfg.print_every = print_every < 1 ? 1 : print_every;
        cfg.adaptive_cfl = adaptive != 0;
        cfg.dt = dt;
        cfg.cfl = cfl;
        cfg.t_end = t_end;
        cfg.output_dt = output_dt;
        cfg.reynolds = reynolds;
        cfg.output_dir = output_dir;
        oracle::Counters c;
        int st = 0;
        const auto row = single ? oracle::run_taylor_green<float>(cfg, &c, &st)
                                : oracle::run_taylor_green<double>(cfg, &c, &st);
        fill_row(row, final_row10);
        fill_counters(c, counters6);
        *steps_taken = st;
    });
}

// Legacy solver restatement (legacy.hpp): physical velocity phys_in
// (3 * nx*ny*nz doubles, x fastest; NULL = taylor_green(1, 1)) -> to_spectral,
// project, truncate; then `steps` RK3 steps of dt. Writes the diagnostics rows
// (t, E, dissipation, enstrophy, max divergence) after 0..steps steps to
// diag_out ((steps+1)*5 doubles), the nonlinear term of the initial state to
// rhs0_out and the final state to state_out (each 3 * 2*nx*ny*nz doubles, C2C
// interleaved; either may be NULL).
int orc_legacy_run(const int shape[3], const double lengths[3], double nu, int dealias, int steps, double dt,
                   const double* phys_in, double* diag_out, double* rhs0_out, double* state_out) {
    return guarded([&] {
        oracle::LegacyNS s({shape[0], shape[1], shape[2]}, {lengths[0], lengths[1], lengths[2]}, nu, dealias != 0);
        const std::size_t np = s.physical_size(), ns = s.spectral_storage_size();
        oracle::LegacyNS::Vec3 u;
        if (phys_in) {
            for (int c = 0; c < 3; ++c) u[c].assign(phys_in + c * np, phys_in + (c + 1) * np);
        } else {
            u = s.taylor_green();
        }
        auto u_hat = s.to_spectral(u);
        s.project(u_hat);
        s.truncate(u_hat);
        if (rhs0_out) {
            const auto r = s.nonlinear_rhs(u_hat);
            for (int c = 0; c < 3; ++c) std::memcpy(rhs0_out + c * ns, r[c].data(), ns * sizeof(double));
        }
        double t = 0.0;
        for (int k = 0; k <= steps; ++k) {
            if (k > 0) {
                u_hat = s.step_rk3(u_hat, dt);
                t += dt;
            }
            if (diag_out) {
                const auto row = s.diagnostics(u_hat, t);
                const double v[5] = {row.t, row.energy, row.dissipation, row.enstrophy, row.max_spectral_divergence};
                std::memcpy(diag_out + 5 * k, v, sizeof(v));
            }
        }
        if (state_out)
            for (int c = 0; c < 3; ++c) std::memcpy(state_out + c * ns, u_hat[c].data(), ns * sizeof(double));
    });
}

// Threads of the oracle's OpenMP loops (1 = the reference's serial build).
int orc_set_threads(int threads) {
    return guarded([&] {
        if (threads < 1) throw std::runtime_error("threads must be positive");
        omp_set_num_threads(threads);
    });
}

}  // extern "C"

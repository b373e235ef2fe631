// CPU oracle for the Taylor-Green RK3 hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it. The GPU library must never call into it.
//
// It restates, on the host and in plain C++, the algorithm of the reference's
// AVA solver with its serial FFTW backend:
//   - storage: every field is a complex N^3 array, x fastest
//     (src/ava_taylor_green_impl.hpp:103-106, 742-765, 340-412);
//   - transforms: C2C, forward unnormalised with sign e^{-i}; backward sign
//     e^{+i} followed by a multiply with Real{1}/N^3
//     (src/ava_taylor_green_impl.hpp:195-213, 475-478, 547-561);
//   - numerics: src/ava_taylor_green_impl.hpp:767-1179 (metadata, initial
//     condition, projection, RHS, RK3, reductions, diagnostics);
//   - driver: src/ava_taylor_green_impl.hpp:1310-1410.
// FFTW 3.3.10 itself is absent here; its role is taken by the mixed-radix
// Stockham transform below (exact twiddles from long double sincos). The
// reference's reduction grouping (AVA_BLOCK_SIZE, unpinned) is modelled as
// 256-element blocks reduced in Real, summed on the host in double.
//
// Pinning (tests/test_oracle.py): brute-force DFT on small grids, single-mode
// closed forms, Parseval, the reference's own known answers
// (cpp_tests/spectral_ns_tests.cpp:16-51, spectral_Re1600_512.dat row 1),
// the exact RK3 amplification polynomial of a pure shear mode, and the
// published 64^3 FP64 t=20 end state and FFT call counts (Benchmark.md:172-177,
// 584-591).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

constexpr double kPi = 3.141592653589793238462643383279502884;

template <class Real>
struct Cx {
    Real x;
    Real y;
};

template <class Real>
inline Cx<Real> cx_add(Cx<Real> a, Cx<Real> b) { return {a.x + b.x, a.y + b.y}; }
template <class Real>
inline Cx<Real> cx_sub(Cx<Real> a, Cx<Real> b) { return {a.x - b.x, a.y - b.y}; }
template <class Real>
inline Cx<Real> cx_mul(Cx<Real> a, Cx<Real> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class Real>
inline Cx<Real> cx_scale(Cx<Real> a, Real s) { return {s * a.x, s * a.y}; }

// ---------------------------------------------------------------------------
// 1D mixed-radix Stockham transform (radix 4, 2, 3, 5, generic odd factor).
// Stage s with radix R and Ns = product of earlier radices maps
//   x_r = a[j + r*N/R] * w_{Ns*R}^{r*(j mod Ns)},  y = DFT_R(x),
//   b[(j/Ns)*Ns*R + (j mod Ns) + r*Ns] = y_r.
// ---------------------------------------------------------------------------
template <class Real>
class Line1d {
public:
    explicit Line1d(int n) : n_(n) {
        if (n < 1) throw std::runtime_error("transform length must be positive");
        int m = n;
        for (int p : {4, 2, 3, 5}) {
            while (m % p == 0) {
                radix_.push_back(p);
                m /= p;
            }
        }
        for (int p = 7; m > 1; p += 2) {
            while (m % p == 0) {
                if (p > 64) throw std::runtime_error("transform length has a prime factor above 64");
                radix_.push_back(p);
                m /= p;
            }
        }
        int ns = 1;
        for (int r : radix_) {
            stage_offset_.push_back(static_cast<int>(tw_.size()));
            for (int k = 0; k < ns; ++k) {
                for (int q = 0; q < r; ++q) {
                    const long double ang =
                        -2.0L * static_cast<long double>(kPi) * static_cast<long double>(q * k) /
                        static_cast<long double>(ns * r);
                    tw_.push_back({static_cast<Real>(std::cos(ang)), static_cast<Real>(std::sin(ang))});
                }
            }
            ns *= r;
        }
        twb_.resize(tw_.size());
        for (std::size_t i = 0; i < tw_.size(); ++i) twb_[i] = {tw_[i].x, -tw_[i].y};
        for (int r : radix_) {
            std::vector<Cx<Real>> roots(r);
            for (int q = 0; q < r; ++q) {
                const long double ang = -2.0L * static_cast<long double>(kPi) * q / r;
                roots[q] = {static_cast<Real>(std::cos(ang)), static_cast<Real>(std::sin(ang))};
            }
            roots_.push_back(roots);
        }
    }

    int size() const { return n_; }

    // sign = -1 forward, +1 backward (unnormalised). `a` holds the input and
    // receives the output; `work` must hold n elements.
    void exec(Cx<Real>* a, Cx<Real>* work, int sign) const { exec_batch(a, work, 1, 1, sign); }

    // The same transform of `count` lines stored element-major with stride B
    // between elements (line b at a[e*B + b], count <= B): every element goes
    // through exactly the operations of exec, lines side by side, so the
    // results are those of exec line by line. `work` must hold n*B elements.
    void exec_batch(Cx<Real>* a, Cx<Real>* work, int B, int count, int sign) const {
        Cx<Real>* src = a;
        Cx<Real>* dst = work;
        int ns = 1;
        const std::vector<Cx<Real>>& tws = sign > 0 ? twb_ : tw_;
        for (std::size_t s = 0; s < radix_.size(); ++s) {
            const int r = radix_[s];
            const int m = n_ / r;
            const Cx<Real>* tw = tws.data() + stage_offset_[s];
            for (int j = 0; j < m; ++j) {
                const int k = j % ns;
                const Cx<Real>* w = tw + k * r;
                const int base = (j / ns) * ns * r + k;
                const Cx<Real>* in = src + static_cast<std::size_t>(j) * B;
                Cx<Real>* out = dst + static_cast<std::size_t>(base) * B;
                const std::size_t is = static_cast<std::size_t>(m) * B, os = static_cast<std::size_t>(ns) * B;
                if (r == 4) {
                    for (int b = 0; b < count; ++b) {
                        const Cx<Real> x0 = in[b], x1 = cx_mul(in[is + b], w[1]), x2 = cx_mul(in[2 * is + b], w[2]),
                                       x3 = cx_mul(in[3 * is + b], w[3]);
                        const Cx<Real> p = cx_add(x0, x2), q = cx_sub(x0, x2);
                        const Cx<Real> c = cx_add(x1, x3), d = cx_sub(x1, x3);
                        const Cx<Real> dr = sign < 0 ? Cx<Real>{d.y, -d.x} : Cx<Real>{-d.y, d.x};
                        out[b] = cx_add(p, c);
                        out[2 * os + b] = cx_sub(p, c);
                        out[os + b] = cx_add(q, dr);
                        out[3 * os + b] = cx_sub(q, dr);
                    }
                } else if (r == 2) {
                    for (int b = 0; b < count; ++b) {
                        const Cx<Real> x0 = in[b], x1 = cx_mul(in[is + b], w[1]);
                        out[b] = cx_add(x0, x1);
                        out[os + b] = cx_sub(x0, x1);
                    }
                } else {
                    Cx<Real> x[64];
                    for (int b = 0; b < count; ++b) {
                        x[0] = in[b];
                        for (int q = 1; q < r; ++q) x[q] = cx_mul(in[q * is + b], w[q]);
                        butterfly(x, r, s, sign);
                        for (int q = 0; q < r; ++q) out[q * os + b] = x[q];
                    }
                }
            }
            std::swap(src, dst);
            ns *= r;
        }
        if (src != a) {
            for (int e = 0; e < n_; ++e)
                for (int b = 0; b < count; ++b) a[static_cast<std::size_t>(e) * B + b] = src[static_cast<std::size_t>(e) * B + b];
        }
    }

private:
    void butterfly(Cx<Real>* x, int r, std::size_t s, int sign) const {
        if (r == 2) {
            const Cx<Real> a = x[0], b = x[1];
            x[0] = cx_add(a, b);
            x[1] = cx_sub(a, b);
        } else if (r == 4) {
            const Cx<Real> a = cx_add(x[0], x[2]), b = cx_sub(x[0], x[2]);
            const Cx<Real> c = cx_add(x[1], x[3]), d = cx_sub(x[1], x[3]);
            // rotate d by -i (forward) or +i (backward)
            const Cx<Real> dr = sign < 0 ? Cx<Real>{d.y, -d.x} : Cx<Real>{-d.y, d.x};
            x[0] = cx_add(a, c);
            x[2] = cx_sub(a, c);
            x[1] = cx_add(b, dr);
            x[3] = cx_sub(b, dr);
        } else {
            Cx<Real> y[64];
            const auto& roots = roots_[s];
            for (int k = 0; k < r; ++k) {
                Cx<Real> acc = x[0];
                for (int q = 1; q < r; ++q) {
                    Cx<Real> w = roots[(q * k) % r];
                    if (sign > 0) w.y = -w.y;
                    acc = cx_add(acc, cx_mul(x[q], w));
                }
                y[k] = acc;
            }
            for (int k = 0; k < r; ++k) x[k] = y[k];
        }
    }

    int n_;
    std::vector<int> radix_;
    std::vector<int> stage_offset_;
    std::vector<Cx<Real>> tw_, twb_;  // forward twiddles and their conjugates
    std::vector<std::vector<Cx<Real>>> roots_;
};

// 3D C2C transform of a [nz][ny][nx] array (x fastest), unnormalised.
template <class Real>
class Grid3d {
public:
    Grid3d(int nx, int ny, int nz) : nx_(nx), ny_(ny), nz_(nz), fx_(nx), fy_(ny), fz_(nz) {}

    std::size_t size() const { return static_cast<std::size_t>(nx_) * ny_ * nz_; }

    // Lines of each axis go through Line1d::exec_batch kBlock at a time,
    // gathered element-major into a per-thread buffer.
    void exec(Cx<Real>* a, int sign) const {
        const int nx = nx_, ny = ny_, nz = nz_;
        constexpr int kBlock = 16;
        // x lines: contiguous rows; a block is kBlock consecutive rows
        const long rows = static_cast<long>(ny) * nz;
#pragma omp parallel
        {
            std::vector<Cx<Real>> buf(static_cast<std::size_t>(nx) * kBlock), work(buf.size());
#pragma omp for schedule(static)
            for (long task = 0; task < (rows + kBlock - 1) / kBlock; ++task) {
                const long r0 = task * kBlock;
                const int bw = static_cast<int>(std::min<long>(kBlock, rows - r0));
                Cx<Real>* base = a + static_cast<std::size_t>(r0) * nx;
                for (int b = 0; b < bw; ++b)
                    for (int x = 0; x < nx; ++x) buf[static_cast<std::size_t>(x) * kBlock + b] = base[static_cast<std::size_t>(b) * nx + x];
                fx_.exec_batch(buf.data(), work.data(), kBlock, bw, sign);
                for (int b = 0; b < bw; ++b)
                    for (int x = 0; x < nx; ++x) base[static_cast<std::size_t>(b) * nx + x] = buf[static_cast<std::size_t>(x) * kBlock + b];
            }
        }
        // y lines: stride nx; a block is kBlock adjacent x of one plane
#pragma omp parallel
        {
            std::vector<Cx<Real>> buf(static_cast<std::size_t>(ny) * kBlock), work(buf.size());
#pragma omp for schedule(static)
            for (long task = 0; task < static_cast<long>(nz) * ((nx + kBlock - 1) / kBlock); ++task) {
                const int z = static_cast<int>(task / ((nx + kBlock - 1) / kBlock));
                const int x0 = static_cast<int>(task % ((nx + kBlock - 1) / kBlock)) * kBlock;
                const int bw = std::min(kBlock, nx - x0);
                Cx<Real>* plane = a + static_cast<std::size_t>(z) * nx * ny;
                for (int y = 0; y < ny; ++y)
                    for (int b = 0; b < bw; ++b)
                        buf[static_cast<std::size_t>(y) * kBlock + b] = plane[static_cast<std::size_t>(y) * nx + x0 + b];
                fy_.exec_batch(buf.data(), work.data(), kBlock, bw, sign);
                for (int y = 0; y < ny; ++y)
                    for (int b = 0; b < bw; ++b)
                        plane[static_cast<std::size_t>(y) * nx + x0 + b] = buf[static_cast<std::size_t>(y) * kBlock + b];
            }
        }
        // z lines: stride nx*ny; a block is kBlock adjacent points of a plane
        const std::size_t plane = static_cast<std::size_t>(nx) * ny;
#pragma omp parallel
        {
            std::vector<Cx<Real>> buf(static_cast<std::size_t>(nz) * kBlock), work(buf.size());
#pragma omp for schedule(static)
            for (long task = 0; task < static_cast<long>((plane + kBlock - 1) / kBlock); ++task) {
                const std::size_t p0 = static_cast<std::size_t>(task) * kBlock;
                const int bw = static_cast<int>(std::min<std::size_t>(kBlock, plane - p0));
                for (int z = 0; z < nz; ++z)
                    for (int b = 0; b < bw; ++b) buf[static_cast<std::size_t>(z) * kBlock + b] = a[z * plane + p0 + b];
                fz_.exec_batch(buf.data(), work.data(), kBlock, bw, sign);
                for (int z = 0; z < nz; ++z)
                    for (int b = 0; b < bw; ++b) a[z * plane + p0 + b] = buf[static_cast<std::size_t>(z) * kBlock + b];
            }
        }
    }

private:
    int nx_, ny_, nz_;
    Line1d<Real> fx_, fy_, fz_;
};

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

struct Counters {
    long long forward_calls = 0;
    long long backward_calls = 0;
    long long step_calls = 0;
    long long rhs_calls = 0;
    long long diagnostics_calls = 0;
    long long max_velocity_calls = 0;
};

// signed mode convention: Nyquist index maps to +n/2
// (src/ava_taylor_green_impl.hpp:123-125).
inline int signed_index(int i, int n) { return i <= n / 2 ? i : i - n; }

// spherical truncation of radius n/2 evaluated in double
// (src/ava_taylor_green_impl.hpp:131-134).
inline bool keep_sphere(int mx, int my, int mz, int n) {
    const double h = 0.5 * static_cast<double>(n);
    return (mx * mx + my * my + mz * mz) / (h * h) <= 1.0;
}

// Host restatement of TaylorGreenSolver::Impl (src/ava_taylor_green_impl.hpp:658-1179),
// C2C storage, one complex N^3 array per field.
template <class Real>
class TaylorGreen {
public:
    using C = Cx<Real>;

    TaylorGreen(int n, double reynolds, bool initialise = true)
        : n_(n), nu_(static_cast<Real>(1.0 / reynolds)), grid_(n > 0 ? n : 1, n > 0 ? n : 1, n > 0 ? n : 1) {
        if (n <= 0) throw std::runtime_error("grid size must be positive");
        if (reynolds <= 0.0) throw std::runtime_error("Reynolds number must be positive");
        total_ = static_cast<std::size_t>(n) * n * n;
        for (auto* f : {&u_hat_, &rhs_, &g_, &phys_, &shift_hat_, &shift_phys_}) {
            for (auto& c : *f) c.assign(total_, C{0, 0});
        }
        product_.assign(total_, C{0, 0});
        product_shift_.assign(total_, C{0, 0});
        product_hat_.assign(total_, C{0, 0});
        product_shift_hat_.assign(total_, C{0, 0});
        build_tables();
        if (initialise) initial_condition();
    }

    int n() const { return n_; }
    std::size_t total() const { return total_; }
    const Counters& counters() const { return counters_; }
    std::vector<C>& u_hat(int c) { return u_hat_[c]; }

    // --- transforms (FftPlan::forward/backward + scale_inverse) -------------
    void forward(std::vector<C>& in, std::vector<C>& out) {
        ++counters_.forward_calls;
        out = in;
        grid_.exec(out.data(), -1);
    }
    void backward(std::vector<C>& in, std::vector<C>& out) {
        ++counters_.backward_calls;
        out = in;
        grid_.exec(out.data(), +1);
        const Real inv = Real{1} / static_cast<Real>(total_);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < static_cast<long>(total_); ++i) out[i] = cx_scale(out[i], inv);
    }

    // --- initial condition (impl.hpp:796-817): cell centres, Real arithmetic
    void initial_condition() {
        const int n = n_;
        const Real rpi = static_cast<Real>(kPi);
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            const int ix = static_cast<int>(id % n);
            const int iy = static_cast<int>((id / n) % n);
            const int iz = static_cast<int>(id / (static_cast<long>(n) * n));
            const Real x = (static_cast<Real>(ix) + Real{0.5}) * Real{2} * rpi / static_cast<Real>(n);
            const Real y = (static_cast<Real>(iy) + Real{0.5}) * Real{2} * rpi / static_cast<Real>(n);
            const Real z = (static_cast<Real>(iz) + Real{0.5}) * Real{2} * rpi / static_cast<Real>(n);
            phys_[0][id] = {std::sin(x) * std::cos(y) * std::cos(z), Real{0}};
            phys_[1][id] = {-std::cos(x) * std::sin(y) * std::cos(z), Real{0}};
            phys_[2][id] = {Real{0}, Real{0}};
        }
        for (int c = 0; c < 3; ++c) forward(phys_[c], u_hat_[c]);
        project_truncate(u_hat_);
    }

    // --- Leray projection + truncation in place (impl.hpp:819-845) ----------
    void project_truncate(std::vector<C> (&f)[3]) {
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            const int ix = mx_[id], iy = my_[id], iz = mz_[id];
            const Real k2 = static_cast<Real>(ix * ix + iy * iy + iz * iz);
            if (k2 > Real{0} && keep_[id]) {
                const C ux = f[0][id], uy = f[1][id], uz = f[2][id];
                const C dot = cx_add(cx_add(cx_scale(ux, static_cast<Real>(ix)), cx_scale(uy, static_cast<Real>(iy))),
                                     cx_scale(uz, static_cast<Real>(iz)));
                const C fac = cx_scale(dot, Real{1} / k2);
                f[0][id] = cx_sub(ux, cx_scale(fac, static_cast<Real>(ix)));
                f[1][id] = cx_sub(uy, cx_scale(fac, static_cast<Real>(iy)));
                f[2][id] = cx_sub(uz, cx_scale(fac, static_cast<Real>(iz)));
            } else {
                f[0][id] = f[1][id] = f[2][id] = C{0, 0};
            }
        }
    }

    // --- RHS (impl.hpp:847-977) ---------------------------------------------
    void compute_rhs() {
        for (int c = 0; c < 3; ++c) backward(u_hat_[c], phys_[c]);
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            for (int c = 0; c < 3; ++c) {
                rhs_[c][id] = C{0, 0};
                shift_hat_[c][id] = cx_mul(u_hat_[c][id], phase_[id]);
            }
        }
        for (int c = 0; c < 3; ++c) backward(shift_hat_[c], shift_phys_[c]);
        static constexpr int pairs[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        for (const auto& p : pairs) accumulate_product(p[0], p[1]);
        add_viscous_project();
        ++counters_.rhs_calls;
    }

    void accumulate_product(int a, int b) {
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            product_[id] = {phys_[a][id].x * phys_[b][id].x, Real{0}};
            product_shift_[id] = {shift_phys_[a][id].x * shift_phys_[b][id].x, Real{0}};
        }
        forward(product_, product_hat_);
        forward(product_shift_, product_shift_hat_);
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            const int kk[3] = {mx_[id], my_[id], mz_[id]};
            const C ph = phase_[id];
            C v = cx_scale(cx_add(product_hat_[id], cx_mul(product_shift_hat_[id], C{ph.x, -ph.y})), Real{0.5});
            if (!keep_[id]) v = C{0, 0};
            // -i k v
            const Real kb = static_cast<Real>(kk[b]);
            rhs_[a][id] = cx_add(rhs_[a][id], C{kb * v.y, -kb * v.x});
            if (a != b) {
                const Real ka = static_cast<Real>(kk[a]);
                rhs_[b][id] = cx_add(rhs_[b][id], C{ka * v.y, -ka * v.x});
            }
        }
    }

    void add_viscous_project() {
        const Real visc = nu_;
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            const int ix = mx_[id], iy = my_[id], iz = mz_[id];
            const Real k2 = static_cast<Real>(ix * ix + iy * iy + iz * iz);
            C r[3];
            for (int c = 0; c < 3; ++c) r[c] = cx_sub(rhs_[c][id], cx_scale(u_hat_[c][id], visc * k2));
            if (k2 > Real{0} && keep_[id]) {
                const C dot = cx_add(cx_add(cx_scale(r[0], static_cast<Real>(ix)), cx_scale(r[1], static_cast<Real>(iy))),
                                     cx_scale(r[2], static_cast<Real>(iz)));
                const C fac = cx_scale(dot, Real{1} / k2);
                rhs_[0][id] = cx_sub(r[0], cx_scale(fac, static_cast<Real>(ix)));
                rhs_[1][id] = cx_sub(r[1], cx_scale(fac, static_cast<Real>(iy)));
                rhs_[2][id] = cx_sub(r[2], cx_scale(fac, static_cast<Real>(iz)));
            } else {
                rhs_[0][id] = rhs_[1][id] = rhs_[2][id] = C{0, 0};
            }
        }
    }

    // --- Williamson low-storage RK3 (impl.hpp:979-1045) ---------------------
    void step(double dt) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < 3; ++s) stage(dt, s);
        ++counters_.step_calls;
        step_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // One of the three stages: RHS, g = alpha_s g + r, u += beta_s g, project.
    void stage(double dt, int s) {
        const Real rdt = static_cast<Real>(dt);
        const Real alpha[3] = {Real{0}, -Real{5} / Real{9}, -Real{153} / Real{128}};
        const Real beta[3] = {rdt / Real{3}, Real{15} * rdt / Real{16}, Real{8} * rdt / Real{15}};
        compute_rhs();
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            for (int c = 0; c < 3; ++c) {
                g_[c][id] = s == 0 ? rhs_[c][id] : cx_add(cx_scale(g_[c][id], alpha[s]), rhs_[c][id]);
                u_hat_[c][id] = cx_add(u_hat_[c][id], cx_scale(g_[c][id], beta[s]));
            }
        }
        project_truncate(u_hat_);
    }

    double step_seconds() const { return step_seconds_; }

    // The dealiased product of two spectral fields exactly as one
    // accumulate_product call forms it (impl.hpp:866-936): physical and
    // half-cell-shifted products, two forward transforms, 0.5*(P + Ps*conj(phi)),
    // truncation mask.
    void dealiased_product(const std::vector<C>& a_hat, const std::vector<C>& b_hat, std::vector<C>& out) {
        std::vector<C> a = a_hat, b = b_hat, as(total_), bs(total_), pa, pb, psa, psb;
        for (std::size_t id = 0; id < total_; ++id) {
            as[id] = cx_mul(a_hat[id], phase_[id]);
            bs[id] = cx_mul(b_hat[id], phase_[id]);
        }
        backward(a, pa);
        backward(b, pb);
        backward(as, psa);
        backward(bs, psb);
        std::vector<C> p(total_), ps(total_), ph, phs;
        for (std::size_t id = 0; id < total_; ++id) {
            p[id] = {pa[id].x * pb[id].x, Real{0}};
            ps[id] = {psa[id].x * psb[id].x, Real{0}};
        }
        forward(p, ph);
        forward(ps, phs);
        out.assign(total_, C{0, 0});
        for (std::size_t id = 0; id < total_; ++id) {
            const C q = phase_[id];
            C v = cx_scale(cx_add(ph[id], cx_mul(phs[id], C{q.x, -q.y})), Real{0.5});
            if (!keep_[id]) v = C{0, 0};
            out[id] = v;
        }
    }

    const std::vector<int>& mode_x() const { return mx_; }
    const std::vector<int>& mode_y() const { return my_; }
    const std::vector<int>& mode_z() const { return mz_; }
    const std::vector<unsigned char>& keep() const { return keep_; }
    const std::vector<C>& phase() const { return phase_; }

    // --- reductions (impl.hpp:1047-1103) ------------------------------------
    template <class F>
    double block_sum(std::size_t count, F&& f) const {
        constexpr std::size_t kBlock = 256;
        const std::size_t blocks = (count + kBlock - 1) / kBlock;
        std::vector<Real> part(blocks);
#pragma omp parallel for schedule(static)
        for (long b = 0; b < static_cast<long>(blocks); ++b) {
            Real cache[kBlock];
            for (std::size_t t = 0; t < kBlock; ++t) {
                const std::size_t id = b * kBlock + t;
                cache[t] = id < count ? f(id) : Real{0};
            }
            for (std::size_t stride = kBlock / 2; stride > 0; stride /= 2)
                for (std::size_t t = 0; t < stride; ++t) cache[t] += cache[t + stride];
            part[b] = cache[0];
        }
        double s = 0.0;
        for (Real v : part) s += static_cast<double>(v);
        return s;
    }
    template <class F>
    double block_max(std::size_t count, F&& f) const {
        double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
        for (long id = 0; id < static_cast<long>(count); ++id) m = std::max(m, static_cast<double>(f(id)));
        return m;
    }

    double max_velocity() {
        for (int c = 0; c < 3; ++c) backward(u_hat_[c], phys_[c]);
        const auto& p = phys_;
        const double m = block_max(total_, [&](std::size_t id) {
            return p[0][id].x * p[0][id].x + p[1][id].x * p[1][id].x + p[2][id].x * p[2][id].x;
        });
        ++counters_.max_velocity_calls;
        return std::sqrt(m);
    }

    Diagnostics diagnostics(double t) {
        for (int c = 0; c < 3; ++c) backward(u_hat_[c], phys_[c]);
        const auto& p = phys_;
        const double e = block_sum(total_, [&](std::size_t id) {
            return Real{0.5} * (p[0][id].x * p[0][id].x + p[1][id].x * p[1][id].x + p[2][id].x * p[2][id].x);
        });
        const double v2 = block_max(total_, [&](std::size_t id) {
            return p[0][id].x * p[0][id].x + p[1][id].x * p[1][id].x + p[2][id].x * p[2][id].x;
        });
        // vorticity i k x u (impl.hpp:1147-1157)
#pragma omp parallel for schedule(static)
        for (long id = 0; id < static_cast<long>(total_); ++id) {
            const Real kx = static_cast<Real>(mx_[id]), ky = static_cast<Real>(my_[id]), kz = static_cast<Real>(mz_[id]);
            const C u0 = u_hat_[0][id], u1 = u_hat_[1][id], u2 = u_hat_[2][id];
            shift_hat_[0][id] = {-(ky * u2.y - kz * u1.y), ky * u2.x - kz * u1.x};
            shift_hat_[1][id] = {-(kz * u0.y - kx * u2.y), kz * u0.x - kx * u2.x};
            shift_hat_[2][id] = {-(kx * u1.y - ky * u0.y), kx * u1.x - ky * u0.x};
        }
        for (int c = 0; c < 3; ++c) backward(shift_hat_[c], shift_phys_[c]);
        const auto& w = shift_phys_;
        const double ens = block_sum(total_, [&](std::size_t id) {
            return Real{0.5} * (w[0][id].x * w[0][id].x + w[1][id].x * w[1][id].x + w[2][id].x * w[2][id].x);
        });
        const auto& u = u_hat_;
        const double div = block_max(total_, [&](std::size_t id) {
            const C d = cx_add(cx_add(cx_scale(u[0][id], static_cast<Real>(mx_[id])),
                                      cx_scale(u[1][id], static_cast<Real>(my_[id]))),
                               cx_scale(u[2][id], static_cast<Real>(mz_[id])));
            return std::sqrt(d.x * d.x + d.y * d.y);
        });
        const double inv = 1.0 / static_cast<double>(total_);
        Diagnostics row;
        row.t = t;
        row.kinetic_energy = e * inv;
        row.enstrophy = ens * inv;
        row.dissipation = 2.0 * static_cast<double>(nu_) * row.enstrophy;
        row.max_spectral_divergence = div;
        row.max_velocity = std::sqrt(v2);
        ++counters_.diagnostics_calls;
        return row;
    }

private:
    void build_tables() {
        mx_.resize(total_);
        my_.resize(total_);
        mz_.resize(total_);
        keep_.resize(total_);
        phase_.resize(total_);
        const int n = n_;
        for (std::size_t id = 0; id < total_; ++id) {
            const int ix = static_cast<int>(id % n), iy = static_cast<int>((id / n) % n),
                      iz = static_cast<int>(id / (static_cast<std::size_t>(n) * n));
            mx_[id] = signed_index(ix, n);
            my_[id] = signed_index(iy, n);
            mz_[id] = signed_index(iz, n);
            const double ang = (static_cast<double>(mx_[id] + my_[id] + mz_[id]) * kPi) / static_cast<double>(n);
            phase_[id] = {static_cast<Real>(std::cos(ang)), static_cast<Real>(std::sin(ang))};
            keep_[id] = keep_sphere(mx_[id], my_[id], mz_[id], n) ? 1 : 0;
        }
    }

    int n_;
    Real nu_;
    Grid3d<Real> grid_;
    std::size_t total_ = 0;
    std::vector<C> u_hat_[3], rhs_[3], g_[3], phys_[3], shift_hat_[3], shift_phys_[3];
    std::vector<C> product_, product_shift_, product_hat_, product_shift_hat_, phase_;
    std::vector<int> mx_, my_, mz_;
    std::vector<unsigned char> keep_;
    Counters counters_;
    double step_seconds_ = 0.0;
};

// Restatement of the driver loop (src/ava_taylor_green_impl.hpp:1310-1410) for
// the oracle solver. Writes the same CSV header/format and returns the last row.
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
    std::string output_dir = "output_oracle";
    bool quiet = true;
};

template <class Real>
Diagnostics run_taylor_green(const RunConfig& cfg, Counters* counters_out = nullptr, int* steps_out = nullptr) {
    if (cfg.dt <= 0.0 || cfg.cfl <= 0.0 || cfg.output_dt <= 0.0)
        throw std::runtime_error("--dt, --cfl, and --output-dt must be positive");
    std::filesystem::create_directories(cfg.output_dir);
    std::ofstream csv(cfg.output_dir + "/taylor_green_diagnostics.csv");
    csv << std::setprecision(17);
    csv << "t,kinetic_energy,dissipation,enstrophy,max_spectral_divergence,max_velocity,dt,cfl,main_dt,main_cfl\n";
    TaylorGreen<Real> solver(cfg.n, cfg.reynolds);
    const double h = 2.0 * kPi / static_cast<double>(cfg.n);
    double t = 0.0;
    int step = 0;
    int outputs = 0;
    Diagnostics last;
    auto emit = [&](double time, bool force, double dt_, double cfl_, double mdt, double mcfl) {
        Diagnostics row = solver.diagnostics(time);
        row.dt = dt_;
        row.cfl = cfl_;
        row.main_dt = mdt;
        row.main_cfl = mcfl;
        csv << row.t << ',' << row.kinetic_energy << ',' << row.dissipation << ',' << row.enstrophy << ','
            << row.max_spectral_divergence << ',' << row.max_velocity << ',' << dt_ << ',' << cfl_ << ',' << mdt
            << ',' << mcfl << '\n';
        if (!cfg.quiet && (force || outputs % cfg.print_every == 0)) {
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
            const double vmax = solver.max_velocity();
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
            solver.step(dt_step);
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
                const double fixed_cfl = solver.max_velocity() * cfg.dt / h;
                emit(t, step == cfg.steps, step == 0 ? 0.0 : cfg.dt, step == 0 ? 0.0 : fixed_cfl,
                     step == 0 ? 0.0 : cfg.dt, step == 0 ? 0.0 : fixed_cfl);
            }
            if (step != cfg.steps) {
                solver.step(cfg.dt);
                t += cfg.dt;
            }
        }
    }
    if (counters_out) *counters_out = solver.counters();
    if (steps_out) *steps_out = step;
    return last;
}

}  // namespace oracle

// Solvers on the pair-buffer pass kernels.
//
// Group<T>: `count` independent solves of one grid size n that share every
// launch (the member index is a grid dimension of each pass kernel). The
// Taylor-Green solver (TaylorGreenSolver, src/ava_taylor_green_impl.hpp:658-1308)
// is a group of one; the batched ensemble holds one group per distinct size and
// captures all groups' stages into one hipGraph per step.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "common.hpp"
#include "ensemble_passes.hpp"

// identifies the library build in the autotuner's cache file (build.py
// passes a hash of the sources)
#ifndef SPH_BUILD_ID
#define SPH_BUILD_ID "unversioned"
#endif
#include "spectral_hip.h"

namespace sph {

// ---------------------------------------------------------------------------
// RK3 coefficient blocks (alpha_s, beta_s*dt for s = 0..2, nu), uploaded
// through kernel arguments so that no host staging buffer can be overwritten
// while a copy is pending.
// ---------------------------------------------------------------------------
template <typename T>
struct CoefBlock {
    T v[32 * kCoefStride];
    int count;
};

template <typename T>
__global__ void k_set_coef_block(T* dst, CoefBlock<T> b) {
    const int i = threadIdx.x;
    if (i < b.count * kCoefStride) dst[i] = b.v[i];
}

// Williamson coefficients computed in T exactly as the reference writes them
// (impl.hpp:998, 1011, 1014, 1027, 1030).
template <typename T>
void rk_coefficients(double dt, T nu, T* out) {
    const T rdt = static_cast<T>(dt);
    out[0] = T(0);
    out[1] = rdt / T(3);
    out[2] = -T(5) / T(9);
    out[3] = T(15) * rdt / T(16);
    out[4] = -T(153) / T(128);
    out[5] = T(8) * rdt / T(15);
    out[6] = nu;
    out[7] = T(0);
}

// ---------------------------------------------------------------------------
// Seeded random-phase solenoidal field, E(k) ~ k^4 exp(-2 (k/4)^2).
//
// Half-spectrum layout [kx][kz][ky], interleaved (re, im) doubles. Every kept
// mode (|m|^2 <= (n/2)^2, no Nyquist index) gets a random-phase complex
// 3-vector from SplitMix64 keyed by (seed, signed mode, component), projected
// perpendicular to k and scaled to |u_hat|^2 ~ E(k) / (4 pi k^2); the kx = 0
// plane is made Hermitian. The field is finally scaled to kinetic energy
// <|u|^2>/2 = energy (unnormalised forward transform, so
// <|u|^2> = sum over the full spectrum of |u_hat|^2 / n^6).
// ---------------------------------------------------------------------------
inline std::uint64_t splitmix64(std::uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

inline void random_solenoidal_field(int n, std::uint64_t seed, double energy, double* u0, double* u1, double* u2) {
    if (n < 4 || n % 2) throw std::runtime_error("random field needs an even grid size >= 4");
    if (!(energy > 0.0)) throw std::runtime_error("random field energy must be positive");
    const int H = n / 2 + 1;
    const std::size_t modes = static_cast<std::size_t>(H) * n * n;
    double* u[3] = {u0, u1, u2};
    for (int c = 0; c < 3; ++c) std::fill(u[c], u[c] + 2 * modes, 0.0);
    auto sgn = [n](int i) { return i <= n / 2 ? i : i - n; };
    auto slot = [n](int kx, int kz, int ky) { return (static_cast<std::size_t>(kx) * n + kz) * n + ky; };
    double sum = 0.0;
    for (int kx = 0; kx < H; ++kx)
        for (int kz = 0; kz < n; ++kz)
            for (int ky = 0; ky < n; ++ky) {
                const int mx = kx, my = sgn(ky), mz = sgn(kz);
                const long k2 = static_cast<long>(mx) * mx + static_cast<long>(my) * my + static_cast<long>(mz) * mz;
                if (k2 == 0 || 4 * k2 > static_cast<long>(n) * n) continue;
                if (mx == n / 2 || my == n / 2 || mz == n / 2) continue;
                if (kx == 0 && !(my > 0 || (my == 0 && mz > 0))) continue;  // filled from its partner
                const std::uint64_t key =
                    (static_cast<std::uint64_t>(mx + n) * (2 * n + 1) + static_cast<std::uint64_t>(my + n)) *
                        (2 * n + 1) +
                    static_cast<std::uint64_t>(mz + n);
                std::complex<double> v[3];
                for (int c = 0; c < 3; ++c) {
                    const std::uint64_t r = splitmix64(seed * 0x2545F4914F6CDD1DULL + key * 3 + c);
                    const double th = 2.0 * kPi * static_cast<double>(r >> 11) * 0x1.0p-53;
                    v[c] = std::polar(1.0, th);
                }
                const double kk[3] = {double(mx), double(my), double(mz)};
                const std::complex<double> dot = kk[0] * v[0] + kk[1] * v[1] + kk[2] * v[2];
                double norm = 0.0;
                for (int c = 0; c < 3; ++c) {
                    v[c] -= kk[c] * dot / static_cast<double>(k2);
                    norm += std::norm(v[c]);
                }
                if (norm < 1e-20) continue;
                const double k = std::sqrt(static_cast<double>(k2));
                const double ek = k * k * k * k * std::exp(-2.0 * (k / 4.0) * (k / 4.0));
                const double amp = std::sqrt(ek / (4.0 * kPi * k * k) / norm);
                const std::size_t i = slot(kx, kz, ky);
                for (int c = 0; c < 3; ++c) {
                    const std::complex<double> w = amp * v[c];
                    u[c][2 * i] = w.real();
                    u[c][2 * i + 1] = w.imag();
                }
                double e = 0.0;
                for (int c = 0; c < 3; ++c) e += std::norm(amp * v[c]);
                if (kx == 0) {
                    const std::size_t j = slot(0, (n - kz) % n, (n - ky) % n);
                    for (int c = 0; c < 3; ++c) {
                        u[c][2 * j] = u[c][2 * i];
                        u[c][2 * j + 1] = -u[c][2 * i + 1];
                    }
                    sum += 2.0 * e;  // the mode and its partner
                } else {
                    sum += 2.0 * e;  // the mode and its conjugate at -k (not stored)
                }
            }
    if (!(sum > 0.0)) throw std::runtime_error("random field has no kept modes");
    const double n6 = std::pow(static_cast<double>(n), 6);
    const double scale = std::sqrt(energy * n6 / (0.5 * sum));
    for (int c = 0; c < 3; ++c)
        for (std::size_t q = 0; q < 2 * modes; ++q) u[c][q] *= scale;
}

// ---------------------------------------------------------------------------
// A group of `count` equal-size solves.
// ---------------------------------------------------------------------------
template <typename T>
class Group {
public:
    Group(int n_in, int count_in, const std::vector<double>& nu) : n(n_in), count(count_in) {
        if (count < 1) throw std::runtime_error("a group needs at least one member");
        if (static_cast<int>(nu.size()) != count) throw std::runtime_error("one viscosity per member");
        pt_ = table_or_throw<T>(n);
        H = n / 2 + 1;
        modes = static_cast<std::int64_t>(H) * n * n;
        npts = static_cast<std::int64_t>(n) * n * n;
        su = 3 * modes;
        sw = 12 * modes;
        U_.resize(count * su);
        G_.resize(count * su);
        W_.resize(count * sw);
        // physical fields only for diagnostics and the initial condition,
        // one member at a time
        P_.resize(6 * npts);
        coef_.resize(static_cast<std::size_t>(count) * kCoefStride);
        // the autotuner's timing launches of the forward-z pass read them
        hip_check(hipMemset(coef_.get(), 0, count * kCoefStride * sizeof(T)), "hipMemset");
        partial_.resize(3 * kReduceBlocks + 8);
        result_.resize(8);
        upload(tw_, twiddle_table<T>(n));
        upload(phase_, phase_table<T>(n));
        hip_check(hipMemset(G_.get(), 0, count * su * sizeof(cx<T>)), "hipMemset");
        hip_check(hipMemset(U_.get(), 0, count * su * sizeof(cx<T>)), "hipMemset");
        hip_check(hipMemset(W_.get(), 0, count * sw * sizeof(cx<T>)), "hipMemset");
        // the solver streams are non-blocking: finish the clears first
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        nu_.resize(count);
        for (int m = 0; m < count; ++m) nu_[m] = static_cast<T>(nu[m]);
        last_dt_.assign(count, std::nan(""));
        kb_ = n;
        autotune();
    }

    // Tuned passes: inverse y, x pass, forward y.
    static constexpr int kTuned = kTunedPasses;
    // per tuned pass: the variant kept, its lines per tile, and the mean
    // launch time of every variant (ms[pass * kMaxVariants + variant]; 0 when
    // not timed or not defined)
    // (in the layout kept: the transposed layout's variants and times when
    // it was chosen)
    // tuning: the first three tuned passes (inverse y, x, forward y);
    // tuned_pass: any of the four
    void tuning(int* variant, int* lines, double* ms) const {
        for (int p = 0; p < 3; ++p) tuned_pass(p, variant + p, lines + p, ms + p * kMaxVariants);
    }
    void tuned_pass(int p, int* variant, int* lines, double* ms) const {
        if (p < 0 || p >= kTuned) throw std::runtime_error("tuned pass out of range");
        const bool xt = xt_ && p < 3;  // the forward-z pass is the same in both layouts
        const int v = cur_var(p);
        *variant = v;
        *lines = xt && p != 1 ? pt_->xt_lines[v] : pt_->tile_lines[p][v];
        for (int q = 0; q < kMaxVariants; ++q) ms[q] = zb_ ? btune_ms_[p][q] : (xt ? xtune_ms_[p][q] : tune_ms_[p][q]);
    }
    // the variant of tuned pass p in the layout in use
    int cur_var(int p) const { return zb_ ? bvar_[p] : (xt_ && p < 3 ? xvar_[p] : var_[p]); }
    // the radix stages of the kept variant of tuned pass p (the y passes and
    // the LDS-staged x pass use radix_plan, the z-type passes their order)
    RadixPlan radix_plan_of(int p) const {
        if (p < 0 || p >= kTuned) throw std::runtime_error("tuned pass out of range");
        const bool xt = xt_ && p < 3;
        const int v = cur_var(p);
        if (p == 0 || p == 2 || (p == 1 && !xt && v == kXLdsVariant)) return radix_plan(n);
        return z_radix_plan(n, pt_->radix_order[p][v]);
    }
    // the work layout kept (0: line layout, 1: transposed between the y
    // passes and the x pass, 2: ky-blocked spectral side) and the timings
    // it was chosen by (0 where a layout was not timed): ms[0], ms[1] the
    // summed time of passes 1-3 with the line / transposed layout's fastest
    // variants, ms[2] passes 1-4 in the ky-blocked layout, ms[3] passes 1-4
    // with the faster of the other two (the forward-z pass is the same in
    // those)
    void layout(int* kind, double* ms) const {
        *kind = zb_ ? 2 : (xt_ ? 1 : 0);
        for (int i = 0; i < 4; ++i) ms[i] = layout_ms_[i];
    }
    // ky block of the state and of the spectral side of the pair fields
    int ky_block() const { return kb_; }
    // the transposed layout in use (the one-launch-per-pass ensemble kernels
    // have no form for it)
    bool transposed_layout() const { return xt_; }
    // pairs between kx rows / z planes of the x pass's fields: [kx][z][y] in
    // the line layout, [z][kx][y] in the ky-blocked one (an x tile's n/2+1
    // rows then lie in one plane of (n/2+1)*n pairs instead of n/2+1 planes
    // n*n pairs apart)
    long long x_krow() const { return xz_ ? n : static_cast<long long>(n) * n; }
    long long x_zrow() const { return xz_ ? static_cast<long long>(H) * n : n; }

    // pass kernels: 0 (inverse z) runs only in prepare_pairs, 1-4 per stage
    static constexpr int kPasses = 5;
    int n = 0, count = 0, H = 0;
    std::int64_t modes = 0, npts = 0, su = 0, sw = 0;  // per-member strides: state, work

    T nu(int m) const { return nu_[m]; }
    cx<T>* U() const { return U_.get(); }
    cx<T>* G() const { return G_.get(); }

    // Upload the coefficient blocks of the members whose dt changed.
    void set_dt(const double* dt, hipStream_t s) {
        bool changed = false;
        for (int m = 0; m < count; ++m) changed |= !(dt[m] == last_dt_[m]);
        if (!changed) return;
        for (int m0 = 0; m0 < count; m0 += 32) {
            CoefBlock<T> b{};
            b.count = std::min(32, count - m0);
            for (int m = 0; m < b.count; ++m) rk_coefficients<T>(dt[m0 + m], nu_[m0 + m], b.v + m * kCoefStride);
            k_set_coef_block<T><<<1, 32 * kCoefStride, 0, s>>>(coef_.get() + m0 * kCoefStride, b);
        }
        for (int m = 0; m < count; ++m) last_dt_[m] = dt[m];
        hip_check(hipGetLastError(), "coefficient upload");
    }

    // Inverse z of the state u (member 0 of the launch) into the velocity
    // pairs, the first three pair fields of W.
    PairInvZArgs<T> inv_z_args(const cx<T>* u, cx<T>* W) const {
        PairInvZArgs<T> a{};
        a.u3 = u;
        for (int c = 0; c < 3; ++c) a.out[c] = W + c * 2 * modes;
        a.ny = n;
        a.n = n;
        a.phase = phase_.get();
        a.tw = tw_.get();
        a.ms_u = su;
        a.ms_w = sw;
        a.kb = kb_;
        return a;
    }

    // Arguments of the x pass (pair fields at src, member 0) and of the
    // forward-z pass (state u, g and pair fields W of member `first`).
    PairXArgs<T> x_args(cx<T>* src) const {
        const std::int64_t M2 = 2 * modes;
        PairXArgs<T> a{};
        for (int c = 0; c < 3; ++c) a.ab[c] = src + c * M2;
        for (int q = 0; q < 6; ++q) a.out[q] = src + q * M2;
        a.ny = n;
        a.nz = n;
        a.krow = x_krow();
        a.zrow = x_zrow();
        a.scale = T(1) / static_cast<T>(npts);
        a.tw = tw_.get();
        a.ms_w = sw;
        return a;
    }
    PairFwdZArgs<T> fwd_z_args(cx<T>* u, cx<T>* g, cx<T>* W, int first, int stage) const {
        const std::int64_t M2 = 2 * modes;
        PairFwdZArgs<T> a{};
        for (int q = 0; q < 6; ++q) a.in[q] = W + q * M2;
        a.u3 = u;
        a.g3 = g;
        a.ny = n;
        a.n = n;
        a.coef = coef_.get() + first * kCoefStride;
        a.stage = stage;
        a.phase = phase_.get();
        a.tw = tw_.get();
        a.ms_u = su;
        a.ms_w = sw;
        a.kb = kb_;
        return a;
    }

    // Everything the one-launch-per-pass ensemble kernels need of this group
    // (every member; ensemble_passes.hpp).
    EnsGroupArgs<T> ens_args() const {
        EnsGroupArgs<T> e{};
        cx<T>* W = W_.get();
        e.n = n;
        e.tw = tw_.get();
        e.inv = inv_z_args(U_.get(), W);
        for (int st = 0; st < 3; ++st) e.fwd[st] = fwd_z_args(U_.get(), G_.get(), W, 0, st);
        e.x = x_args(zb_ ? X_.get() : W);
        e.lines = W;
        e.npairlines[0] = 3LL * H * n;
        e.npairlines[1] = 6LL * H * n;
        e.member_stride = sw;
        e.blocked = zb_ ? 1 : 0;
        e.xlines = zb_ ? X_.get() : nullptr;
        e.xs_kx = 2 * x_krow();
        e.xs_z = 2 * x_zrow();
        e.field_stride = 2 * modes;
        return e;
    }

    // pass k (0 inverse z, 1 inverse y, 2 inverse x + products + forward x,
    // 3 forward y, 4 forward z + update + the next stage's inverse z) of
    // stage `stage` for members [first, first + members); u and g address
    // member 0 of the state family. Pass 0 runs only when the velocity pairs
    // in W are not those of the state (pairs_valid()); pass 4 leaves them so.
    // vmax2 (pass 2 only, register-direct x pass): max |u|^2 of the state is
    // reduced into it as well.
    void launch_pass(int k, int stage, cx<T>* u, cx<T>* g, int first, int members, hipStream_t s,
                     unsigned long long* vmax2 = nullptr) const {
        const std::int64_t M2 = 2 * modes;
        cx<T>* W = W_.get() + first * sw;
        u += first * su;
        g += first * su;
        // transposed layout: the inverse y pass moves the velocity pairs from
        // W (line layout) to X (transposed), the x pass works in X, the
        // forward y pass moves the product pairs back to W; ky-blocked
        // layout: the same moves between W (ky-blocked) and X (line layout)
        cx<T>* X = xt_ || zb_ ? X_.get() + first * sw : nullptr;
        const int zv = zb_ ? bvar_[3] : var_[3];
        switch (k) {
            case 0:
                pt_->pinv_z(inv_z_args(u, W), PZ_RHS, n, H, members, zv, s);
                break;
            case 1:
                if (xt_)
                    pt_->plines_t(W, X, 3, M2, sw, members, true, tw_.get(), xvar_[0], s);
                else if (zb_)
                    pt_->plines_b(W, X, 3, M2, sw, 2 * x_krow(), 2 * x_zrow(), members, true, tw_.get(), bvar_[0], s);
                else
                    pt_->plines(W, 3LL * H * n, sw, members, true, tw_.get(), var_[0], s);
                break;
            case 2: {
                PairXArgs<T> a = x_args(X ? X : W);
                a.vmax2 = vmax2;
                const int xv = cur_var(1);
                if (vmax2)
                    pt_->px_products_vmax(a, xv, xt_, s);
                else
                    pt_->px_products(a, n, n, members, xv, xt_, s);
                break;
            }
            case 3:
                if (xt_)
                    pt_->plines_t(X, W, 6, M2, sw, members, false, tw_.get(), xvar_[2], s);
                else if (zb_)
                    pt_->plines_b(X, W, 6, M2, sw, 2 * x_krow(), 2 * x_zrow(), members, false, tw_.get(), bvar_[2],
                                  s);
                else
                    pt_->plines(W, 6LL * H * n, sw, members, false, tw_.get(), var_[2], s);
                break;
            default: {
                const PairFwdZArgs<T> a = fwd_z_args(u, g, W, first, stage);
                pt_->pfwd_z_rk(a, inv_z_args(u, W), n, H, members, zv, s);
                break;
            }
        }
    }

    // Velocity pairs of the current state for the first stage: the last pass
    // of every stage leaves them in W, so they are recomputed only after the
    // state was set from outside (initial condition, set_uhat) or W was used
    // for something else (time_passes).
    void prepare_pairs(hipStream_t s) {
        if (pairs_valid_) return;
        launch_pass(0, 0, U_.get(), G_.get(), 0, count, s);
        pairs_valid_ = true;
    }
    void invalidate_pairs() { pairs_valid_ = false; }
    bool pairs_valid() const { return pairs_valid_; }

    // passes 1-4 of a stage (prepare_pairs first)
    void launch_stage(int stage, hipStream_t s) const {
        for (int k = 1; k < kPasses; ++k) launch_pass(k, stage, U_.get(), G_.get(), 0, count, s);
    }

    // A step split at the first use of dt, for the adaptive time step: the
    // head is stage 0's inverse y and x passes, which need no dt and whose x
    // pass forms the physical velocity of the current state anyway, so it
    // also reduces max |u|^2 into *vmax2; stage 0's forward y pass needs no
    // dt either (launch_step_mid); the tail is the rest of the step, from
    // stage 0's forward z and update on. One solve per group, register-direct
    // x pass only.
    bool x_reduces_velocity() const { return count == 1 && (xt_ || cur_var(1) != kXLdsVariant); }
    void launch_step_head(unsigned long long* vmax2, hipStream_t s) const {
        k_zero_vmax<<<1, kVmaxSlots, 0, s>>>(vmax2);
        launch_pass(1, 0, U_.get(), G_.get(), 0, 1, s);
        launch_pass(2, 0, U_.get(), G_.get(), 0, 1, s, vmax2);
    }
    void launch_step_mid(hipStream_t s) const { launch_pass(3, 0, U_.get(), G_.get(), 0, 1, s); }
    void launch_step_tail(hipStream_t s) const {
        launch_pass(4, 0, U_.get(), G_.get(), 0, 1, s);
        for (int st = 1; st < 3; ++st) launch_stage(st, s);
    }

    // member m: u_hat -> physical (the group's one physical buffer).
    // PZ_DIAG: u in P[0..2], omega in P[3..5]; PZ_VEL: u in P[0..2]. The
    // transforms run in the back half of W (product pair fields 3-5, unused
    // between stages), so the velocity pairs for the next step stay intact.
    void inverse(int m, int mode, hipStream_t s) const {
        const std::int64_t M2 = 2 * modes;
        cx<T>* W = W_.get() + m * sw + 3 * M2;
        T* P = P_.get();
        PairInvZArgs<T> a{};
        a.u3 = U_.get() + m * su;
        for (int c = 0; c < 3; ++c) a.out[c] = W + c * M2;
        a.ny = n;
        a.n = n;
        a.phase = phase_.get();
        a.tw = tw_.get();
        a.kb = kb_;
        // ky-blocked layout: the y pass moves the pairs into the back half of
        // X (line layout), where the x pass reads them
        cx<T>* const XY = zb_ ? X_.get() + m * sw + 3 * M2 : W;
        PairInvXArgs<T> x{};
        x.ny = n;
        x.nz = n;
        x.krow = x_krow();
        x.zrow = x_zrow();
        x.scale = T(1) / static_cast<T>(npts);
        x.tw = tw_.get();
        if (mode == PZ_DIAG) {
            pt_->pinv_z(a, PZ_DIAG, n, H, 1, 0, s);
            if (zb_)
                pt_->plines_b(W, XY, 3, M2, 0, 2 * x_krow(), 2 * x_zrow(), 1, true, tw_.get(), bvar_[0], s);
            else
                pt_->plines(W, 3LL * H * n, 0, 1, true, tw_.get(), var_[0], s);
            for (int c = 0; c < 3; ++c) {
                x.ab[c] = XY + c * M2;
                x.oa[c] = P + c * npts;
                x.ob[c] = P + (3 + c) * npts;
            }
            x.npairs = 3;
        } else {
            pt_->pinv_z(a, PZ_VEL, n, H, 1, 0, s);
            if (zb_)
                pt_->plines_b(W, XY, 2, M2, 0, 2 * x_krow(), 2 * x_zrow(), 1, true, tw_.get(), bvar_[0], s);
            else
                pt_->plines(W, 2LL * H * n, 0, 1, true, tw_.get(), var_[0], s);
            x.ab[0] = XY;
            x.oa[0] = P;
            x.ob[0] = P + npts;
            x.ab[1] = XY + M2;
            x.oa[1] = P + 2 * npts;
            x.ob[1] = nullptr;
            x.npairs = 2;
        }
        pt_->pinv_x(x, n, n, 1, s);
        hip_check(hipGetLastError(), "inverse transform launch");
    }

    // impl.hpp:1122-1179 for member m (reductions in FP64)
    sph_diagnostics diagnostics(int m, double t, hipStream_t s) {
        inverse(m, PZ_DIAG, s);
        T* P = P_.get();
        k_reduce_phys<T><<<kReduceBlocks, kReduceThreads, 0, s>>>(P, P + npts, P + 2 * npts, P + 3 * npts,
                                                                   P + 4 * npts, P + 5 * npts, npts, partial_.get());
        k_reduce_final<<<1, kReduceThreads, 0, s>>>(partial_.get(), kReduceBlocks, 3, 0b010, result_.get());
        k_reduce_div3<T><<<kReduceBlocks, kReduceThreads, 0, s>>>(U_.get() + m * su, n, kb_, partial_.get());
        k_reduce_final<<<1, kReduceThreads, 0, s>>>(partial_.get(), kReduceBlocks, 1, 0b1, result_.get() + 3);
        double r[4];
        hip_check(hipMemcpyAsync(r, result_.get(), sizeof(r), hipMemcpyDeviceToHost, s), "hipMemcpy diagnostics");
        hip_check(hipStreamSynchronize(s), "diagnostics");
        const double inv = 1.0 / static_cast<double>(npts);
        sph_diagnostics d{};
        d.t = t;
        d.kinetic_energy = r[0] * inv;
        d.enstrophy = r[2] * inv;
        d.dissipation = 2.0 * static_cast<double>(nu_[m]) * d.enstrophy;
        d.max_spectral_divergence = r[3];
        d.max_velocity = std::sqrt(r[1]);
        return d;
    }

    // impl.hpp:1105-1120 for member m
    double max_velocity(int m, hipStream_t s) {
        inverse(m, PZ_VEL, s);
        T* P = P_.get();
        k_reduce_phys<T><<<kReduceBlocks, kReduceThreads, 0, s>>>(P, P + npts, P + 2 * npts, nullptr, nullptr,
                                                                   nullptr, npts, partial_.get());
        k_reduce_final<<<1, kReduceThreads, 0, s>>>(partial_.get(), kReduceBlocks, 3, 0b010, result_.get());
        double r[3];
        hip_check(hipMemcpyAsync(r, result_.get(), sizeof(r), hipMemcpyDeviceToHost, s), "hipMemcpy max_velocity");
        hip_check(hipStreamSynchronize(s), "max_velocity");
        return std::sqrt(r[1]);
    }

    // component c of member m (component-interleaved state: strided copies)
    // Component c of member m. The state is component-interleaved, so the
    // member's whole state moves in one contiguous copy (a strided copy of
    // 8-byte rows is orders of magnitude slower) and c is picked on the host.
    void get_uhat(int m, int c, double* host, hipStream_t s) const {
        if (c < 0 || c > 2) throw std::runtime_error("component out of range");
        const std::vector<cx<T>> tmp = download_state(m, s);
        for (std::int64_t i = 0; i < modes; ++i) {
            host[2 * i] = tmp[3 * i + c].x;
            host[2 * i + 1] = tmp[3 * i + c].y;
        }
    }

    void set_uhat(int m, int c, const double* host, hipStream_t s) {
        if (c < 0 || c > 2) throw std::runtime_error("component out of range");
        std::vector<cx<T>> tmp = download_state(m, s);
        for (std::int64_t i = 0; i < modes; ++i)
            tmp[3 * i + c] = {static_cast<T>(host[2 * i]), static_cast<T>(host[2 * i + 1])};
        upload_state(m, tmp, s);
    }

    // all three components of member m at once
    void set_state(int m, const double* const* u, hipStream_t s) {
        std::vector<cx<T>> tmp(static_cast<std::size_t>(su));
        for (std::int64_t i = 0; i < modes; ++i)
            for (int c = 0; c < 3; ++c)
                tmp[3 * i + c] = {static_cast<T>(u[c][2 * i]), static_cast<T>(u[c][2 * i + 1])};
        upload_state(m, tmp, s);
    }

    // impl.hpp:796-817 for member m: TG field at cell centres, 3 forward
    // transforms, project_truncate
    void taylor_green(int m, hipStream_t s) {
        const std::int64_t M = modes;
        T* P = P_.get();
        cx<T>* W = W_.get() + m * sw;
        k_taylor_green<T><<<2048, 256, 0, s>>>(P, P + npts, P + 2 * npts, n);
        FwdXArgs<T> fx{};
        for (int c = 0; c < 3; ++c) {
            fx.in[c] = P + c * npts;
            fx.out[c] = W + c * M;
        }
        fx.nin = 3;
        fx.ny = n;
        fx.nz = n;
        fx.tw = tw_.get();
        pt_->fwd_x(fx, n, n, s);
        pt_->lines(W, 3LL * H * n, false, tw_.get(), s);
        FwdZArgs<T> fz{};
        for (int c = 0; c < 3; ++c) fz.in[c] = W + c * M;
        fz.u3 = U_.get() + m * su;
        fz.ny = n;
        fz.n = n;
        fz.tw = tw_.get();
        fz.kb = kb_;
        pt_->fwd_z(fz, FZ_PROJECT, n, H, s);
        hip_check(hipGetLastError(), "initial condition launch");
        pairs_valid_ = false;
    }

    // Mean duration of each pass kernel (all members) over `reps` launches;
    // the RK pass updates scratch copies of the state so the state is kept
    // (the work buffer is not: the step after this recomputes its pairs).
    void time_passes(int reps, double* ms, double* bytes, hipStream_t s) {
        pairs_valid_ = false;
        if (reps < 1) reps = 1;
        hip_check(hipStreamSynchronize(s), "time_passes");
        DeviceArray<cx<T>> u2(count * su), g2(count * su);
        hip_check(hipMemcpyAsync(u2.get(), U_.get(), count * su * sizeof(cx<T>), hipMemcpyDeviceToDevice, s), "copy");
        hip_check(hipMemcpyAsync(g2.get(), G_.get(), count * su * sizeof(cx<T>), hipMemcpyDeviceToDevice, s), "copy");
        if (std::isnan(last_dt_[0])) {
            std::vector<double> zero(count, 0.0);
            set_dt(zero.data(), s);
        }
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        for (int k = 0; k < kPasses; ++k) {
            // warm-up; the passes before k make its inputs genuine
            for (int j = 0; j <= k; ++j) launch_pass(j, 1, u2.get(), g2.get(), 0, count, s);
            hip_check(hipEventRecord(a, s), "hipEventRecord");
            for (int r = 0; r < reps; ++r) launch_pass(k, 1, u2.get(), g2.get(), 0, count, s);
            hip_check(hipEventRecord(b, s), "hipEventRecord");
            hip_check(hipEventSynchronize(b), "hipEventSynchronize");
            float e = 0.f;
            hip_check(hipEventElapsedTime(&e, a, b), "hipEventElapsedTime");
            ms[k] = static_cast<double>(e) / reps;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        hip_check(hipGetLastError(), "time_passes");
        pass_bytes(bytes);
        hip_check(hipStreamSynchronize(s), "time_passes");
    }

    // Algorithmic bytes per launch of each pass (each array read or written
    // once), C = one complex half-spectrum field of every member: inverse z
    // 3C + 6C; inverse y 6C + 6C; x pass 6C + 12C; forward y 12C + 12C;
    // forward z + update + next inverse z 12C + u, g (6C) in, u, g (6C) and
    // the next stage's velocity pairs (6C) out (its state comes from LDS),
    // except that stage 0 reads no g and stage 2 stores none: 27C, 30C, 27C,
    // 28C per launch over a step.
    void pass_bytes(double* bytes) const {
        const double C = static_cast<double>(modes) * 2.0 * sizeof(T) * count;
        const double b5[kPasses] = {9 * C, 12 * C, 18 * C, 24 * C, 28 * C};
        for (int k = 0; k < kPasses; ++k) bytes[k] = b5[k];
    }

private:
    // Autotuner: every variant of each tuned pass (tile width, and for the
    // x pass also the kernel form) is launched on this group's zeroed work
    // buffer, one warm-up and three timed launches between HIP events, and
    // the fastest is kept. Results are cached per (precision, n, members) for
    // the process; SPH_AUTOTUNE=0 keeps variant 0 everywhere,
    // SPH_TILE_VARIANT=v forces variant v wherever it exists.
    struct Choice {
        int var[kTuned], xvar[kTuned], bvar[kTuned];
        double ms[kTuned][kMaxVariants], xms[kTuned][kMaxVariants], bms[kTuned][kMaxVariants], layout_ms[4];
        int layout;  // 0 line, 1 transposed, 2 ky-blocked
    };
    void apply(const Choice& c) {
        for (int p = 0; p < kTuned; ++p) {
            var_[p] = c.var[p];
            xvar_[p] = c.xvar[p];
            bvar_[p] = c.bvar[p];
            for (int v = 0; v < kMaxVariants; ++v) {
                tune_ms_[p][v] = c.ms[p][v];
                xtune_ms_[p][v] = c.xms[p][v];
                btune_ms_[p][v] = c.bms[p][v];
            }
        }
        for (int i = 0; i < 4; ++i) layout_ms_[i] = c.layout_ms[i];
        use_layout(c.layout);
    }
    // The autotuner's choices across processes: with SPH_TUNE_CACHE set to a
    // file path, a choice is looked up there by (library build, device,
    // precision, n, members, forced layout) before timing, and a new one is
    // appended after it. One line per choice:
    //   sph-tune-2 <build> <device> <bytes> <n> <members> <forced layout> <layout>
    // (forced layout: SPH_LAYOUT or -1, plus 10 with SPH_X_ZMAJOR=0)
    //     <var x4> <xvar x4> <bvar x4> <layout ms x4> <ms x16> <xms x16> <bms x16>
    // (the last matching line wins). Entries whose variants do not exist in
    // this build or whose layout does not fit are ignored.
    static std::string tune_key(int layout_key, int n_, int members) {
        int dev = 0;
        hipDeviceProp_t prop{};
        std::string device = "unknown";
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            device = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
        for (char& ch : device)
            if (ch == ' ') ch = '_';
        return std::string("sph-tune-2 ") + SPH_BUILD_ID + " " + device + " " + std::to_string(sizeof(T)) + " " +
               std::to_string(n_) + " " + std::to_string(members) + " " + std::to_string(layout_key);
    }
    bool choice_valid(const Choice& c, bool xt_allowed, bool zb_allowed) const {
        if (c.layout < 0 || c.layout > 2) return false;
        if (c.layout == 1 && !xt_allowed) return false;
        if (c.layout == 2 && !zb_allowed) return false;
        for (int p = 0; p < kTuned; ++p) {
            if (c.var[p] < 0 || c.var[p] >= pt_->variants[p]) return false;
            if (c.bvar[p] < 0 || c.bvar[p] >= pt_->variants[p]) return false;
            if (c.xvar[p] < 0 || c.xvar[p] >= kMaxVariants) return false;
        }
        return true;
    }
    static bool load_tune_file(const char* path, const std::string& key, Choice& out) {
        std::FILE* f = std::fopen(path, "r");
        if (!f) return false;
        bool found = false;
        char line[4096];
        while (std::fgets(line, sizeof(line), f)) {
            const std::size_t klen = key.size();
            if (std::strncmp(line, key.c_str(), klen) != 0 || line[klen] != ' ') continue;
            Choice c{};
            const char* q = line + klen;
            char* end = nullptr;
            auto next = [&](double& v) {
                v = std::strtod(q, &end);
                if (end == q) return false;
                q = end;
                return true;
            };
            double v = 0.0;
            bool ok = next(v);
            c.layout = static_cast<int>(v);
            for (int p = 0; ok && p < kTuned; ++p) {
                ok = next(v);
                c.var[p] = static_cast<int>(v);
            }
            for (int p = 0; ok && p < kTuned; ++p) {
                ok = next(v);
                c.xvar[p] = static_cast<int>(v);
            }
            for (int p = 0; ok && p < kTuned; ++p) {
                ok = next(v);
                c.bvar[p] = static_cast<int>(v);
            }
            for (int i = 0; ok && i < 4; ++i) ok = next(c.layout_ms[i]);
            for (int p = 0; ok && p < kTuned; ++p)
                for (int i = 0; ok && i < kMaxVariants; ++i) ok = next(c.ms[p][i]);
            for (int p = 0; ok && p < kTuned; ++p)
                for (int i = 0; ok && i < kMaxVariants; ++i) ok = next(c.xms[p][i]);
            for (int p = 0; ok && p < kTuned; ++p)
                for (int i = 0; ok && i < kMaxVariants; ++i) ok = next(c.bms[p][i]);
            if (ok) {
                out = c;
                found = true;
            }
        }
        std::fclose(f);
        return found;
    }
    static void append_tune_file(const char* path, const std::string& key, const Choice& c) {
        std::string line = key + " " + std::to_string(c.layout);
        char buf[64];
        for (int p = 0; p < kTuned; ++p) line += " " + std::to_string(c.var[p]);
        for (int p = 0; p < kTuned; ++p) line += " " + std::to_string(c.xvar[p]);
        for (int p = 0; p < kTuned; ++p) line += " " + std::to_string(c.bvar[p]);
        auto add = [&](double v) {
            std::snprintf(buf, sizeof(buf), " %.6g", v);
            line += buf;
        };
        for (int i = 0; i < 4; ++i) add(c.layout_ms[i]);
        for (int p = 0; p < kTuned; ++p)
            for (int i = 0; i < kMaxVariants; ++i) add(c.ms[p][i]);
        for (int p = 0; p < kTuned; ++p)
            for (int i = 0; i < kMaxVariants; ++i) add(c.xms[p][i]);
        for (int p = 0; p < kTuned; ++p)
            for (int i = 0; i < kMaxVariants; ++i) add(c.bms[p][i]);
        line += "\n";
        // one write in append mode: concurrent processes add whole lines
        if (std::FILE* f = std::fopen(path, "a")) {
            std::fputs(line.c_str(), f);
            std::fclose(f);
        }
    }
    void autotune() {
        // SPH_LAYOUT=0 / 1 / 2 forces the line / transposed / ky-blocked work
        // layout (the latter two where the length supports them and the
        // second work buffer fits)
        const char* lay = std::getenv("SPH_LAYOUT");
        const int forced_layout = lay && lay[0] >= '0' && lay[0] <= '2' && lay[1] == '\0' ? lay[0] - '0' : -1;
        const bool fits = (pt_->xt || pt_->zb) && forced_layout != 0 && second_buffer_fits();
        const bool xt_allowed = pt_->xt && fits && forced_layout != 2;
        const bool zb_allowed = pt_->zb && fits && forced_layout != 1;
        const char* force = std::getenv("SPH_TILE_VARIANT");
        if (force && force[0] >= '0' && force[0] < '0' + kMaxVariants && force[1] == '\0') {
            const int f = force[0] - '0';
            for (int p = 0; p < kTuned; ++p) {
                var_[p] = std::min(f, pt_->variants[p] - 1);
                bvar_[p] = p == 0 || p == 2 ? std::min(f, 2) : var_[p];
                // transposed layout: y passes 0..2; x pass 0, 1 and the
                // alternative radix order (no LDS-staged form)
                const bool x_alt = f == kXOrderVariant && pt_->variants[1] > kXOrderVariant;
                xvar_[p] = p == 3 ? 0 : (p == 1 ? (x_alt ? f : std::min(f, 1)) : std::min(f, 2));
            }
            use_layout(forced_layout == 1 && xt_allowed ? 1 : (forced_layout == 2 && zb_allowed ? 2 : 0));
            return;
        }
        static std::mutex mu;
        static std::map<std::tuple<int, int, int, int>, Choice> cache;
        // the forced layout and SPH_X_ZMAJOR=0 (+10) key the choice
        const char* xz = std::getenv("SPH_X_ZMAJOR");
        const int layout_key = forced_layout + (xz && xz[0] == '0' ? 10 : 0);
        const auto key = std::make_tuple(static_cast<int>(sizeof(T)), n, count, layout_key);
        {
            std::lock_guard<std::mutex> lock(mu);
            const auto it = cache.find(key);
            if (it != cache.end()) {
                apply(it->second);
                return;
            }
        }
        const char* env = std::getenv("SPH_AUTOTUNE");
        if (env && env[0] == '0') {
            use_layout(forced_layout == 1 && xt_allowed ? 1 : (forced_layout == 2 && zb_allowed ? 2 : 0));
            return;
        }
        const char* tune_file = std::getenv("SPH_TUNE_CACHE");
        const std::string file_key = tune_file && tune_file[0] ? tune_key(layout_key, n, count) : std::string();
        if (!file_key.empty()) {
            Choice c{};
            if (load_tune_file(tune_file, file_key, c) && choice_valid(c, xt_allowed, zb_allowed)) {
                apply(c);
                std::lock_guard<std::mutex> lock(mu);
                cache.emplace(key, c);
                return;
            }
        }
        Stream stream;
        const hipStream_t s = stream.get();
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        const int pass_of[kTuned] = {1, 2, 3, 4};
        // every variant of the tuned passes in the current layout (nvar[p]
        // of pass p; 0: not tuned here); returns the summed time of the
        // fastest of passes 1-3, the ones the line and transposed layouts
        // differ in (the forward-z pass's fastest is in ms[3][var[3]])
        auto tune_layout = [&](int (&var)[kTuned], double (&ms)[kTuned][kMaxVariants], const int* nvar,
                               bool transposed) {
            double total = 0.0;
            for (int p = 0; p < kTuned; ++p) {
                if (nvar[p] < 1) continue;
                int best = 0;
                for (int v = 0; v < nvar[p]; ++v) {
                    // the transposed layout has no LDS-staged x pass
                    if (transposed && p == 1 && v == kXLdsVariant) continue;
                    var[p] = v;
                    launch_pass(pass_of[p], 0, U_.get(), G_.get(), 0, count, s);
                    hip_check(hipEventRecord(a, s), "hipEventRecord");
                    for (int r = 0; r < 3; ++r) launch_pass(pass_of[p], 0, U_.get(), G_.get(), 0, count, s);
                    hip_check(hipEventRecord(b, s), "hipEventRecord");
                    hip_check(hipEventSynchronize(b), "hipEventSynchronize");
                    float e = 0.f;
                    hip_check(hipEventElapsedTime(&e, a, b), "hipEventElapsedTime");
                    ms[p][v] = e / 3.0;
                    if (ms[p][v] < ms[p][best]) best = v;
                }
                var[p] = best;
                if (p < 3) total += ms[p][best];
            }
            return total;
        };
        // the line layout's forward-z pass also serves the transposed layout
        int keep = 0;
        double best_other = 0.0;
        use_layout(0);
        // the line layout in full unless another one is forced and available
        // (a forced layout this length or device lacks falls back to it)
        const bool forced_other = (forced_layout == 1 && xt_allowed) || (forced_layout == 2 && zb_allowed);
        if (!forced_other) {
            layout_ms_[0] = tune_layout(var_, tune_ms_, pt_->variants, false);
            best_other = layout_ms_[0];
        } else {
            const int nz[kTuned] = {0, 0, 0, pt_->variants[3]};
            tune_layout(var_, tune_ms_, nz, false);
        }
        if (xt_allowed) {
            use_layout(1);
            const int nx[kTuned] = {3, pt_->variants[1], 3, 0};
            layout_ms_[1] = tune_layout(xvar_, xtune_ms_, nx, true);
            if (forced_layout == 1 || layout_ms_[1] < layout_ms_[0]) {
                keep = 1;
                best_other = layout_ms_[1];
            }
        }
        layout_ms_[3] = best_other + tune_ms_[3][var_[3]];
        if (zb_allowed) {
            use_layout(2);
            const int nb[kTuned] = {3, pt_->variants[1], 3, pt_->variants[3]};
            layout_ms_[2] = tune_layout(bvar_, btune_ms_, nb, false) + btune_ms_[3][bvar_[3]];
            if (forced_layout == 2 || layout_ms_[2] < layout_ms_[3]) keep = 2;
        }
        use_layout(keep);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        // the timing launches left transforms of zeros in W: clear it again
        hip_check(hipMemsetAsync(W_.get(), 0, count * sw * sizeof(cx<T>), s), "hipMemset");
        hip_check(hipStreamSynchronize(s), "autotune");
        Choice c{};
        for (int p = 0; p < kTuned; ++p) {
            c.var[p] = var_[p];
            c.xvar[p] = xvar_[p];
            c.bvar[p] = bvar_[p];
            for (int v = 0; v < kMaxVariants; ++v) {
                c.ms[p][v] = tune_ms_[p][v];
                c.xms[p][v] = xtune_ms_[p][v];
                c.bms[p][v] = btune_ms_[p][v];
            }
        }
        for (int i = 0; i < 4; ++i) c.layout_ms[i] = layout_ms_[i];
        c.layout = keep;
        if (!file_key.empty()) append_tune_file(tune_file, file_key, c);
        std::lock_guard<std::mutex> lock(mu);
        cache.emplace(key, c);
    }

    // The transposed and the ky-blocked layouts need a second work buffer of
    // 12C per member: taken only with room to spare on the device.
    bool second_buffer_fits() const {
        std::size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
        const std::size_t need = static_cast<std::size_t>(count) * sw * sizeof(cx<T>);
        return need + (std::size_t(4) << 30) <= free_b;
    }
    // 0: line layout; 1: transposed [z][y][kx] between the y passes and the
    // x pass; 2: ky-blocked state and spectral pair fields (z tiles are
    // contiguous blocks), line layout in X for the x pass. Only while the
    // state holds no data (construction), since the state's layout changes.
    void use_layout(int kind) {
        xt_ = kind == 1;
        zb_ = kind == 2;
        kb_ = zb_ ? pt_->zb_kb : n;
        const char* xz = std::getenv("SPH_X_ZMAJOR");
        xz_ = zb_ && !(xz && xz[0] == '0');
        if (kind != 0 && !X_.get()) X_.resize(static_cast<std::size_t>(count) * sw);
        if (kind == 0) X_.release();
    }

    // Member m's state in the plain mode order ((kx*n + kz)*n + ky, three
    // components per mode), whatever the device layout.
    std::vector<cx<T>> download_state(int m, hipStream_t s) const {
        std::vector<cx<T>> tmp(static_cast<std::size_t>(su));
        hip_check(hipMemcpyAsync(tmp.data(), U_.get() + m * su, su * sizeof(cx<T>), hipMemcpyDeviceToHost, s),
                  "hipMemcpy get_uhat");
        hip_check(hipStreamSynchronize(s), "get_uhat");
        if (kb_ == n) return tmp;
        std::vector<cx<T>> out(tmp.size());
        for_each_mode_block([&](std::int64_t plain, std::int64_t dev) {
            for (int c = 0; c < 3; ++c) out[3 * plain + c] = tmp[3 * dev + c];
        });
        return out;
    }
    void upload_state(int m, const std::vector<cx<T>>& plain_order, hipStream_t s) {
        const std::vector<cx<T>>* src = &plain_order;
        std::vector<cx<T>> dev_order;
        if (kb_ != n) {
            dev_order.resize(plain_order.size());
            for_each_mode_block([&](std::int64_t plain, std::int64_t dev) {
                for (int c = 0; c < 3; ++c) dev_order[3 * dev + c] = plain_order[3 * plain + c];
            });
            src = &dev_order;
        }
        hip_check(hipMemcpyAsync(U_.get() + m * su, src->data(), su * sizeof(cx<T>), hipMemcpyHostToDevice, s),
                  "hipMemcpy set_uhat");
        hip_check(hipStreamSynchronize(s), "set_uhat");
        pairs_valid_ = false;
    }
    // f(plain mode index, device mode index) over all modes of one member
    // (ky blocks of kb_, spectral_mode_index)
    template <typename F>
    void for_each_mode_block(F&& f) const {
        for (int kx = 0; kx < H; ++kx)
            for (int kz = 0; kz < n; ++kz)
                for (int ky = 0; ky < n; ++ky)
                    f(spectral_mode_index(kx, kz, ky, n, n), spectral_mode_index(kx, kz, ky, n, kb_));
    }

    const PassTable<T>* pt_ = nullptr;
    DeviceArray<cx<T>> U_, G_, W_, tw_, phase_;
    DeviceArray<T> P_, coef_;
    DeviceArray<double> partial_, result_;
    std::vector<T> nu_;
    std::vector<double> last_dt_;
    bool pairs_valid_ = false;     // W holds the velocity pairs of U (every member)
    int var_[kTuned] = {0, 0, 0};  // variant of each tuned pass (line layout)
    double tune_ms_[kTuned][kMaxVariants] = {};
    // transposed work layout (k_plines_t, XT x pass): in use, its second work
    // buffer, variants and times; ky-blocked layout (k_plines_b): in use,
    // the ky block, variants and times; layout_ms_ as reported by layout()
    bool xt_ = false;
    DeviceArray<cx<T>> X_;
    int xvar_[kTuned] = {0, 0, 0};
    double xtune_ms_[kTuned][kMaxVariants] = {};
    bool zb_ = false;
    bool xz_ = false;  // the x pass's fields [z][kx][y] (ky-blocked layout)
    int kb_ = 0;
    int bvar_[kTuned] = {0, 0, 0};
    double btune_ms_[kTuned][kMaxVariants] = {};
    double layout_ms_[4] = {0.0, 0.0, 0.0, 0.0};
};

// Per-step GPU time from event pairs, resolved lazily.
class StepTimer {
public:
    ~StepTimer() {
        for (auto& p : pending_) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
    std::pair<hipEvent_t, hipEvent_t> begin(hipStream_t s) {
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
        return {e0, e1};
    }
    void end(std::pair<hipEvent_t, hipEvent_t> p, hipStream_t s) {
        hip_check(hipEventRecord(p.second, s), "hipEventRecord");
        pending_.push_back(p);
        ++steps;
        if (pending_.size() > 64) resolve_one();
    }
    double seconds() {
        while (!pending_.empty()) resolve_one();
        return total_;
    }
    // GPU time of each of the last kKeep steps (ms), oldest first
    static constexpr std::size_t kKeep = 4096;
    const std::deque<float>& step_ms() {
        while (!pending_.empty()) resolve_one();
        return recent_;
    }
    long long steps = 0;

private:
    void resolve_one() {
        auto p = pending_.front();
        pending_.pop_front();
        hip_check(hipEventSynchronize(p.second), "hipEventSynchronize");
        float e = 0.f;
        hip_check(hipEventElapsedTime(&e, p.first, p.second), "hipEventElapsedTime");
        total_ += 1e-3 * e;
        recent_.push_back(e);
        if (recent_.size() > kKeep) recent_.pop_front();
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    std::deque<std::pair<hipEvent_t, hipEvent_t>> pending_;
    std::deque<float> recent_;
    double total_ = 0.0;
};

// A captured step: every stage of every group, in order.
class StepGraph {
public:
    ~StepGraph() { reset(); }
    template <class F>
    void launch(hipStream_t s, F&& record) {
        if (!exec_) {
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            record();
            hip_check(hipStreamEndCapture(s, &graph_), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
        }
        hip_check(hipGraphLaunch(exec_, s), "hipGraphLaunch");
    }
    void reset() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
    }

private:
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

// ---------------------------------------------------------------------------
// Taylor-Green solver (one member)
// ---------------------------------------------------------------------------
class SolverBase {
public:
    virtual ~SolverBase() = default;
    virtual void step(double dt) = 0;
    // One step whose dt = dt_of(max |u| of the current state); dt_of returns
    // a value <= 0 to take no step. Returns the dt used (or that value).
    virtual double step_cfl(const std::function<double(double)>& dt_of) = 0;
    virtual sph_diagnostics diagnostics(double t) = 0;
    virtual double max_velocity() = 0;
    virtual void get_uhat(int c, double* host) = 0;
    virtual void set_uhat(int c, const double* host) = 0;
    virtual void profile(sph_profile* out) = 0;
    // GPU time (ms, HIP events around each step) of up to `max` of the most
    // recent steps, oldest first; returns how many were written
    virtual int step_times(double* ms, int max) = 0;
    virtual void set_graph(bool on) = 0;
    virtual void synchronize() = 0;
    virtual void time_passes(int reps, double* ms, double* bytes) = 0;
    virtual void time_step_passes(double dt, int steps, double* ms, double* bytes) = 0;
    virtual void step_adaptive(double dt_max, double cfl, double* vmax, double* dt_used) = 0;
    virtual bool fused_velocity_max() const = 0;
    virtual void tuning(int* variant, int* lines, double* ms) const = 0;
    virtual void tuned_pass(int pass, int* variant, int* lines, double* ms) const = 0;
    virtual RadixPlan radix_plan_of(int pass) const = 0;
    virtual void layout(int* kind, double* ms) const = 0;
    int n = 0;
    std::int64_t modes = 0;
};

template <typename T>
class Solver final : public SolverBase {
public:
    Solver(int n_in, double reynolds) : g_(n_in, 1, {1.0 / reynolds}) {
        n = n_in;
        modes = g_.modes;
        g_.taylor_green(0, stream_.get());
        prof_.forward_calls += 3;
        hip_check(hipStreamSynchronize(stream_.get()), "TaylorGreenSolver initialisation");
    }
    ~Solver() override { (void)hipStreamSynchronize(stream_.get()); }

    void step(double dt) override {
        const hipStream_t s = stream_.get();
        g_.set_dt(&dt, s);
        auto ev = timer_.begin(s);
        g_.prepare_pairs(s);
        if (use_graph_) {
            graph_.launch(s, [&] {
                for (int st = 0; st < 3; ++st) g_.launch_stage(st, s);
            });
        } else {
            for (int st = 0; st < 3; ++st) g_.launch_stage(st, s);
            hip_check(hipGetLastError(), "TaylorGreenSolver::step");
        }
        timer_.end(ev, s);
        prof_.step_calls += 1;
        prof_.rhs_calls += 3;
        prof_.forward_calls += 36;
        prof_.backward_calls += 18;
    }

    // max |u| comes from stage 0's x pass (Group::launch_step_head) instead
    // of transforms of its own; the same physical field, formed by the x
    // pass's transforms (equal to max_velocity() to rounding). The host waits
    // for that value only: stage 0's forward y pass is queued behind its
    // copy, so the dt choice and the tail's launch overlap it on the GPU.
    double step_cfl(const std::function<double(double)>& dt_of) override {
        const hipStream_t s = stream_.get();
        if (!g_.x_reduces_velocity()) {
            const double dt = dt_of(max_velocity());
            if (dt > 0.0) step(dt);
            return dt;
        }
        auto ev = timer_.begin(s);
        g_.prepare_pairs(s);
        g_.launch_step_head(vmax2_.get(), s);
        hip_check(hipMemcpyAsync(vmax2_host_.get(), vmax2_.get(), kVmaxSlots * kVmaxStride * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, s),
                  "hipMemcpy vmax");
        hip_check(hipEventRecord(vmax_ready_.get(), s), "hipEventRecord");
        g_.launch_step_mid(s);
        hip_check(hipGetLastError(), "TaylorGreenSolver::step_cfl");
        hip_check(hipEventSynchronize(vmax_ready_.get()), "step_cfl");
        unsigned long long bits = 0;
        for (int k = 0; k < kVmaxSlots; ++k) bits = std::max(bits, vmax2_host_.get()[k * kVmaxStride]);
        double vmax2 = 0.0;
        std::memcpy(&vmax2, &bits, sizeof(vmax2));
        prof_.max_velocity_calls += 1;
        prof_.backward_calls += 3;
        const double dt = dt_of(std::sqrt(vmax2));
        if (!(dt > 0.0)) {
            // the head and the forward y pass consumed the velocity pairs; no
            // step is taken
            g_.invalidate_pairs();
            timer_.end(ev, s);
            return dt;
        }
        g_.set_dt(&dt, s);
        if (use_graph_)
            tail_graph_.launch(s, [&] { g_.launch_step_tail(s); });
        else
            g_.launch_step_tail(s);
        hip_check(hipGetLastError(), "TaylorGreenSolver::step_cfl");
        timer_.end(ev, s);
        prof_.step_calls += 1;
        prof_.rhs_calls += 3;
        prof_.forward_calls += 36;
        prof_.backward_calls += 18;
        return dt;
    }

    sph_diagnostics diagnostics(double t) override {
        const auto t0 = std::chrono::steady_clock::now();
        const sph_diagnostics d = g_.diagnostics(0, t, stream_.get());
        prof_.diagnostics += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        prof_.diagnostics_calls += 1;
        prof_.backward_calls += 6;
        return d;
    }

    double max_velocity() override {
        const auto t0 = std::chrono::steady_clock::now();
        const double v = g_.max_velocity(0, stream_.get());
        prof_.max_velocity += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        prof_.max_velocity_calls += 1;
        prof_.backward_calls += 3;
        return v;
    }

    void get_uhat(int c, double* host) override { g_.get_uhat(0, c, host, stream_.get()); }
    void set_uhat(int c, const double* host) override { g_.set_uhat(0, c, host, stream_.get()); }

    void profile(sph_profile* out) override {
        prof_.step = timer_.seconds();
        *out = prof_;
    }
    int step_times(double* ms, int max) override {
        const std::deque<float>& r = timer_.step_ms();
        const int n = static_cast<int>(std::min<std::size_t>(r.size(), max > 0 ? static_cast<std::size_t>(max) : 0));
        for (int i = 0; i < n; ++i) ms[i] = r[r.size() - n + i];
        return n;
    }
    void set_graph(bool on) override { use_graph_ = on; }
    void synchronize() override { hip_check(hipStreamSynchronize(stream_.get()), "hipStreamSynchronize"); }
    void time_passes(int reps, double* ms, double* bytes) override { g_.time_passes(reps, ms, bytes, stream_.get()); }

    // `steps` real RK3 steps of dt with the pass kernels launched directly
    // (no graph) and a HIP event between consecutive launches on the solver's
    // stream: ms[k] = mean duration of pass k inside the steps (k = 1..4, over
    // steps x 3 stages; ms[0] = the standalone inverse z if the velocity
    // pairs had to be formed first, else 0), ms[5] = mean step time, bytes[k]
    // = algorithmic bytes per launch (Group::time_passes). The state advances.
    void time_step_passes(double dt, int steps, double* ms, double* bytes) override {
        const hipStream_t s = stream_.get();
        if (steps < 1) steps = 1;
        g_.set_dt(&dt, s);
        constexpr int K = Group<T>::kPasses;
        const int per_step = 3 * (K - 1);
        std::vector<hipEvent_t> ev(static_cast<std::size_t>(steps) * per_step + 3);
        for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        std::size_t q = 0;
        hip_check(hipEventRecord(ev[q++], s), "hipEventRecord");
        const bool prep = !g_.pairs_valid();
        g_.prepare_pairs(s);
        hip_check(hipEventRecord(ev[q++], s), "hipEventRecord");
        for (int k = 0; k < steps; ++k)
            for (int st = 0; st < 3; ++st)
                for (int p = 1; p < K; ++p) {
                    g_.launch_pass(p, st, g_.U(), g_.G(), 0, 1, s);
                    hip_check(hipEventRecord(ev[q++], s), "hipEventRecord");
                }
        hip_check(hipGetLastError(), "time_step_passes");
        hip_check(hipEventSynchronize(ev[q - 1]), "hipEventSynchronize");
        auto span = [&](std::size_t a, std::size_t b) {
            float e = 0.f;
            hip_check(hipEventElapsedTime(&e, ev[a], ev[b]), "hipEventElapsedTime");
            return static_cast<double>(e);
        };
        for (int p = 0; p <= K; ++p) ms[p] = 0.0;
        ms[0] = prep ? span(0, 1) : 0.0;
        std::size_t e = 1;
        for (int k = 0; k < steps; ++k)
            for (int st = 0; st < 3; ++st)
                for (int p = 1; p < K; ++p, ++e) ms[p] += span(e, e + 1) / (3.0 * steps);
        ms[K] = span(1, e) / steps;
        for (auto& x : ev) (void)hipEventDestroy(x);
        g_.pass_bytes(bytes);
        prof_.step_calls += steps;
        prof_.rhs_calls += 3 * steps;
        prof_.forward_calls += 36 * steps;
        prof_.backward_calls += 18 * steps;
    }

    void tuning(int* variant, int* lines, double* ms) const override { g_.tuning(variant, lines, ms); }
    void tuned_pass(int pass, int* variant, int* lines, double* ms) const override {
        g_.tuned_pass(pass, variant, lines, ms);
    }
    RadixPlan radix_plan_of(int pass) const override { return g_.radix_plan_of(pass); }
    void layout(int* kind, double* ms) const override { g_.layout(kind, ms); }

    // One adaptive step as the driver takes it (impl.hpp:1363-1378, without
    // the output-time clipping): dt = min(dt_max, cfl * (2 pi / n) / max|u|),
    // max|u| from the step's own x pass where that form runs (step_cfl).
    void step_adaptive(double dt_max, double cfl, double* vmax, double* dt_used) override {
        const double h = 2.0 * kPi / static_cast<double>(n);
        double v = 0.0;
        const double dt = step_cfl([&](double vm) {
            v = vm;
            double d = dt_max;
            if (vm > 1.0e-14) d = std::min(d, cfl * h / vm);
            return d;
        });
        *vmax = v;
        *dt_used = dt;
    }
    bool fused_velocity_max() const override { return g_.x_reduces_velocity(); }

private:
    Stream stream_;
    Group<T> g_;
    bool use_graph_ = true;
    StepGraph graph_;
    // step_cfl: the rest of a step after its forward y pass of stage 0, the
    // max |u|^2 slots on the device and their pinned host copy, the copy's
    // event
    StepGraph tail_graph_;
    DeviceArray<unsigned long long> vmax2_{kVmaxSlots * kVmaxStride};
    PinnedHost<unsigned long long> vmax2_host_{kVmaxSlots * kVmaxStride};
    Event vmax_ready_;
    StepTimer timer_;
    sph_profile prof_{};
};

// ---------------------------------------------------------------------------
// Batched ensemble of independent solves of uneven sizes
// ---------------------------------------------------------------------------
class EnsembleBase {
public:
    virtual ~EnsembleBase() = default;
    virtual void step(const double* dt) = 0;
    virtual sph_diagnostics diagnostics(int member, double t) = 0;
    virtual double max_velocity(int member) = 0;
    virtual void get_uhat(int member, int c, double* host) = 0;
    virtual void set_uhat(int member, int c, const double* host) = 0;
    virtual void set_graph(bool on) = 0;
    virtual void synchronize() = 0;
    virtual double step_seconds(long long* steps) = 0;
    virtual void time_passes(int reps, double* ms, double* bytes) = 0;
    virtual void launch_mode(int* mode, double* ms) const = 0;
    std::vector<int> sizes;
};

template <typename T>
class Ensemble final : public EnsembleBase {
public:
    Ensemble(int count, const int* n, const double* nu, const std::uint64_t* seeds, double energy) {
        if (count < 1) throw std::runtime_error("ensemble needs at least one member");
        sizes.assign(n, n + count);
        // one group per distinct size, members in order of appearance
        for (int j = 0; j < count; ++j) {
            if (n[j] <= 0) throw std::runtime_error("grid size must be positive");
            if (!(nu[j] > 0.0)) throw std::runtime_error("viscosity must be positive");
            if (!pass_table<T>(n[j])) throw std::runtime_error("unsupported transform length " + std::to_string(n[j]));
            int gi = -1;
            for (std::size_t q = 0; q < group_n_.size(); ++q)
                if (group_n_[q] == n[j]) gi = static_cast<int>(q);
            if (gi < 0) {
                gi = static_cast<int>(group_n_.size());
                group_n_.push_back(n[j]);
                group_nu_.emplace_back();
            }
            where_.push_back({gi, static_cast<int>(group_nu_[gi].size())});
            group_nu_[gi].push_back(nu[j]);
        }
        for (std::size_t q = 0; q < group_n_.size(); ++q)
            groups_.push_back(std::make_unique<Group<T>>(group_n_[q], static_cast<int>(group_nu_[q].size()),
                                                         group_nu_[q]));
        dt_.resize(groups_.size());
        for (std::size_t q = 0; q < groups_.size(); ++q) dt_[q].assign(groups_[q]->count, 0.0);
        make_branches();
        // the states are still zero: time the two launch structures on them
        make_fused();
        for (int j = 0; j < count; ++j) {
            const int nj = n[j];
            const std::size_t modes = static_cast<std::size_t>(nj / 2 + 1) * nj * nj;
            std::vector<double> u0(2 * modes), u1(2 * modes), u2(2 * modes);
            random_solenoidal_field(nj, seeds[j], energy, u0.data(), u1.data(), u2.data());
            const double* u[3] = {u0.data(), u1.data(), u2.data()};
            const auto w = locate(j);
            groups_[w.first]->set_state(w.second, u, stream_.get());
        }
    }

    void step(const double* dt) override {
        const hipStream_t s = stream_.get();
        for (std::size_t j = 0; j < where_.size(); ++j) dt_[where_[j].first][where_[j].second] = dt[j];
        for (std::size_t q = 0; q < groups_.size(); ++q) groups_[q]->set_dt(dt_[q].data(), s);
        auto ev = timer_.begin(s);
        for (auto& g : groups_) g->prepare_pairs(s);
        auto record = [&] { record_step(s); };
        // the cooperative per-stage launch is not captured into the graph
        if (use_graph_ && mode_ != 2) {
            graph_.launch(s, record);
        } else {
            record();
            hip_check(hipGetLastError(), "ensemble step");
        }
        timer_.end(ev, s);
    }

    sph_diagnostics diagnostics(int member, double t) override {
        const auto w = locate(member);
        return groups_[w.first]->diagnostics(w.second, t, stream_.get());
    }
    void launch_mode(int* mode, double* ms) const override {
        *mode = mode_;
        for (int i = 0; i < 3; ++i) ms[i] = mode_ms_[i];
    }
    double max_velocity(int member) override {
        const auto w = locate(member);
        return groups_[w.first]->max_velocity(w.second, stream_.get());
    }
    void get_uhat(int member, int c, double* host) override {
        const auto w = locate(member);
        groups_[w.first]->get_uhat(w.second, c, host, stream_.get());
    }
    void set_uhat(int member, int c, const double* host) override {
        const auto w = locate(member);
        groups_[w.first]->set_uhat(w.second, c, host, stream_.get());
    }
    void set_graph(bool on) override { use_graph_ = on; }
    void synchronize() override { hip_check(hipStreamSynchronize(stream_.get()), "hipStreamSynchronize"); }
    double step_seconds(long long* steps) override {
        if (steps) *steps = timer_.steps;
        return timer_.seconds();
    }
    // Per pass, summed over the groups (one launch per pass, stage and group).
    void time_passes(int reps, double* ms, double* bytes) override {
        constexpr int K = Group<T>::kPasses;
        for (int k = 0; k < K; ++k) ms[k] = bytes[k] = 0.0;
        for (auto& g : groups_) {
            double mk[K], bk[K];
            g->time_passes(reps, mk, bk, stream_.get());
            for (int k = 0; k < K; ++k) {
                ms[k] += mk[k];
                bytes[k] += bk[k];
            }
        }
    }

private:
    std::pair<int, int> locate(int member) const {
        if (member < 0 || member >= static_cast<int>(where_.size())) throw std::runtime_error("member out of range");
        return where_[member];
    }

    // The three stages of a step. Groups in the one-launch-per-pass set run as
    // four launches per stage on the main stream; every other group runs its
    // own chain of four launches per stage on a branch of its own (fork/join
    // through events), so small grids run beside large ones either way.
    void record_step(hipStream_t s) {
        std::vector<std::size_t> own;
        for (std::size_t q = 0; q < groups_.size(); ++q)
            if (!(mode_ > 0 && is_fused(q))) own.push_back(q);
        const bool fused = mode_ > 0 && !fused_.empty();
        if (!fused && own.size() == 1) {
            for (int st = 0; st < 3; ++st) groups_[own[0]]->launch_stage(st, s);
            return;
        }
        if (!own.empty()) {
            hip_check(hipEventRecord(fork_, s), "hipEventRecord");
            for (std::size_t q : own) {
                const hipStream_t b = branch_[q]->get();
                hip_check(hipStreamWaitEvent(b, fork_, 0), "hipStreamWaitEvent");
                for (int st = 0; st < 3; ++st) groups_[q]->launch_stage(st, b);
                hip_check(hipEventRecord(join_[q], b), "hipEventRecord");
            }
        }
        if (fused && mode_ == 2)
            for (int st = 0; st < 3; ++st) ens_launch_stage<T>(table_.get(), stage_grids_, stage_blocks_, st, s);
        else if (fused)
            for (int st = 0; st < 3; ++st)
                for (int k = 1; k < Group<T>::kPasses; ++k) ens_launch<T>(k, table_.get(), grid_[k], st, s);
        for (std::size_t q : own) hip_check(hipStreamWaitEvent(s, join_[q], 0), "hipStreamWaitEvent");
    }

    bool is_fused(std::size_t q) const { return std::find(fused_.begin(), fused_.end(), q) != fused_.end(); }

    // One launch per pass for the groups of the sizes ensemble_passes.hpp
    // fuses (largest first), and the choice between that and one chain per
    // group: SPH_ENSEMBLE_FUSED=0 / 1 forces it, otherwise both are timed on
    // the still-zero states (one warm-up and two timed steps each, launched
    // directly) and the faster is kept.
    void make_fused() {
        for (std::size_t q = 0; q < groups_.size(); ++q)
            if (ens_size(groups_[q]->n) && !groups_[q]->transposed_layout() &&
                static_cast<int>(fused_.size()) < kEnsGroups)
                fused_.push_back(q);
        if (fused_.empty()) return;
        stage_blocks_ = ens_stage_blocks<T>();
        std::sort(fused_.begin(), fused_.end(), [&](std::size_t a, std::size_t b) { return groups_[a]->n > groups_[b]->n; });
        std::vector<EnsGroupArgs<T>> args;
        std::vector<int> fn, fm;
        for (std::size_t q : fused_) {
            args.push_back(groups_[q]->ens_args());
            fn.push_back(groups_[q]->n);
            fm.push_back(groups_[q]->count);
        }
        table_.resize(args.size());
        hip_check(hipMemcpy(table_.get(), args.data(), args.size() * sizeof(args[0]), hipMemcpyHostToDevice),
                  "hipMemcpy ensemble table");
        for (int k = 0; k < Group<T>::kPasses; ++k)
            grid_[k] = ens_grid<T>(k, static_cast<int>(fused_.size()), fn.data(), fm.data());
        for (int k = 1; k < Group<T>::kPasses; ++k) stage_grids_.pass[k - 1] = grid_[k];
        // SPH_ENSEMBLE_FUSED=0 / 1 / 2 forces per-size chains / one launch
        // per pass / one launch per stage (where cooperative launches exist)
        const char* env = std::getenv("SPH_ENSEMBLE_FUSED");
        if (env && env[0] >= '0' && env[0] <= '2' && env[1] == '\0') {
            mode_ = env[0] - '0';
            if (mode_ == 2 && stage_blocks_ < 8) throw std::runtime_error("one launch per stage needs cooperative launches");
            return;
        }
        const hipStream_t s = stream_.get();
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        for (int mode = 0; mode < 3; ++mode) {
            if (mode == 2 && stage_blocks_ < 8) break;
            mode_ = mode;
            record_step(s);
            hip_check(hipEventRecord(a, s), "hipEventRecord");
            for (int r = 0; r < 2; ++r) record_step(s);
            hip_check(hipEventRecord(b, s), "hipEventRecord");
            hip_check(hipEventSynchronize(b), "hipEventSynchronize");
            float e = 0.f;
            hip_check(hipEventElapsedTime(&e, a, b), "hipEventElapsedTime");
            mode_ms_[mode] = e / 2.0;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        hip_check(hipGetLastError(), "ensemble launch timing");
        mode_ = 0;
        for (int mode = 1; mode < 3; ++mode)
            if (mode_ms_[mode] > 0.0 && mode_ms_[mode] < mode_ms_[mode_]) mode_ = mode;
    }

    void make_branches() {
        if (groups_.size() < 2) return;
        hip_check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
        join_.assign(groups_.size(), nullptr);
        for (auto& e : join_) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        for (std::size_t q = 0; q < groups_.size(); ++q) branch_.push_back(std::make_unique<Stream>());
    }

    Stream stream_;
    std::vector<int> group_n_;
    std::vector<std::vector<double>> group_nu_;
    std::vector<std::unique_ptr<Group<T>>> groups_;
    std::vector<std::pair<int, int>> where_;  // member -> (group, index in group)
    std::vector<std::vector<double>> dt_;
    bool use_graph_ = true;
    StepGraph graph_;
    StepTimer timer_;
    // fork/join of the per-size branches (default-priority streams)
    std::vector<std::unique_ptr<Stream>> branch_;
    hipEvent_t fork_ = nullptr;
    std::vector<hipEvent_t> join_;
    // one launch per pass: the fused groups (indices into groups_), their
    // argument table on the device, the grid of each pass, whether in use,
    // and the timed step of each launch structure (ms: per-group chains,
    // one launch per pass, one launch per stage); mode_ the one in use (0,
    // 1, 2), stage_blocks_ the resident grid of the per-stage launch
    std::vector<std::size_t> fused_;
    DeviceArray<EnsGroupArgs<T>> table_;
    EnsGrid grid_[Group<T>::kPasses] = {};
    EnsStageGrids stage_grids_{};
    int stage_blocks_ = 0;
    int mode_ = 0;
    double mode_ms_[3] = {0.0, 0.0, 0.0};

public:
    ~Ensemble() override {
        (void)hipStreamSynchronize(stream_.get());
        for (auto& b : branch_) (void)hipStreamSynchronize(b->get());
        if (fork_) (void)hipEventDestroy(fork_);
        for (auto e : join_)
            if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace sph

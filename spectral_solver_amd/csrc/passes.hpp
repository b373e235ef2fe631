// Pass kernels of the 3D real<->complex transforms, with the pointwise work of
// the RK3 right-hand side folded into the first and last passes.
//
// Storage (per field):
//   physical  real  [z][y][x]            offset (z*ny + y)*nx + x
//   spectral  cx    [kx][kz][ky]         offset (kx*nz + kz)*ny + ky,
//             kx in [0, nx/2] (half spectrum), ky and kz in [0, n).
//   mixed     the same spectral layout with some axes still in physical space.
//
// Passes (inverse = spectral -> physical, forward = physical -> spectral):
//   inv_z  : z lines of a tile of NL adjacent ky at fixed kx (128-byte row
//            segments); RHS mode emits u_hat and the half-cell-shifted
//            u_hat*phi (src/ava_taylor_green_impl.hpp:866-873) for 3
//            components -> 6 fields; DIAG mode emits u_hat and
//            omega_hat = i k x u_hat (impl.hpp:1147-1157).
//   lines  : y lines, contiguous, in place, any number of fields.
//   inv_x  : x lines of NL adjacent y rows at fixed z; two half spectra A, B
//            are packed as Z = A + iB with Hermitian extension (taking Re of
//            the kx = 0 and kx = n/2 planes, which is what Re() of the
//            reference's C2C backward keeps), inverse-transformed, and split
//            into the real rows a = Re z, b = Im z, times 1/N^3
//            (impl.hpp:555-561).
//   fwd_x  : PRODUCTS mode forms the six products u_a*u_b and the six
//            shifted products s_a*s_b (impl.hpp:903-906) in registers, packs
//            each pair as p + i q, transforms, and separates the two half
//            spectra: P = (Z_k + conj Z_{n-k})/2, Q = (Z_k - conj Z_{n-k})/(2i).
//   fwd_z  : RK mode transforms the 12 product spectra along z and, per mode,
//            applies the dealias combine and mask (impl.hpp:926-936), the -ik
//            accumulation in the reference's order (00,01,02,11,12,22)
//            (impl.hpp:881-886), the viscous term and RHS projection
//            (impl.hpp:955-974), the Williamson update (impl.hpp:994-1033) and
//            project_truncate of u_hat (impl.hpp:819-845), writing u_hat and g.
#pragma once

#include "fft_device.hpp"

namespace sph {

enum InvZMode { IZ_PLAIN = 0, IZ_RHS = 1, IZ_DIAG = 2 };
enum FwdXMode { FX_PLAIN = 0, FX_PRODUCTS = 1 };
enum FwdZMode { FZ_PLAIN = 0, FZ_PROJECT = 1, FZ_RK = 2 };

__device__ __forceinline__ int signed_mode(int i, int n) { return i <= n / 2 ? i : i - n; }

// keep (mx^2+my^2+mz^2)/(n/2)^2 <= 1, evaluated exactly in integers
// (impl.hpp:131-134 evaluates it in double; both agree for every integer input).
__device__ __forceinline__ bool keep_mode(int k2, int n) {
    return 4LL * k2 <= static_cast<long long>(n) * n;
}

template <typename T>
struct InvZArgs {
    const cx<T>* in[3];
    cx<T>* out[6];
    int nfields;         // PLAIN: in[f] -> out[f]
    int ny;              // number of ky values (tile dimension)
    int n;               // cube size for modes and phase (RHS, DIAG)
    const cx<T>* phase;  // e^{i pi s/n}, indexed by s + n, s = mx+my+mz
    const cx<T>* tw;     // twiddles of the z length
};

template <typename T>
struct InvXArgs {
    const cx<T>* a[3];
    const cx<T>* b[3];  // may be null (treated as 0)
    T* oa[3];
    T* ob[3];  // may be null (discarded)
    int npairs;
    int ny;
    int nz;
    T scale;
    const cx<T>* tw;
};

template <typename T>
struct FwdXArgs {
    const T* in[6];
    cx<T>* out[12];  // PRODUCTS: pair p -> out[2p] (plain), out[2p+1] (shifted)
    int nin;         // PLAIN: 1..3 inputs, pairs (in0,in1) (in2,0) -> out[0..nin)
    int ny;
    int nz;
    const cx<T>* tw;
};

template <typename T>
struct FwdZArgs {
    const cx<T>* in[12];
    cx<T>* out[3];  // PLAIN / PROJECT outputs
    cx<T>* u[3];    // RK: state (read-write)
    cx<T>* g[3];    // RK: low-storage register (read-write)
    int nfields;    // PLAIN
    int ny;
    int n;
    T nu;
    const T* coef;  // RK: alpha_s, beta_s*dt per stage, in device memory
    int stage;      // RK: 0, 1, 2
    const cx<T>* phase;
    const cx<T>* tw;
};

// ---------------------------------------------------------------------------
// inverse z pass
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE>
__global__ void __launch_bounds__((Tile<T, N>::NT)) k_inv_z(InvZArgs<T> a) {
    using TL = Tile<T, N>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, S = TL::STRIDE;
    __shared__ cx<T> buf[TL::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const int ky0 = blockIdx.x * NL;
    const int kx = blockIdx.y;
    load_twiddles<T, N, NT>(twl, a.tw, tid);

    // line-fast items: thread owns line l and elements e0 + i*STEP
    static_assert(NT % NL == 0, "tile threads must cover whole rows");
    constexpr int STEP = NT / NL;
    const int l = tid % NL, e0 = tid / NL;
    const bool lane_ok = ky0 + l < a.ny;
    const int off0 = (kx * N + e0) * a.ny + ky0 + l;
    const int stride_off = STEP * a.ny;
    const int li0 = e0 * S + l;
#define SPH_IN(i) ((TL::ELEMS % NT == 0) || (e0 + (i) * STEP < N))
#define SPH_OK(i) (lane_ok && SPH_IN(i))
#define SPH_OFF(i) (off0 + (i) * stride_off)
#define SPH_LI(i) (SPH_IN(i) ? li0 + (i) * STEP * S : 0)
    if constexpr (MODE == IZ_PLAIN) {
        for (int f = 0; f < a.nfields; ++f) {
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_IN(i)) buf[SPH_LI(i)] = SPH_OK(i) ? a.in[f][SPH_OFF(i)] : czero<T>();
            __syncthreads();
            fft_tile<T, N, NL, NT, true>(buf, twl, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_OK(i)) a.out[f][SPH_OFF(i)] = buf[SPH_LI(i)];
            __syncthreads();
        }
    } else if constexpr (MODE == IZ_RHS) {
#pragma unroll 1
        for (int c = 0; c < 3; ++c) {
            cx<T> u[E];
#pragma unroll
            for (int i = 0; i < E; ++i) {
                u[i] = SPH_OK(i) ? a.in[c][SPH_OFF(i)] : czero<T>();
                if (SPH_IN(i)) buf[SPH_LI(i)] = u[i];
            }
            __syncthreads();
            fft_tile<T, N, NL, NT, true>(buf, twl, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_OK(i)) a.out[2 * c][SPH_OFF(i)] = buf[SPH_LI(i)];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int idx = tid + i * NT;
                const int my = signed_mode(ky0 + idx % NL, a.n), mz = signed_mode(idx / NL, a.n);
                const cx<T> ph = SPH_OK(i) ? a.phase[kx + my + mz + a.n] : czero<T>();
                if (SPH_IN(i)) buf[SPH_LI(i)] = cmul(u[i], ph);
            }
            __syncthreads();
            fft_tile<T, N, NL, NT, true>(buf, twl, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_OK(i)) a.out[2 * c + 1][SPH_OFF(i)] = buf[SPH_LI(i)];
            __syncthreads();
        }
    } else {  // IZ_DIAG
        cx<T> u0[E], u1[E], u2[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            u0[i] = SPH_OK(i) ? a.in[0][SPH_OFF(i)] : czero<T>();
            u1[i] = SPH_OK(i) ? a.in[1][SPH_OFF(i)] : czero<T>();
            u2[i] = SPH_OK(i) ? a.in[2][SPH_OFF(i)] : czero<T>();
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_IN(i)) buf[SPH_LI(i)] = c == 0 ? u0[i] : (c == 1 ? u1[i] : u2[i]);
            __syncthreads();
            fft_tile<T, N, NL, NT, true>(buf, twl, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_OK(i)) a.out[2 * c][SPH_OFF(i)] = buf[SPH_LI(i)];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int idx = tid + i * NT;
                const T ix = T(kx), iy = T(signed_mode(ky0 + idx % NL, a.n)), iz = T(signed_mode(idx / NL, a.n));
                cx<T> w;
                if (c == 0) {
                    w = {-(iy * u2[i].y - iz * u1[i].y), iy * u2[i].x - iz * u1[i].x};
                } else if (c == 1) {
                    w = {-(iz * u0[i].y - ix * u2[i].y), iz * u0[i].x - ix * u2[i].x};
                } else {
                    w = {-(ix * u1[i].y - iy * u0[i].y), ix * u1[i].x - iy * u0[i].x};
                }
                if (SPH_IN(i)) buf[SPH_LI(i)] = w;
            }
            __syncthreads();
            fft_tile<T, N, NL, NT, true>(buf, twl, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_OK(i)) a.out[2 * c + 1][SPH_OFF(i)] = buf[SPH_LI(i)];
            __syncthreads();
        }
    }
}

#undef SPH_IN
#undef SPH_OK
#undef SPH_OFF
#undef SPH_LI

// ---------------------------------------------------------------------------
// contiguous lines (y pass), in place
// ---------------------------------------------------------------------------
template <typename T, int N, bool INV>
__global__ void __launch_bounds__((Tile<T, N>::NT)) k_lines(cx<T>* data, long long nlines, const cx<T>* tw) {
    using TL = Tile<T, N>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, S = TL::STRIDE;
    __shared__ cx<T> buf[TL::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const long long line0 = static_cast<long long>(blockIdx.x) * NL;
    cx<T>* base = data + line0 * N;
    const int valid_lines = static_cast<int>(nlines - line0 < NL ? nlines - line0 : NL);
    load_twiddles<T, N, NT>(twl, tw, tid);
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int idx = tid + i * NT;
        const int e = idx % N, l = idx / N;
        if (idx < TL::ELEMS) buf[e * S + l] = l < valid_lines ? base[idx] : czero<T>();
    }
    __syncthreads();
    fft_tile<T, N, NL, NT, INV>(buf, twl, tid);
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int idx = tid + i * NT;
        const int e = idx % N, l = idx / N;
        if (idx < TL::ELEMS && l < valid_lines) base[idx] = buf[e * S + l];
    }
}

// ---------------------------------------------------------------------------
// inverse x pass (complex-to-real, two fields per complex transform)
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__((Tile<T, N>::NT)) k_inv_x(InvXArgs<T> a) {
    using TL = Tile<T, N>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, S = TL::STRIDE;
    constexpr int H = N / 2 + 1;
    constexpr int LOADS = NL * H;
    constexpr int LPT = (LOADS + NT - 1) / NT;
    __shared__ cx<T> buf[TL::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * NL;
    const int z = blockIdx.y;
    load_twiddles<T, N, NT>(twl, a.tw, tid);
#pragma unroll 1
    for (int p = 0; p < a.npairs; ++p) {
        const cx<T>* A = a.a[p];
        const cx<T>* B = a.b[p];
#pragma unroll
        for (int i = 0; i < LPT; ++i) {
            const int idx = tid + i * NT;
            if (LOADS % NT == 0 || idx < LOADS) {
                const int l = idx % NL, k = idx / NL;
                cx<T> va = czero<T>(), vb = czero<T>();
                if (y0 + l < a.ny) {
                    const int off = (k * a.nz + z) * a.ny + y0 + l;
                    va = A[off];
                    if (B) vb = B[off];
                }
                if (k == 0 || k == N / 2) {
                    va.y = T(0);
                    vb.y = T(0);
                }
                buf[k * S + l] = {va.x - vb.y, va.y + vb.x};
                if (k > 0 && k < N / 2) buf[(N - k) * S + l] = {va.x + vb.y, vb.x - va.y};
            }
        }
        __syncthreads();
        fft_tile<T, N, NL, NT, true>(buf, twl, tid);
        T* OA = a.oa[p];
        T* OB = a.ob[p];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int idx = tid + i * NT;
            const int x = idx % N, l = idx / N;
            if (idx < TL::ELEMS && y0 + l < a.ny) {
                const cx<T> v = buf[x * S + l];
                const int o = (z * a.ny + y0 + l) * N + x;
                OA[o] = v.x * a.scale;
                if (OB) OB[o] = v.y * a.scale;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// forward x pass (real-to-complex, two real rows per complex transform)
// ---------------------------------------------------------------------------
template <typename T, int N>
__device__ __forceinline__ void fwd_x_separate_store(const cx<T>* buf, cx<T>* P, cx<T>* Q, int y0, int z, int ny,
                                                     int nz, int tid) {
    using TL = Tile<T, N>;
    constexpr int NL = TL::NL, NT = TL::NT, S = TL::STRIDE;
    constexpr int H = N / 2 + 1;
    constexpr int ITEMS = NL * H;
    constexpr int IPT = (ITEMS + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int idx = tid + i * NT;
        if (ITEMS % NT == 0 || idx < ITEMS) {
            const int l = idx % NL, k = idx / NL;
            if (y0 + l < ny) {
                const cx<T> zk = buf[k * S + l];
                const cx<T> zm = buf[((N - k) % N) * S + l];
                const int off = (k * nz + z) * ny + y0 + l;
                P[off] = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
                if (Q) Q[off] = {T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
            }
        }
    }
}

template <typename T, int N, int MODE>
__global__ void __launch_bounds__((Tile<T, N>::NT)) k_fwd_x(FwdXArgs<T> a) {
    using TL = Tile<T, N>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, S = TL::STRIDE;
    __shared__ cx<T> buf[TL::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * NL;
    const int z = blockIdx.y;
    load_twiddles<T, N, NT>(twl, a.tw, tid);

    int off[E], li[E];
    bool ok[E], in[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int idx = tid + i * NT;
        const int x = idx % N, l = idx / N;
        in[i] = (TL::ELEMS % NT == 0) || (idx < TL::ELEMS);
        ok[i] = in[i] && (y0 + l < a.ny);
        off[i] = (z * a.ny + y0 + l) * N + x;
        li[i] = in[i] ? x * S + l : 0;
    }
    if constexpr (MODE == FX_PRODUCTS) {
        T u[3][E], s[3][E];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                u[c][i] = ok[i] ? a.in[c][off[i]] : T(0);
                s[c][i] = ok[i] ? a.in[3 + c][off[i]] : T(0);
            }
        }
        constexpr int PA[6] = {0, 0, 0, 1, 1, 2};
        constexpr int PB[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int p = 0; p < 6; ++p) {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                if (in[i]) buf[li[i]] = {u[PA[p]][i] * u[PB[p]][i], s[PA[p]][i] * s[PB[p]][i]};
            }
            __syncthreads();
            fft_tile<T, N, NL, NT, false>(buf, twl, tid);
            fwd_x_separate_store<T, N>(buf, a.out[2 * p], a.out[2 * p + 1], y0, z, a.ny, a.nz, tid);
            __syncthreads();
        }
    } else {  // FX_PLAIN
#pragma unroll 1
        for (int p = 0; 2 * p < a.nin; ++p) {
            const T* fa = a.in[2 * p];
            const T* fb = 2 * p + 1 < a.nin ? a.in[2 * p + 1] : nullptr;
#pragma unroll
            for (int i = 0; i < E; ++i) {
                if (in[i]) buf[li[i]] = {ok[i] ? fa[off[i]] : T(0), (ok[i] && fb) ? fb[off[i]] : T(0)};
            }
            __syncthreads();
            fft_tile<T, N, NL, NT, false>(buf, twl, tid);
            fwd_x_separate_store<T, N>(buf, a.out[2 * p], fb ? a.out[2 * p + 1] : nullptr, y0, z, a.ny, a.nz,
                                       tid);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// forward z pass (+ fused RHS completion and RK3 update)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void project_or_zero(cx<T>& f0, cx<T>& f1, cx<T>& f2, int ix, int iy, int iz, int k2,
                                                bool keep) {
    const T rk2 = T(k2);
    if (k2 > 0 && keep) {
        const T fx = T(ix), fy = T(iy), fz = T(iz);
        const cx<T> dot = (f0 * fx + f1 * fy) + f2 * fz;
        const cx<T> fac = dot * (T(1) / rk2);
        f0 = f0 - fac * fx;
        f1 = f1 - fac * fy;
        f2 = f2 - fac * fz;
    } else {
        f0 = f1 = f2 = czero<T>();
    }
}

template <typename T, int N, int MODE>
__global__ void __launch_bounds__((Tile<T, N>::NT)) k_fwd_z(FwdZArgs<T> a) {
    using TL = Tile<T, N>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, S = TL::STRIDE;
    __shared__ cx<T> buf[TL::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const int ky0 = blockIdx.x * NL;
    const int kx = blockIdx.y;
    load_twiddles<T, N, NT>(twl, a.tw, tid);

    // line-fast items: thread owns line l and elements e0 + i*STEP
    static_assert(NT % NL == 0, "tile threads must cover whole rows");
    constexpr int STEP = NT / NL;
    const int l = tid % NL, e0 = tid / NL;
    const bool lane_ok = ky0 + l < a.ny;
    const int off0 = (kx * N + e0) * a.ny + ky0 + l;
    const int stride_off = STEP * a.ny;
    const int li0 = e0 * S + l;
#define SPH_IN(i) ((TL::ELEMS % NT == 0) || (e0 + (i) * STEP < N))
#define SPH_OK(i) (lane_ok && SPH_IN(i))
#define SPH_OFF(i) (off0 + (i) * stride_off)
#define SPH_LI(i) (SPH_IN(i) ? li0 + (i) * STEP * S : 0)
    auto transform = [&](const cx<T>* src) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            if (SPH_IN(i)) buf[SPH_LI(i)] = SPH_OK(i) ? src[SPH_OFF(i)] : czero<T>();
        }
        __syncthreads();
        fft_tile<T, N, NL, NT, false>(buf, twl, tid);
    };

    if constexpr (MODE == FZ_PLAIN) {
#pragma unroll 1
        for (int f = 0; f < a.nfields; ++f) {
            transform(a.in[f]);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (SPH_OK(i)) a.out[f][SPH_OFF(i)] = buf[SPH_LI(i)];
            __syncthreads();
        }
    } else if constexpr (MODE == FZ_PROJECT) {
        cx<T> v[3][E];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            transform(a.in[c]);
#pragma unroll
            for (int i = 0; i < E; ++i) v[c][i] = buf[SPH_LI(i)];
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < E; ++i) {
            if (!SPH_OK(i)) continue;
            const int idx = tid + i * NT;
            const int ix = kx, iy = signed_mode(ky0 + idx % NL, a.n), iz = signed_mode(idx / NL, a.n);
            const int k2 = ix * ix + iy * iy + iz * iz;
            project_or_zero(v[0][i], v[1][i], v[2][i], ix, iy, iz, k2, keep_mode(k2, a.n));
#pragma unroll
            for (int c = 0; c < 3; ++c) a.out[c][SPH_OFF(i)] = v[c][i];
        }
    } else {  // FZ_RK
        // Every thread owns the same line l = tid % NL for all of its items
        // (NT is a multiple of NL), so my is per-thread and mz per item.
        static_assert(NT % NL == 0, "tile threads must cover whole rows");
        const int my = signed_mode(ky0 + l, a.n);
        const T fx = T(kx), fy = T(my);
        cx<T> r0[E], r1[E], r2[E];
#pragma unroll
        for (int i = 0; i < E; ++i) r0[i] = r1[i] = r2[i] = czero<T>();
        // When two tile buffers fit in LDS, the plain product's transform is
        // parked in the second while the shifted product is transformed in the
        // first; otherwise it is held in registers.
        constexpr bool kPark = (2 * TL::LDS + N) * sizeof(cx<T>) <= 160 * 1024;
        __shared__ cx<T> bufp[kPark ? TL::LDS : 1];
        cx<T> Preg[kPark ? 1 : E];
        // products in the reference's accumulation order (00,01,02,11,12,22)
#pragma unroll 1
        for (int p = 0; p < 6; ++p) {
            const int pa = p < 3 ? 0 : (p < 5 ? 1 : 2);
            const int pb = p < 3 ? p : (p < 5 ? p - 2 : 2);
            if constexpr (kPark) {
#pragma unroll
                for (int i = 0; i < E; ++i)
                    if (SPH_IN(i)) bufp[SPH_LI(i)] = SPH_OK(i) ? a.in[2 * p][SPH_OFF(i)] : czero<T>();
                __syncthreads();
                fft_tile<T, N, NL, NT, false>(bufp, twl, tid);
            } else {
                transform(a.in[2 * p]);
#pragma unroll
                for (int i = 0; i < E; ++i) Preg[i] = buf[SPH_LI(i)];
                __syncthreads();
            }
            transform(a.in[2 * p + 1]);
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int mz = signed_mode(e0 + i * STEP, a.n);
                const int k2 = kx * kx + my * my + mz * mz;
                const cx<T> Q = buf[SPH_LI(i)];
                cx<T> Pv;
                if constexpr (kPark)
                    Pv = bufp[SPH_LI(i)];
                else
                    Pv = Preg[i];
                const cx<T> ph = SPH_OK(i) ? a.phase[kx + my + mz + a.n] : czero<T>();
                cx<T> v = (Pv + cmul(Q, cx<T>{ph.x, -ph.y})) * T(0.5);
                if (!keep_mode(k2, a.n)) v = czero<T>();
                const T fz = T(mz);
                const T kb = pb == 0 ? fx : (pb == 1 ? fy : fz);
                const T ka = pa == 0 ? fx : (pa == 1 ? fy : fz);
                const cx<T> da = {kb * v.y, -kb * v.x};  // -i k_b v into r_a
                const cx<T> db = {ka * v.y, -ka * v.x};  // -i k_a v into r_b
                if (pa == 0)
                    r0[i] = r0[i] + da;
                else if (pa == 1)
                    r1[i] = r1[i] + da;
                else
                    r2[i] = r2[i] + da;
                if (pa != pb) {
                    if (pb == 1)
                        r1[i] = r1[i] + db;
                    else
                        r2[i] = r2[i] + db;
                }
            }
            __syncthreads();
        }
        const T alpha = a.coef[2 * a.stage];
        const T beta = a.coef[2 * a.stage + 1];
        const bool first = a.stage == 0;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            if (!SPH_OK(i)) continue;
            const int mz = signed_mode(e0 + i * STEP, a.n);
            const int k2 = kx * kx + my * my + mz * mz;
            const bool keep = keep_mode(k2, a.n);
            cx<T> u[3] = {a.u[0][SPH_OFF(i)], a.u[1][SPH_OFF(i)], a.u[2][SPH_OFF(i)]};
            cx<T> r[3] = {r0[i], r1[i], r2[i]};
            const T vk2 = a.nu * T(k2);
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] = r[c] - u[c] * vk2;
            project_or_zero(r[0], r[1], r[2], kx, my, mz, k2, keep);
            cx<T> g[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                g[c] = first ? r[c] : a.g[c][SPH_OFF(i)] * alpha + r[c];
                u[c] = u[c] + g[c] * beta;
            }
            project_or_zero(u[0], u[1], u[2], kx, my, mz, k2, keep);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a.g[c][SPH_OFF(i)] = g[c];
                a.u[c][SPH_OFF(i)] = u[c];
            }
        }
    }
}

#undef SPH_IN
#undef SPH_OK
#undef SPH_OFF
#undef SPH_LI

// ---------------------------------------------------------------------------
// reductions (FP64 partials)
// ---------------------------------------------------------------------------
constexpr int kReduceThreads = 256;
constexpr int kReduceBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// partial[3*b + 0] = sum 0.5|u|^2, [1] = max |u|^2, [2] = sum 0.5|w|^2
template <typename T>
__global__ void __launch_bounds__(kReduceThreads) k_reduce_phys(const T* u0, const T* u1, const T* u2, const T* w0,
                                                               const T* w1, const T* w2, long long count,
                                                               double* partial) {
    double se = 0.0, mv = 0.0, sw = 0.0;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const double a = u0[i], b = u1[i], c = u2[i];
        const double q = a * a + b * b + c * c;
        se += 0.5 * q;
        mv = fmax(mv, q);
        if (w0) {
            const double d = w0[i], e = w1[i], f = w2[i];
            sw += 0.5 * (d * d + e * e + f * f);
        }
    }
    __shared__ double red[3][kReduceThreads / 64];
    se = wave_sum(se);
    mv = wave_max(mv);
    sw = wave_sum(sw);
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    if (lane == 0) {
        red[0][wv] = se;
        red[1][wv] = mv;
        red[2][wv] = sw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int w = 0; w < kReduceThreads / 64; ++w) {
            a += red[0][w];
            b = fmax(b, red[1][w]);
            c += red[2][w];
        }
        partial[3 * blockIdx.x + 0] = a;
        partial[3 * blockIdx.x + 1] = b;
        partial[3 * blockIdx.x + 2] = c;
    }
}

// partial[b] = max |k . u_hat| over the half spectrum (unnormalised units,
// impl.hpp:1167-1170)
template <typename T>
__global__ void __launch_bounds__(kReduceThreads) k_reduce_div(const cx<T>* u0, const cx<T>* u1, const cx<T>* u2,
                                                              int n, double* partial) {
    const long long count = static_cast<long long>(n / 2 + 1) * n * n;
    double m = 0.0;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int ky = static_cast<int>(i % n);
        const int kz = static_cast<int>((i / n) % n);
        const int kx = static_cast<int>(i / (static_cast<long long>(n) * n));
        const T fx = T(kx), fy = T(signed_mode(ky, n)), fz = T(signed_mode(kz, n));
        const cx<T> d = (u0[i] * fx + u1[i] * fy) + u2[i] * fz;
        m = fmax(m, static_cast<double>(sqrt(d.x * d.x + d.y * d.y)));
    }
    __shared__ double red[kReduceThreads / 64];
    m = wave_max(m);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < kReduceThreads / 64; ++w) b = fmax(b, red[w]);
        partial[blockIdx.x] = b;
    }
}

// out[j] = reduce over blocks of partial[stride*b + j]; op bit j: 1 = max, 0 = sum
static __global__ void __launch_bounds__(kReduceThreads) k_reduce_final(const double* partial, int nblocks, int stride,
                                                                int max_mask, double* out) {
    __shared__ double red[kReduceThreads / 64];
    for (int j = 0; j < stride; ++j) {
        const bool is_max = (max_mask >> j) & 1;
        double v = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
            const double x = partial[stride * b + j];
            v = is_max ? fmax(v, x) : v + x;
        }
        v = is_max ? wave_max(v) : wave_sum(v);
        if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double acc = 0.0;
            for (int w = 0; w < kReduceThreads / 64; ++w) acc = is_max ? fmax(acc, red[w]) : acc + red[w];
            out[j] = acc;
        }
        __syncthreads();
    }
}

// Taylor-Green initial condition at cell centres, evaluated in double and
// rounded once to T (impl.hpp:805-812 uses Real arithmetic; in double this is
// the same expression).
template <typename T>
__global__ void __launch_bounds__(256) k_taylor_green(T* p0, T* p1, T* p2, int n) {
    const long long count = static_cast<long long>(n) * n * n;
    const double rpi = 3.141592653589793238462643383279502884;
    for (long long id = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; id < count;
         id += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int ix = static_cast<int>(id % n);
        const int iy = static_cast<int>((id / n) % n);
        const int iz = static_cast<int>(id / (static_cast<long long>(n) * n));
        const double x = (static_cast<double>(ix) + 0.5) * 2.0 * rpi / static_cast<double>(n);
        const double y = (static_cast<double>(iy) + 0.5) * 2.0 * rpi / static_cast<double>(n);
        const double z = (static_cast<double>(iz) + 0.5) * 2.0 * rpi / static_cast<double>(n);
        p0[id] = static_cast<T>(sin(x) * cos(y) * cos(z));
        p1[id] = static_cast<T>(-cos(x) * sin(y) * cos(z));
        p2[id] = T(0);
    }
}

}  // namespace sph

// Kernels of the standalone 3D real<->complex transforms (the FFT backend
// handle and the solver's initial condition), the projection and the
// reductions shared with the solver's pass kernels (pair_passes.hpp).
//
// Storage (per field):
//   physical  real  [z][y][x]            offset (z*ny + y)*nx + x
//   spectral  cx    [kx][kz][ky]         offset (kx*nz + kz)*ny + ky,
//             kx in [0, nx/2] (half spectrum), ky and kz in [0, n).
//   mixed     the same spectral layout with some axes still in physical space.
//
// Passes (inverse = spectral -> physical, forward = physical -> spectral):
//   inv_z  : z lines of a tile of NL adjacent ky at fixed kx (128-byte row
//            segments), any number of fields.
//   lines  : y lines, contiguous, in place, any number of fields.
//   inv_x  : x lines of NL adjacent y rows at fixed z; two half spectra A, B
//            are packed as Z = A + iB with Hermitian extension (taking Re of
//            the kx = 0 and kx = n/2 planes, which is what Re() of the
//            reference's C2C backward keeps), inverse-transformed, and split
//            into the real rows a = Re z, b = Im z, times 1/N^3
//            (impl.hpp:555-561).
//   fwd_x  : two real rows per complex transform, packed as p + i q,
//            transformed, and separated into the two half spectra:
//            P = (Z_k + conj Z_{n-k})/2, Q = (Z_k - conj Z_{n-k})/(2i).
//   fwd_z  : z lines; PROJECT mode applies project_truncate
//            (impl.hpp:819-845) to the three transformed components and may
//            write them component-interleaved (the solver's state).
#pragma once

#include "fft_device.hpp"

namespace sph {

enum FwdZMode { FZ_PLAIN = 0, FZ_PROJECT = 1 };

__device__ __forceinline__ int signed_mode(int i, int n) { return i <= n / 2 ? i : i - n; }

// keep (mx^2+my^2+mz^2)/(n/2)^2 <= 1, evaluated exactly in integers
// (impl.hpp:131-134 evaluates it in double; both agree for every integer input).
__device__ __forceinline__ bool keep_mode(int k2, int n) {
    return 4LL * k2 <= static_cast<long long>(n) * n;
}

// Raw buffer access: a work-group-uniform resource (scalar registers) plus a
// 32-bit byte offset per lane, so one offset register per slot serves every
// array of the same layout. Accesses beyond `bytes` read 0 / are dropped.
// Resource word 3 for gfx9-family raw buffers: 0x00020000.
// Cache policy of global loads and stores (the aux operand of the buffer
// and global instructions): bit 1 is gfx950's non-temporal hint, NT. The y
// passes stream whole contiguous tiles that no other work-group touches and
// store them with NT (512^3 FP32: inverse y 1.21 -> 1.14 ms, forward y 2.38 ->
// 2.26 ms); the z passes keep the default (their 24/48-byte state rows share
// 128-byte lines with the neighbouring tile, which the L2 serves: with NT the
// forward-z pass took 3.8 instead of 3.4 ms), and so do accesses of partial
// 128-byte lines (the transposed layout's rows).
constexpr int kPolicyDefault = 0;
constexpr int kPolicyStream = 2;

template <typename T, int POL = kPolicyDefault>
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const cx<T>* base, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<cx<T>*>(base), 0, static_cast<int>(bytes), 0x00020000)) {}
    __device__ __forceinline__ cx<T> ld(uint32_t off) const {
        if constexpr (sizeof(T) == 4) {
            return __builtin_bit_cast(cx<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, POL));
        } else {
            return __builtin_bit_cast(cx<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, POL));
        }
    }
    __device__ __forceinline__ void st(uint32_t off, cx<T> v) const {
        if constexpr (sizeof(T) == 4) {
            using V = int __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, off, 0, POL);
        } else {
            using V = int __attribute__((ext_vector_type(4)));
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, v), r, off, 0, POL);
        }
    }
};

// Plain-pointer accesses with cache policy POL.
template <int POL, typename T>
__device__ __forceinline__ cx<T> gld(const cx<T>* p) {
    if constexpr (POL == kPolicyStream) {
        using V = T __attribute__((ext_vector_type(2)));
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
        return {v.x, v.y};
    } else {
        return *p;
    }
}
template <int POL, typename T>
__device__ __forceinline__ void gst(cx<T>* p, cx<T> v) {
    if constexpr (POL == kPolicyStream) {
        using V = T __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(V{v.x, v.y}, reinterpret_cast<V*>(p));
    } else {
        *p = v;
    }
}

// Work-groups are dealt to the 8 XCDs round-robin by linear id, and each XCD
// has its own L2. Renumber them so that consecutive logical tiles (x fastest)
// run on the same XCD at about the same time: tiles whose rows share a 128-byte
// line (the 24/48-byte-per-mode state rows of the z passes) then meet in one
// L2 instead of fetching the line twice. Identity on the last total % 8 ids.
struct TileId {
    unsigned x, y, z;
};
__device__ __forceinline__ TileId xcd_tile_id() {
    constexpr unsigned kXcds = 8;
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned total = gx * gy * gridDim.z;
    const unsigned p = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned full = total / kXcds * kXcds;
    const unsigned l = p < full ? (p % kXcds) * (full / kXcds) + p / kXcds : p;
    return {l % gx, (l / gx) % gy, l / (gx * gy)};
}

// Pointer q of an argument array, selected with a uniform index: a chain of
// scalar selects, so the array is never copied to private memory (dynamic
// indexing of a by-value kernel argument would be).
template <typename P, int K>
__device__ __forceinline__ P pick(P const (&arr)[K], int q) {
    P r = arr[0];
#pragma unroll
    for (int k = 1; k < K; ++k) r = q == k ? arr[k] : r;
    return r;
}

template <typename T>
struct InvZArgs {
    const cx<T>* in[3];
    cx<T>* out[3];
    int nfields;      // in[f] -> out[f]
    int ny;           // number of ky values (tile dimension)
    const cx<T>* tw;  // twiddles of the z length
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
    const T* in[3];
    cx<T>* out[3];
    int nin;  // 1..3 inputs, pairs (in0,in1) (in2,0) -> out[0..nin)
    int ny;
    int nz;
    const cx<T>* tw;
};

template <typename T>
struct FwdZArgs {
    const cx<T>* in[3];
    cx<T>* out[3];  // PLAIN / PROJECT outputs
    cx<T>* u3;      // PROJECT: if set, the output is this component-interleaved state
    int nfields;    // PLAIN
    int ny;
    int n;          // cube size for the modes (PROJECT)
    const cx<T>* tw;
    int kb;         // PROJECT into u3: ky block of the state (0 or ny: plain layout)
};

// ---------------------------------------------------------------------------
// inverse z pass (register-direct first/last stages, see zfft)
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__((ZTile<T, N>::NT)) k_inv_z(InvZArgs<T> a) {
    using Z = ZTile<T, N>;
    constexpr int NL = Z::NL, NT = Z::NT, T0 = Z::T0, R0 = Z::R0, TLS = Z::TLAST, RL = Z::RL;
    __shared__ cx<T> lds[Z::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const int ky = blockIdx.x * NL + tid % NL;
    const int kx = blockIdx.y;
    const bool ok = ky < a.ny;
    // loads of lines beyond ny read line ny-1 (in bounds) and are never stored
    const uint32_t lane = static_cast<uint32_t>((tid / NL) * a.ny + (ok ? ky : a.ny - 1)) * sizeof(cx<T>);
    const uint32_t rowb = static_cast<uint32_t>(a.ny) * sizeof(cx<T>);
    const size_t plane = static_cast<size_t>(kx) * N * a.ny;
    const uint32_t plane_bytes = N * rowb;
    load_twiddles<T, N, NT>(twl, a.tw, tid);
    auto load = [&](const cx<T>* src, cx<T> (&v)[T0][R0]) {
        const Buf<T> b(src + plane, plane_bytes);
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r) v[i][r] = b.ld(lane + (i * Z::STEP + r * (N / R0)) * rowb);
    };
    auto store = [&](cx<T>* dst, const cx<T> (&w)[TLS][RL]) {
        if (!ok) return;
        const Buf<T> b(dst + plane, plane_bytes);
#pragma unroll
        for (int i = 0; i < TLS; ++i)
#pragma unroll
            for (int r = 0; r < RL; ++r) b.st(lane + (i * Z::STEP + r * (N / RL)) * rowb, w[i][r]);
    };
    cx<T> v[1][T0][R0];
    cx<T> w[1][TLS][RL];
    for (int f = 0; f < a.nfields; ++f) {
        load(pick(a.in, f), v[0]);
        zfft<T, N, 1, true>(v, w, lds, twl, tid);
        store(pick(a.out, f), w[0]);
    }
}

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
        const cx<T>* A = pick(a.a, p);
        const cx<T>* B = pick(a.b, p);
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
        T* OA = pick(a.oa, p);
        T* OB = pick(a.ob, p);
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

template <typename T, int N>
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
#pragma unroll 1
    for (int p = 0; 2 * p < a.nin; ++p) {
        const T* fa = pick(a.in, 2 * p);
        const T* fb = 2 * p + 1 < a.nin ? pick(a.in, 2 * p + 1) : nullptr;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            if (in[i]) buf[li[i]] = {ok[i] ? fa[off[i]] : T(0), (ok[i] && fb) ? fb[off[i]] : T(0)};
        }
        __syncthreads();
        fft_tile<T, N, NL, NT, false>(buf, twl, tid);
        fwd_x_separate_store<T, N>(buf, pick(a.out, 2 * p), fb ? pick(a.out, 2 * p + 1) : nullptr, y0, z, a.ny,
                                   a.nz, tid);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// forward z pass (+ projection)
// ---------------------------------------------------------------------------
// Leray projection of f (impl.hpp:829-843): f -= k (k.f) * (1/k2) for kept
// modes with k2 > 0, zero otherwise; evaluated without branches.
template <typename T>
__device__ __forceinline__ void project_or_zero(cx<T>& f0, cx<T>& f1, cx<T>& f2, int ix, int iy, int iz, int k2,
                                                bool keep) {
    const bool live = k2 > 0 && keep;
    const T inv = T(1) / T(live ? k2 : 1);
    const T fx = T(ix), fy = T(iy), fz = T(iz);
    const cx<T> dot = (f0 * fx + f1 * fy) + f2 * fz;
    const cx<T> fac = dot * inv;
    const cx<T> g0 = f0 - fac * fx, g1 = f1 - fac * fy, g2 = f2 - fac * fz;
    f0 = live ? g0 : czero<T>();
    f1 = live ? g1 : czero<T>();
    f2 = live ? g2 : czero<T>();
}

template <typename T, int N, int MODE>
__global__ void __launch_bounds__((ZTile<T, N>::NT)) k_fwd_z(FwdZArgs<T> a) {
    using Z = ZTile<T, N>;
    constexpr int NL = Z::NL, NT = Z::NT, T0 = Z::T0, R0 = Z::R0, TLS = Z::TLAST, RL = Z::RL;
    __shared__ cx<T> lds[Z::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const int ky = blockIdx.x * NL + tid % NL;
    const int kx = blockIdx.y;
    const bool ok = ky < a.ny;
    const uint32_t lane = static_cast<uint32_t>((tid / NL) * a.ny + (ok ? ky : a.ny - 1)) * sizeof(cx<T>);
    const uint32_t rowb = static_cast<uint32_t>(a.ny) * sizeof(cx<T>);
    const size_t plane = static_cast<size_t>(kx) * N * a.ny;
    const uint32_t plane_bytes = N * rowb;
    load_twiddles<T, N, NT>(twl, a.tw, tid);
    auto load = [&](const cx<T>* src, cx<T> (&v)[T0][R0]) {
        const Buf<T> b(src + plane, plane_bytes);
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r) v[i][r] = b.ld(lane + (i * Z::STEP + r * (N / R0)) * rowb);
    };
    auto last_off = [&](int i, int r) -> uint32_t { return lane + (i * Z::STEP + r * (N / RL)) * rowb; };
    cx<T> v[1][T0][R0];
    cx<T> w[1][TLS][RL];
    if constexpr (MODE == FZ_PLAIN) {
        for (int f = 0; f < a.nfields; ++f) {
            load(pick(a.in, f), v[0]);
            zfft<T, N, 1, false>(v, w, lds, twl, tid);
            if (ok) {
                const Buf<T> o(pick(a.out, f) + plane, plane_bytes);
#pragma unroll
                for (int i = 0; i < TLS; ++i)
#pragma unroll
                    for (int r = 0; r < RL; ++r) o.st(last_off(i, r), w[0][i][r]);
            }
        }
    } else if constexpr (MODE == FZ_PROJECT) {
        cx<T> u[3][TLS][RL];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            load(a.in[c], v[0]);
            zfft<T, N, 1, false>(v, w, lds, twl, tid);
#pragma unroll
            for (int i = 0; i < TLS; ++i)
#pragma unroll
                for (int r = 0; r < RL; ++r) u[c][i][r] = w[0][i][r];
        }
        if (ok) {
            const int my = signed_mode(ky, a.n);
#pragma unroll
            for (int i = 0; i < TLS; ++i)
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int e = Z::e_last(tid, i, r);
                    const int mz = signed_mode(e, a.n);
                    const int k2 = kx * kx + my * my + mz * mz;
                    project_or_zero(u[0][i][r], u[1][i][r], u[2][i][r], kx, my, mz, k2, keep_mode(k2, a.n));
                    if (a.u3) {
                        const Buf<T> o(a.u3 + plane * 3, plane_bytes * 3);
                        const uint32_t kb = a.kb > 0 ? a.kb : a.ny;
                        const uint32_t off3 =
                            ((ky / kb) * N * kb + Z::e_last(tid, i, r) * kb + ky % kb) * 3 * sizeof(cx<T>);
#pragma unroll
                        for (int c = 0; c < 3; ++c) o.st(off3 + c * sizeof(cx<T>), u[c][i][r]);
                    } else {
#pragma unroll
                        for (int c = 0; c < 3; ++c) Buf<T>(a.out[c] + plane, plane_bytes).st(last_off(i, r), u[c][i][r]);
                    }
                }
        }
    }
}

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
// (launched from the solver's translation unit only)
[[maybe_unused]] static __global__ void __launch_bounds__(kReduceThreads)
    k_reduce_final(const double* partial, int nblocks, int stride, int max_mask, double* out) {
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

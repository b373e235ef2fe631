// Pass kernels of the solver's RK3 stage on pair-interleaved work buffers.
//
// Everything the nonlinear term transforms comes in pairs: (u_hat, u_hat*phi)
// on the inverse side (unshifted / half-cell-shifted velocity,
// src/ava_taylor_green_impl.hpp:851-879), (u_a u_b, s_a s_b) on the forward
// side (impl.hpp:903-911), (u_hat, omega_hat) for the diagnostics
// (impl.hpp:1129-1160). The work buffers store the two members of a pair side
// by side per mode:
//   pair field  [kx][kz or z][ky][2]        offset ((kx*n + kz)*ny + ky)*2 + f
// and the state is component-interleaved:
//   u_hat, g    [kx][kz][ky][3]             offset ((kx*n + kz)*ny + ky)*3 + c.
// The spectral side (the state, and the pair fields between the z passes and
// the y passes) may instead be ky-blocked, kb ky per block (the z tile's
// NL/2), so that a z tile is one contiguous block of each field:
//   [kx][ky/kb][kz][ky%kb][2 or 3]  mode offset kx*n*ny + (ky/kb)*n*kb + kz*kb + ky%kb
// (kb = ny is the plain layout above). The z-pass arguments carry kb.
// In the z passes a tile is NL/2 adjacent ky times both pair members, i.e. one
// 128-byte row segment per element; lane l transforms member f = l & 1 of line
// ky = l >> 1, and the two lanes of a line split the modes after the
// transform (a lane swap), so each lane keeps RHS accumulators for half of
// its modes only.
#pragma once

#include "passes.hpp"

namespace sph {

enum PairZMode { PZ_RHS = 0, PZ_DIAG = 1, PZ_VEL = 2 };

// Batched launches: blockIdx.z (blockIdx.y for the y lines) is a member of an
// ensemble of equal-size solves; every pointer below addresses member 0 and
// member m adds m times the stride of its family (ms_u: state, ms_w: work
// pairs, ms_p: physical fields), all in complex (state, work) or real
// (physical) elements. A single solve uses zero strides.
template <typename T>
struct PairInvZArgs {
    const cx<T>* u3;  // state, component-interleaved
    cx<T>* out[3];    // pair fields
    int ny;
    int n;
    const cx<T>* phase;
    const cx<T>* tw;
    long long ms_u;
    long long ms_w;
    int kb;  // ky block of the state and of the pair fields (ny: plain layout)
};

template <typename T>
struct PairInvXArgs {
    const cx<T>* ab[3];  // pair fields (A, B) in mixed space, rows krow / zrow apart (below)
    T* oa[3];
    T* ob[3];  // may be null (B discarded)
    int npairs;
    int ny;
    int nz;
    // pairs between kx rows and between z planes of the mixed-space fields:
    // [kx][z][y] (krow = nz*ny, zrow = ny) or [z][kx][y] (krow = ny, zrow =
    // (n/2+1)*ny)
    long long krow, zrow;
    T scale;
    const cx<T>* tw;
    long long ms_w;
    long long ms_p;
};

// coef: per member 8 values in device memory: alpha_s, beta_s*dt for the
// three stages, then nu.
constexpr int kCoefStride = 8;

template <typename T>
struct PairFwdZArgs {
    const cx<T>* in[6];  // pair fields after the x and y transforms
    cx<T>* u3;
    cx<T>* g3;
    int ny;
    int n;
    const T* coef;
    int stage;
    const cx<T>* phase;
    const cx<T>* tw;
    long long ms_u;
    long long ms_w;
    int kb;  // ky block of the state and of the pair fields (ny: plain layout)
};

// Mode offset of (kx, kz, ky) in one member's state or pair field of an n^3
// grid (n*n modes per kx plane) with ky blocks of kb (kb = n: the plain
// layout, (kx*n + kz)*n + ky), and its inverse.
__host__ __device__ __forceinline__ long long spectral_mode_index(int kx, int kz, int ky, int n, int kb) {
    return static_cast<long long>(kx) * n * n + static_cast<long long>(ky / kb) * n * kb +
           static_cast<long long>(kz) * kb + ky % kb;
}
__host__ __device__ __forceinline__ void spectral_mode_of(long long i, int n, int kb, int& kx, int& kz, int& ky) {
    kx = static_cast<int>(i / (static_cast<long long>(n) * n));
    const int rem = static_cast<int>(i % (static_cast<long long>(n) * n));
    kz = (rem % (n * kb)) / kb;
    ky = (rem / (n * kb)) * kb + rem % kb;
}

// Mode offset within a kx plane of element (kz) 0 of line ky, and the
// offset step between consecutive kz, for ky blocks of kb (kb = ny: the plain
// layout, kz * ny + ky); spectral_mode_index of (0, kz, ky) = base + kz*step.
struct ZRow {
    uint32_t base, step;
    __host__ __device__ __forceinline__ ZRow(uint32_t ky, uint32_t kb, uint32_t n)
        : base((ky / kb) * n * kb + ky % kb), step(kb) {}
};

// both members of a pair at one mode: one 16-byte (FP32) / 32-byte access
template <typename T>
struct alignas(4 * sizeof(T)) cx2 {
    cx<T> a;
    cx<T> b;
};

// cx2 through a raw buffer resource (32-bit offset; reads past the resource's
// extent return zeros)
template <typename T, int POL = kPolicyDefault>
__device__ __forceinline__ cx2<T> ld_cx2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    using V4 = int __attribute__((ext_vector_type(4)));
    cx2<T> q;
    if constexpr (sizeof(T) == 4) {
        q = __builtin_bit_cast(cx2<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, POL));
    } else {
        const V4 lo = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, POL);
        const V4 hi = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, POL);
        q.a = __builtin_bit_cast(cx<T>, lo);
        q.b = __builtin_bit_cast(cx<T>, hi);
    }
    return q;
}

// The x pass prefetches the rows of its second and third velocity pairs into
// LDS while the previous pair is transformed (FP32, line layout, lengths with
// padded ping-pong tiles: 256 and 512; 512^3 x pass 2.35 -> 2.27 ms).
// SPH_X_DMA=0 builds without it.
#ifndef SPH_X_DMA
#define SPH_X_DMA 1
#endif
constexpr bool kXPrefetch = SPH_X_DMA != 0;

// store of a cx2 with cache policy POL
template <int POL, typename T>
__device__ __forceinline__ void gst2(cx2<T>* p, cx2<T> v) {
    if constexpr (POL == kPolicyStream) {
        using V = T __attribute__((ext_vector_type(4)));
        __builtin_nontemporal_store(V{v.a.x, v.a.y, v.b.x, v.b.y}, reinterpret_cast<V*>(p));
    } else {
        *p = v;
    }
}

template <typename T>
__device__ __forceinline__ cx<T> shfl_xor1(cx<T> v) {
    return {__shfl_xor(v.x, 1), __shfl_xor(v.y, 1)};
}

// Cache policy of the z passes' pair-field rows (128-byte segments of one
// tile): FP32 stores non-temporal (512^3: forward z + update + next inverse
// z 3.41 -> 3.33 ms).
template <typename T>
constexpr int z_pair_store_policy() {
    return sizeof(T) == 4 ? kPolicyStream : kPolicyDefault;
}
// and the forward-z pass's loads of them: NT at N = 1024 in FP32 (27.9 ->
// 27.2 ms per launch in the ky-blocked layout; neutral at 512); its state
// accesses keep the default (NT state: 512^3 3.1 -> 4.1 ms)
template <typename T, int N>
constexpr int z_pair_load_policy() {
    return sizeof(T) == 4 && N >= 1024 ? kPolicyStream : kPolicyDefault;
}

// Ping-pong z tiles (zfft's PP) in the solver's z passes: large lengths with
// middle stages, when two field tiles and the twiddles fit in LDS (the
// kernels already run one 1024-thread work-group per CU there).
template <typename T, int N, int LINES = 0>
constexpr bool z_pingpong() {
    return N >= 256 && z_mid_stages<T, N, LINES>() >= 1 &&
           (2 * ZTile<T, N, LINES>::LDS + N) * static_cast<int>(sizeof(cx<T>)) <= kLdsBytes;
}
template <typename T, int N, int LINES = 0>
constexpr int z_lds() {
    return (z_pingpong<T, N, LINES>() ? 2 : 1) * ZTile<T, N, LINES>::LDS;
}

// State of a z tile kept in LDS between the update and the next stage's
// inverse z (k_pfwd_z_rk_inv): component c, element (kz) e, ky ky_local of
// the tile's NL/2; it overlays the transform tiles, so it needs
// 3 * N * NL/2 values of the kernel's LDS.
template <int N, int NL>
__host__ __device__ __forceinline__ int z_stash_at(int c, int e, int kyl) {
    return (c * N + e) * (NL / 2) + kyl;
}
template <typename T, int N, int LINES = 0>
constexpr bool z_stash_fits() {
    return 3 * N * (ZTile<T, N, LINES>::NL / 2) <= z_lds<T, N, LINES>();
}

// ---------------------------------------------------------------------------
// inverse z: state -> pair fields. One tile: the body below runs in
// k_pinv_z and, for the next RK stage, at the end of k_pfwd_z_rk_inv; the
// twiddles are in twl already, lds holds z_lds<T, N>() values.
// ---------------------------------------------------------------------------
// su (optional): the tile's state in LDS, written by pfwd_z_rk_tile (layout
// z_stash_at); the state is then not read from global memory.
template <typename T, int N, int MODE, int LINES = 0>
__device__ __forceinline__ void pinv_z_tile(const PairInvZArgs<T>& a, TileId tile, cx<T>* lds, const cx<T>* twl,
                                            int tid, const cx<T>* su = nullptr) {
    using Z = ZTile<T, N, LINES>;
    constexpr int NL = Z::NL, T0 = Z::T0, R0 = Z::R0, TLS = Z::TLAST, RL = Z::RL;
    static_assert(NL % 2 == 0, "pair tiles need an even line count");
    const int l = tid % NL, f = l & 1;
    const int ky = tile.x * (NL / 2) + (l >> 1);
    const int kx = tile.y;
    const bool ok = ky < a.ny;
    const uint32_t kyc = ok ? ky : a.ny - 1;
    constexpr uint32_t CB = sizeof(cx<T>);
    const uint32_t ny = a.ny;
    // member offsets kept in locals: writing into the argument struct would
    // copy it to private memory
    const long long zw = tile.z * a.ms_w;
    const size_t plane = static_cast<size_t>(kx) * N * ny;
    const Buf<T> U(a.u3 + tile.z * a.ms_u + plane * 3, N * ny * 3 * CB);
    const ZRow zr(kyc, a.kb > 0 ? a.kb : ny, N);
    const uint32_t u_lane = (zr.base + (tid / NL) * zr.step) * 3 * CB, u_row = zr.step * 3 * CB;
    const uint32_t p_lane = ((zr.base + (tid / NL) * zr.step) * 2 + f) * CB, p_row = zr.step * 2 * CB;
    const uint32_t p_bytes = N * ny * 2 * CB;
    cx<T> v[1][T0][R0], w[1][TLS][RL];
    auto uoff = [&](int i, int r, int c) -> uint32_t {
        return u_lane + (i * Z::STEP + r * (N / R0)) * u_row + c * CB;
    };
    auto store = [&](cx<T>* dst) {
        if (!ok) return;
        const Buf<T, z_pair_store_policy<T>()> b(dst + zw + plane * 2, p_bytes);
#pragma unroll
        for (int i = 0; i < TLS; ++i)
#pragma unroll
            for (int r = 0; r < RL; ++r) b.st(p_lane + (i * Z::STEP + r * (N / RL)) * p_row, w[0][i][r]);
    };
    const int my = signed_mode(kyc, a.n);
    // all three components of the lane's modes, read once (one 24/48-byte row
    // per mode of the component-interleaved state)
    cx<T> us[3][T0][R0];
    if (su) {
        const int kyl = static_cast<int>(kyc) - static_cast<int>(tile.x) * (NL / 2);
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) us[c][i][r] = su[z_stash_at<N, NL>(c, Z::e_first(tid, i, r), kyl)];
        // every wave has its values before the first transform writes over
        // the stash
        __syncthreads();
    } else {
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) us[c][i][r] = U.ld(uoff(i, r, c));
    }
    if constexpr (MODE == PZ_RHS) {
        // member 0: u_hat_c, member 1: u_hat_c * phi (impl.hpp:866-873); the
        // shift is applied to all three components before the first
        // transform, so the phases need no registers during the transforms
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const int mz = signed_mode(Z::e_first(tid, i, r), a.n);
                const cx<T> ph = f ? a.phase[kx + my + mz + a.n] : cx<T>{T(1), T(0)};
#pragma unroll
                for (int c = 0; c < 3; ++c) us[c][i][r] = f ? cmul(us[c][i][r], ph) : us[c][i][r];
            }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int i = 0; i < T0; ++i)
#pragma unroll
                for (int r = 0; r < R0; ++r) v[0][i][r] = us[c][i][r];
            zfft<T, N, 1, true, LINES, z_pingpong<T, N, LINES>()>(v, w, lds, twl, tid);
            store(a.out[c]);
        }
    } else if constexpr (MODE == PZ_DIAG) {
        // member 0: u_hat_c, member 1: omega_hat_c = (i k x u_hat)_c (impl.hpp:1147-1157)
        const T ix = T(kx), iy = T(my);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int i = 0; i < T0; ++i)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const cx<T> u0 = us[0][i][r], u1 = us[1][i][r], u2 = us[2][i][r];
                    const T iz = T(signed_mode(Z::e_first(tid, i, r), a.n));
                    cx<T> om;
                    if (c == 0)
                        om = {-(iy * u2.y - iz * u1.y), iy * u2.x - iz * u1.x};
                    else if (c == 1)
                        om = {-(iz * u0.y - ix * u2.y), iz * u0.x - ix * u2.x};
                    else
                        om = {-(ix * u1.y - iy * u0.y), ix * u1.x - iy * u0.x};
                    const cx<T> uc = c == 0 ? u0 : (c == 1 ? u1 : u2);
                    v[0][i][r] = f ? om : uc;
                }
            zfft<T, N, 1, true, LINES, z_pingpong<T, N, LINES>()>(v, w, lds, twl, tid);
            store(a.out[c]);
        }
    } else {  // PZ_VEL: pairs (u0, u1) and (u2, 0)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
#pragma unroll
            for (int i = 0; i < T0; ++i)
#pragma unroll
                for (int r = 0; r < R0; ++r)
                    v[0][i][r] = q == 0 ? (f ? us[1][i][r] : us[0][i][r]) : (f ? czero<T>() : us[2][i][r]);
            zfft<T, N, 1, true, LINES, z_pingpong<T, N, LINES>()>(v, w, lds, twl, tid);
            store(a.out[q]);
        }
    }
}

// (launch bounds: 512-thread tiles ask for 4 waves per SIMD, i.e. two
// work-groups per CU, which caps the registers at 128 like the 1024-thread
// tiles; smaller tiles leave the register budget to the compiler: held to
// 128 the 96^3 forward-z kernel spills 240 bytes per lane)
template <typename T, int N, int MODE, int LINES = 0>
__global__ void __launch_bounds__((ZTile<T, N, LINES>::NT), (ZTile<T, N, LINES>::kMinWaves))
    k_pinv_z(PairInvZArgs<T> a) {
    using Z = ZTile<T, N, LINES>;
    __shared__ cx<T> lds[z_lds<T, N, LINES>()];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    load_twiddles<T, N, Z::NT>(twl, a.tw, tid);
    pinv_z_tile<T, N, MODE, LINES>(a, xcd_tile_id(), lds, twl, tid);
}

// ---------------------------------------------------------------------------
// contiguous y lines of pair fields, in place: a tile is NL/2 pair lines,
// i.e. NL*N consecutive complex values, member f of pair line g is line 2g+f
// ---------------------------------------------------------------------------
// One tile (the twiddles already in twl; buf holds tile_lds values):
// pair lines tile.x*NL/2 .. of member tile.y.
template <typename T, int N>
constexpr bool plines_swizzled(int nl) {
    return sizeof(T) == 4 && nl <= 8;  // narrow FP32 tiles: unpadded rotated rows (tile_at)
}
template <typename T, int N, bool INV, int LINES>
__device__ __forceinline__ void plines_tile(cx<T>* data, long long npairlines, long long member_stride, TileId tile,
                                            cx<T>* buf, const cx<T>* twl, int tid) {
    using TL = Tile<T, N, LINES>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E;
    static_assert(NL % 2 == 0, "pair tiles need an even line count");
    constexpr bool SWZ = plines_swizzled<T, N>(NL);
    const long long g0 = static_cast<long long>(tile.x) * (NL / 2);
    cx<T>* base = data + tile.y * member_stride + g0 * 2 * N;
    const int valid = static_cast<int>(npairlines - g0 < NL / 2 ? npairlines - g0 : NL / 2) * 2 * N;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int m = tid + i * NT;
        const int g = m / (2 * N), rem = m % (2 * N);
        if (m < TL::ELEMS)
            buf[tile_at<NL, SWZ>(rem >> 1, 2 * g + (rem & 1))] = m < valid ? gld<kPolicyStream>(base + m) : czero<T>();
    }
    __syncthreads();
    fft_tile<T, N, NL, NT, INV, 0, 1, SWZ>(buf, twl, tid);
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int m = tid + i * NT;
        const int g = m / (2 * N), rem = m % (2 * N);
        if (m < valid) gst<kPolicyStream>(base + m, buf[tile_at<NL, SWZ>(rem >> 1, 2 * g + (rem & 1))]);
    }
}

template <typename T, int N, bool INV, int LINES = default_lines<T, N>()>
__global__ void __launch_bounds__((Tile<T, N, LINES>::NT))
    k_plines(cx<T>* data, long long npairlines, long long member_stride, const cx<T>* tw) {
    using TL = Tile<T, N, LINES>;
    __shared__ cx<T> buf[tile_lds<TL::NL, plines_swizzled<T, N>(TL::NL)>(N)];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    load_twiddles<T, N, TL::NT>(twl, tw, tid);
    // an XCD's tiles are one contiguous run
    plines_tile<T, N, INV, LINES>(data, npairlines, member_stride, xcd_tile_id(), buf, twl, tid);
}

// ---------------------------------------------------------------------------
// y lines between the line layout and the transposed (kx-fastest) layout of
// the x pass:
//   line layout        [kx][z][y][2]   offset ((kx*n + z)*n + y)*2 + f
//   transposed layout  [z][y][kx][2]   offset ((z*n + y)*H + kx)*2 + f
// (per pair field; fields `field_stride` complex values apart, members
// `member_stride` apart; grid (kx blocks of NL/2, z, field + nfields*member)).
// In the transposed layout an x tile (NL y lines of one z, every kx) is one
// contiguous block instead of H rows n*n pairs apart, which at n = 1024 cost
// the x pass its address translations. A tile here is NL/2 pair lines of
// consecutive kx at one z: on the line-layout side each pair line is 2n
// contiguous values, on the transposed side each y row holds the tile's NL
// values contiguously. INV: line layout (spectral ky) in, transposed
// (physical y) out; forward: transposed in, line layout out. Out of place.
// ---------------------------------------------------------------------------
// the line-layout side of k_plines_t (whole contiguous pair lines) streams
// non-temporally, the transposed side (partial 128-byte lines) keeps the
// default policy (1024^3: inverse y 13.4 -> 12.7 ms)
constexpr int kPtLinePolicy = kPolicyStream;
template <typename T, int N, bool INV, int LINES>
__global__ void __launch_bounds__((Tile<T, N, LINES>::NT))
    k_plines_t(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
               const cx<T>* tw) {
    using TL = Tile<T, N, LINES>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E;
    constexpr int H = N / 2 + 1;
    static_assert(NL % 2 == 0, "pair tiles need an even line count");
    constexpr bool SWZ = sizeof(T) == 4 && NL <= 8;
    __shared__ cx<T> buf[tile_lds<NL, SWZ>(N)];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const TileId tile = xcd_tile_id();
    const int kx0 = tile.x * (NL / 2);
    const int z = tile.y;
    const int field = tile.z % nfields, member = tile.z / nfields;
    const long long fo = field * field_stride + member * member_stride;
    const int lines = (H - kx0 < NL / 2 ? H - kx0 : NL / 2) * 2;  // valid values of a transposed row
    const cx<T>* src = in + fo;
    cx<T>* dst = out + fo;
    // line layout: value rem of pair line kx0 + g; transposed layout: value
    // c of the tile's row y
    auto lin_off = [&](int g, int rem) -> long long {
        return ((static_cast<long long>(kx0 + g) * N + z) * N) * 2 + rem;
    };
    const long long row0 = (static_cast<long long>(z) * N * H + kx0) * 2;
    auto tr_off = [&](int y, int c) -> long long { return row0 + static_cast<long long>(y) * H * 2 + c; };
    load_twiddles<T, N, NT>(twl, tw, tid);
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int m = tid + i * NT;
        if (TL::ELEMS % NT == 0 || m < TL::ELEMS) {
            if constexpr (INV) {
                const int g = m / (2 * N), rem = m % (2 * N);  // pair line g, element rem >> 1, member rem & 1
                buf[tile_at<NL, SWZ>(rem >> 1, 2 * g + (rem & 1))] = 2 * g < lines ? gld<kPtLinePolicy>(src + lin_off(g, rem)) : czero<T>();
            } else {
                const int y = m / NL, c = m % NL;  // row y, value c = 2g + f
                buf[tile_at<NL, SWZ>(y, c)] = c < lines ? gld<kPolicyDefault>(src + tr_off(y, c)) : czero<T>();
            }
        }
    }
    __syncthreads();
    fft_tile<T, N, NL, NT, INV, 0, 1, SWZ>(buf, twl, tid);
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int m = tid + i * NT;
        if (TL::ELEMS % NT == 0 || m < TL::ELEMS) {
            if constexpr (INV) {
                const int y = m / NL, c = m % NL;
                if (c < lines) gst<kPolicyDefault>(dst + tr_off(y, c), buf[tile_at<NL, SWZ>(y, c)]);
            } else {
                const int g = m / (2 * N), rem = m % (2 * N);
                if (2 * g < lines) gst<kPtLinePolicy>(dst + lin_off(g, rem), buf[tile_at<NL, SWZ>(rem >> 1, 2 * g + (rem & 1))]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// y lines between the line layout of the x pass and the ky-blocked spectral
// layout of the z passes (out of place):
//   line layout   [kx][z][y][2] or [z][kx][y][2]: offset kx*xs_kx + z*xs_z + 2y + f
//   ky-blocked    [kx][ky/KB][kz][ky%KB][2]    offset (kx*n*n + (ky/KB)*n*KB + kz*KB + ky%KB)*2 + f
// with KB the default z tile's NL/2, so that a z tile is one contiguous block
// of each field (at n = 1024 the z passes' 64-byte rows 16 KB apart cost them
// their streaming rate). A tile is NL/2 pair lines of consecutive z (kz) of
// one kx: on the line side NL/2 lines of 2N contiguous values, xs_z apart; on
// the blocked side, per ky block, the tile's NL/2 rows of KB modes, NL*KB
// contiguous values. Thread q
// of a side takes that side's q-th value in address order. INV: blocked
// (spectral ky) in, line layout (physical y) out; forward: the other way.
// Grid (pair lines / (NL/2), 1, field + nfields*member).
// ---------------------------------------------------------------------------
template <typename T, int N>
constexpr int zblock_ky() {
    return ZTile<T, N>::NL / 2;
}
// Blocked side of a k_plines_b tile of NL lines: the q-th value in address
// order is element ky of tile line `line` (pair line line/2, member line&1),
// at this offset (complex values) from the tile's first blocked value (pair
// line 0, ky = 0, member 0).
template <int N, int NL, int KB>
__host__ __device__ __forceinline__ long long plines_b_slot(int q, int& ky, int& line) {
    constexpr int RUN = NL * KB;  // contiguous values per ky block
    const int b = q / RUN, r = q % RUN;
    const int g = r / (2 * KB), kyl = (r >> 1) % KB, f = r & 1;
    ky = b * KB + kyl;
    line = 2 * g + f;
    return static_cast<long long>(b) * N * KB * 2 + r;
}
// One tile: pair lines z0 .. z0 + NL/2 - 1 of kx, the field and member
// offsets already applied to src and dst (twl holds the twiddles, buf
// tile_lds values).
template <typename T, int N, bool INV, int LINES>
__device__ __forceinline__ void plines_b_tile(const cx<T>* src, cx<T>* dst, long long xs_kx, long long xs_z, int kx,
                                              int z0, cx<T>* buf, const cx<T>* twl, int tid) {
    using TL = Tile<T, N, LINES>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, KB = zblock_ky<T, N>();
    static_assert(NL % 2 == 0, "pair tiles need an even line count");
    static_assert(N % (NL / 2) == 0 && N % KB == 0, "tiles must not straddle a kx plane or a ky block");
    constexpr bool SWZ = plines_swizzled<T, N>(NL);
    const long long line0 = kx * xs_kx + z0 * xs_z;  // line side: the tile's first value
    const long long blk0 = (static_cast<long long>(kx) * N * N + static_cast<long long>(z0) * KB) * 2;
    // blocked side: q-th value of the tile in address order -> offset, ky, line
    auto blk = [&](int q, int& ky, int& line) -> long long { return blk0 + plines_b_slot<N, NL, KB>(q, ky, line); };
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int q = tid + i * NT;
        if (TL::ELEMS % NT != 0 && q >= TL::ELEMS) break;
        if constexpr (INV) {
            int ky, line;
            const long long o = blk(q, ky, line);
            buf[tile_at<NL, SWZ>(ky, line)] = gld<kPolicyStream>(src + o);
        } else {
            const int g = q / (2 * N), rem = q % (2 * N);
            buf[tile_at<NL, SWZ>(rem >> 1, 2 * g + (rem & 1))] = gld<kPolicyStream>(src + line0 + g * xs_z + rem);
        }
    }
    __syncthreads();
    fft_tile<T, N, NL, NT, INV, 0, 1, SWZ>(buf, twl, tid);
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int q = tid + i * NT;
        if (TL::ELEMS % NT != 0 && q >= TL::ELEMS) break;
        if constexpr (INV) {
            const int g = q / (2 * N), rem = q % (2 * N);
            gst<kPolicyStream>(dst + line0 + g * xs_z + rem, buf[tile_at<NL, SWZ>(rem >> 1, 2 * g + (rem & 1))]);
        } else {
            int ky, line;
            const long long o = blk(q, ky, line);
            gst<kPolicyStream>(dst + o, buf[tile_at<NL, SWZ>(ky, line)]);
        }
    }
}

template <typename T, int N, bool INV, int LINES>
__global__ void __launch_bounds__((Tile<T, N, LINES>::NT))
    k_plines_b(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
               long long xs_kx, long long xs_z, const cx<T>* tw) {
    using TL = Tile<T, N, LINES>;
    constexpr int G = TL::NL / 2;  // pair lines per tile
    __shared__ cx<T> buf[tile_lds<TL::NL, plines_swizzled<T, N>(TL::NL)>(N)];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const TileId tile = xcd_tile_id();
    // consecutive tiles: consecutive z blocks of one kx, whose rows are
    // neighbours in every ky block (kx fastest instead, where the line side
    // is [z][kx][y]: 1024^3 forward y 17.1 -> 21.6 ms)
    const long long g0 = static_cast<long long>(tile.x) * G;  // pair line kx*N + z of the tile's first line
    const int kx = static_cast<int>(g0 / N), z0 = static_cast<int>(g0 % N);
    const int field = tile.z % nfields, member = tile.z / nfields;
    const long long fo = field * field_stride + member * member_stride;
    load_twiddles<T, N, TL::NT>(twl, tw, tid);
    plines_b_tile<T, N, INV, LINES>(in + fo, out + fo, xs_kx, xs_z, kx, z0, buf, twl, tid);
}

// ---------------------------------------------------------------------------
// x-pass building blocks. A tile is NL x lines (y0 .. y0+NL-1, fixed z) in LDS,
// element x of line l at buf[x*S + l]; thread tid handles element
// idx = tid + i*NT (x = idx % N, l = idx / N) of the physical lines.
// ---------------------------------------------------------------------------

// Spectral x lines of pair field AB -> buf as Z = A + iB over all N modes
// (Hermitian extension of both members; Re rule on kx = 0 and N/2).
template <typename T, int N, int LINES>
__device__ __forceinline__ void load_pair_x_lines(cx<T>* buf, const cx<T>* AB, int tid, int y0, int z, int ny,
                                                  long long krow, long long zrow) {
    using TL = Tile<T, N, LINES>;
    constexpr int NL = TL::NL, NT = TL::NT, S = TL::STRIDE;
    constexpr int LOADS = NL * (N / 2 + 1);
    constexpr int LPT = (LOADS + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
        const int idx = tid + i * NT;
        if (LOADS % NT == 0 || idx < LOADS) {
            const int l = idx % NL, k = idx / NL;
            // lines past the last one read a valid row and discard it
            const bool ok = y0 + l < ny;
            const size_t off = k * krow + z * zrow + (ok ? y0 + l : ny - 1);
            const cx2<T> q = reinterpret_cast<const cx2<T>*>(AB)[off];
            cx<T> va = ok ? q.a : czero<T>(), vb = ok ? q.b : czero<T>();
            if (k == 0 || k == N / 2) {
                va.y = T(0);
                vb.y = T(0);
            }
            buf[k * S + l] = {va.x - vb.y, va.y + vb.x};
            if (k > 0 && k < N / 2) buf[(N - k) * S + l] = {va.x + vb.y, vb.x - va.y};
        }
    }
}

// The six products (u_a u_b, s_a s_b), ab in (00,01,02,11,12,22), of the
// physical values held in registers: forward x transform of each pair and
// separation into the pair field out[p] (P_hat, Q_hat) at the tile's modes.
template <typename T, int N, int LINES, int E>
__device__ __forceinline__ void products_forward_x(const T (&u)[3][E], const T (&s)[3][E], cx<T>* buf,
                                                   const cx<T>* twl, cx<T>* const (&out)[6], long long zw, int tid,
                                                   int y0, int z, int ny, long long krow, long long zrow) {
    using TL = Tile<T, N, LINES>;
    constexpr int NL = TL::NL, NT = TL::NT, S = TL::STRIDE;
    static_assert(E == TL::E, "register slots must match the tile");
    constexpr int ITEMS = NL * (N / 2 + 1);
    constexpr int IPT = (ITEMS + NT - 1) / NT;
    constexpr int PA[6] = {0, 0, 0, 1, 1, 2};
    constexpr int PB[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int p = 0; p < 6; ++p) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int idx = tid + i * NT;
            if (TL::ELEMS % NT == 0 || idx < TL::ELEMS)
                buf[(idx % N) * S + idx / N] = {u[PA[p]][i] * u[PB[p]][i], s[PA[p]][i] * s[PB[p]][i]};
        }
        __syncthreads();
        fft_tile<T, N, NL, NT, false>(buf, twl, tid);
        cx<T>* o = out[p] + zw;
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
            const int idx = tid + i * NT;
            if (ITEMS % NT == 0 || idx < ITEMS) {
                const int l = idx % NL, k = idx / NL;
                if (y0 + l < ny) {
                    const cx<T> zk = buf[k * S + l];
                    const cx<T> zm = buf[((N - k) % N) * S + l];
                    const size_t off = k * krow + z * zrow + y0 + l;
                    cx2<T> q;
                    q.a = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
                    q.b = {T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
                    gst2<kPolicyDefault>(reinterpret_cast<cx2<T>*>(o) + off, q);
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// inverse x (complex-to-real) of pair fields: Z = A + iB per row
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__((Tile<T, N>::NT)) k_pinv_x(PairInvXArgs<T> a) {
    using TL = Tile<T, N>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, S = TL::STRIDE;
    __shared__ cx<T> buf[TL::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * NL;
    const int z = blockIdx.y;
    const long long zw = blockIdx.z * a.ms_w, zp = blockIdx.z * a.ms_p;
    load_twiddles<T, N, NT>(twl, a.tw, tid);
#pragma unroll 1
    for (int p = 0; p < a.npairs; ++p) {
        load_pair_x_lines<T, N, NL>(buf, pick(a.ab, p) + zw, tid, y0, z, a.ny, a.krow, a.zrow);
        __syncthreads();
        fft_tile<T, N, NL, NT, true>(buf, twl, tid);
        T* OA = pick(a.oa, p) + zp;
        T* OB = pick(a.ob, p);
        if (OB) OB += zp;
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
// the x pass of the RHS in one kernel: inverse x of the three velocity pairs
// (u_c, s_c) (the arithmetic of k_pinv_x, scale included), the six products
// in registers, forward x of the product pairs. The physical fields never
// touch HBM (impl.hpp:892-911 forms them in full arrays). The product pairs
// may overwrite the velocity pairs (out[c] == ab[c]): a work-group reads all
// of its lines before it writes any, and work-groups own disjoint (y, z)
// lines.
// ---------------------------------------------------------------------------
template <typename T>
struct PairXArgs {
    const cx<T>* ab[3];  // (u_hat_c, (u_hat phi)_c) after the z and y transforms
    cx<T>* out[6];       // product pairs, same layout
    int ny;
    int nz;
    // pairs between kx rows and between z planes: [kx][z][y][2] (krow =
    // nz*ny, zrow = ny) or [z][kx][y][2] (krow = ny, zrow = (n/2+1)*ny; an x
    // tile's rows then lie within one z plane's n/2+1 rows); unused in the
    // transposed layout (XT)
    long long krow, zrow;
    T scale;
    const cx<T>* tw;
    long long ms_w;
    // if set (register-direct form): max |u|^2 over the grid points of the
    // unshifted velocity, as the bits of non-negative doubles (whose order is
    // that of the unsigned integers) combined with atomicMax into
    // kVmaxSlots slots kVmaxStride apart (one 128-byte line each: one
    // address for all wavefronts serialises them, measured 1.1 ms of a 2.5 ms
    // pass at 512^3); the maximum is that of the slots. Zero them first.
    unsigned long long* vmax2;
};
constexpr int kVmaxSlots = 256;
constexpr int kVmaxStride = 16;

// the kVmaxSlots slots = 0 on the stream (a kernel, so that a captured graph
// holds no memset; launched from the solver's translation unit only)
[[maybe_unused]] static __global__ void k_zero_vmax(unsigned long long* p) {
    if (threadIdx.x < kVmaxSlots) p[threadIdx.x * kVmaxStride] = 0ull;
}

// Register-direct form: the x lines are transformed with the z-pass machinery
// (zfft), lanes over y. The inverse first stage loads its rows straight from
// HBM (rows k > N/2 from row N-k, conjugated: the Hermitian extension), the
// inverse last stage leaves u and s in registers at the slots the forward
// first stage reads (both hold elements j + m*N/E of the lane's line), and
// only the forward output goes through LDS once more, for the pair separation
// of Z_k and Z_{N-k}. Same formulas as the LDS-staged form below (results
// agree to rounding; the compiler may contract multiply-adds differently).
// VMAX: also reduce max |u|^2 of the physical velocity into a.vmax2 (the
// adaptive time step's CFL condition, without transforms of its own).
// XT: the transposed layout [z][y][kx][2] in and out (k_plines_t): the tile's
// NL lines are one contiguous block of NL * (N/2+1) pairs, row k of line l at
// l * (N/2+1) + k.
// One tile (lds holds z_lds<T, N, LINES>() values, the twiddles are in twl).
// lds_b: the second tile, where the kernel declares it apart from the first
// (the prefetch into it then stays in flight across the first tile's
// barriers); else lds + Z::LDS.
template <typename T, int N, int LINES = 0, bool VMAX = false, bool XT = false>
__device__ __forceinline__ void px_products_rd_tile(const PairXArgs<T>& a, TileId tile, cx<T>* lds, const cx<T>* twl,
                                                    int tid, cx<T>* lds_b = nullptr) {
    using Z = ZTile<T, N, LINES>;
    constexpr int NL = Z::NL, NT = Z::NT, E = Z::E, S = Z::S;
    constexpr int T0 = Z::T0, R0 = Z::R0, TLS = Z::TLAST, RL = Z::RL;
    // ping-pong transforms; where the pair separation goes through LDS (no
    // mirrored last stage) it uses tile B, which the next transform writes
    // only after its first barrier. Not at N = 1024, where the second tile's
    // addressing pushes the 128-register x tile further into spills (3%
    // faster at 512; at 1024 the x pass takes 33.4 ms with it, 29.1 without)
    constexpr bool PP = N <= 512 && z_pingpong<T, N, LINES>();
    cx<T>* const tile_b = lds_b ? lds_b : lds + Z::LDS;
    cx<T>* const sep = PP ? tile_b : lds;
    // inverse transforms: first-stage butterflies in mirrored order, so the
    // two reads of each spectral row (elements k and N - k) come from one
    // wavefront (ZTile::j_mirror); measured 2% faster at 512, slower at 1024
    // with its 8-line tiles (x pass 35.0 against 29.1 ms)
    constexpr bool MIR = N <= 512 && Z::kMirrorable;
    // forward transforms: last-stage butterflies in the same mirrored order,
    // so Z_k and Z_{N-k} of a line sit in lanes tid and tid ^ NL of one
    // wavefront and the pair separation is a lane exchange instead of a trip
    // through an LDS tile with two barriers
    constexpr bool MIRL = Z::kMirrorableLast;
    // cache policy of the pair rows: NT in FP32 in the line layout (each row
    // segment is 128 bytes of one tile; 512^3: 2.50 -> 2.40 ms), the default
    // in FP64 (5.6 -> 6.2 ms with NT) and in the transposed layout
    constexpr int XPOL = sizeof(T) == 4 && !XT ? kPolicyStream : kPolicyDefault;
    const int l = tid % NL;
    const int y0 = tile.x * NL;
    const int z = tile.y;
    const bool ok = y0 + l < a.ny;
    const long long zw = tile.z * a.ms_w;
    const uint32_t row = static_cast<uint32_t>(a.krow);  // cx2 elements between k rows
    // per-slot 32-bit addressing fits when STEP rows of a cubic grid span
    // less than 2 GiB (everything but FP64 at N = 1024); the launcher only
    // runs this form with ny * nz <= N * N
    constexpr bool kNarrow = static_cast<long long>(Z::STEP) * N * N * sizeof(cx2<T>) < (1LL << 31);
    constexpr uint32_t H = N / 2 + 1;
    // transposed layout: the tile's lines as one block (line l at l * H)
    const size_t xt_block = (static_cast<size_t>(z) * a.ny + y0) * H;
    // u, s at slot m of element j + m*N/E (j = tid / NL)
    T u[3][E], s[3][E];
    // Where this pass runs without ping-pong tiles but the kernel allocates
    // them (N >= 1024), component 0's u and s are parked in the unused second
    // tile after its inverse transform: 2E registers fewer through the other
    // transforms (at 1024^3 the 128-register FP32 tiles spilled 40-52 bytes
    // per lane to scratch, about 14 GB of HBM traffic per launch). Each
    // thread reads back only its own slots.
    constexpr bool PARK = !PP && z_pingpong<T, N, LINES>();
    static_assert(!PARK || E * NT <= Z::LDS, "the parking area fits the second tile");
    cx<T>* const park = tile_b;
    // Prefetch (DMA): the rows of velocity pairs 1 and 2 are loaded straight
    // into tile B (global-to-LDS buffer loads, [k][line] cx2) while the
    // previous pair's inverse transform runs on tile A alone with LDS-only
    // barriers; the third transform and the forward ones use both tiles.
    constexpr int PRE_ENT = H * NL;               // cx2 entries of one pair's rows
    constexpr int PRE_INS = (PRE_ENT + 63) / 64;  // 64-lane load instructions
    // (where the rows fit the second tile: the padded 16-line FP32 tiles)
    constexpr bool DMA = kXPrefetch && sizeof(T) == 4 && PP && !XT && kNarrow && PRE_INS * 64 * 2 <= Z::LDS;
    cx2<T>* const pre = reinterpret_cast<cx2<T>*>(tile_b);
    auto prefetch = [&](int c) {
        const cx2<T>* AB = reinterpret_cast<const cx2<T>*>(a.ab[c] + zw) + static_cast<size_t>(z) * a.zrow + y0;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<cx2<T>*>(AB), 0, static_cast<int>(((H - 1) * row + NL) * sizeof(cx2<T>)), 0x00020000);
        // a fixed, unrolled count per wave (the remainder under a uniform
        // branch), so that the compiler's waits for the first pair's own
        // loads, issued before these, leave the prefetch in flight
        constexpr int NW = NT / 64;
        constexpr int FULL = PRE_INS / NW;
        const int wave = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;
        auto issue = [&](int q) {
            const int e = q * 64 + lane;
            const int k = e / NL, ll = e % NL;
            // rows past the last line and the padding lanes read out of range: zeros
            const uint32_t off = e < PRE_ENT && y0 + ll < a.ny
                                     ? (static_cast<uint32_t>(k) * row + ll) * static_cast<uint32_t>(sizeof(cx2<T>))
                                     : 0xFFFFFFC0u;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(pre + q * 64), 16,
                                                     off, 0, 0, XPOL);
        };
#pragma unroll
        for (int it = 0; it < FULL; ++it) issue(wave + it * NW);
        if (FULL * NW + wave < PRE_INS) issue(FULL * NW + wave);
    };
    auto uval = [&](int c, int m) -> T { return PARK && c == 0 ? park[m * NT + tid].x : u[c][m]; };
    auto sval = [&](int c, int m) -> T { return PARK && c == 0 ? park[m * NT + tid].y : s[c][m]; };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (DMA && c > 0) {
            // pair c's rows are in tile B in every wave
            __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
            __syncthreads();
        }
        const cx2<T>* AB = reinterpret_cast<const cx2<T>*>(a.ab[c] + zw) + (XT ? xt_block : static_cast<size_t>(z) * a.zrow + y0);
        // XT: one resource for the tile's block; lanes past the last line read zeros
        const __amdgpu_buffer_rsrc_t rsx = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<cx2<T>*>(AB), 0, static_cast<int>(NL * H * sizeof(cx2<T>)), 0x00020000);
        cx<T> v[1][T0][R0], w[1][TLS][RL];
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                // slot m holds element e = j + m*STEP: over the work-group the
                // rows k = min(e, N - e) span at most STEP rows from kmin, so
                // one resource per slot (based at row kmin) and a 32-bit lane
                // offset address them; lanes past the last line read zeros
                constexpr int STEP = Z::STEP;
                const int m = i + r * T0;
                const int lo = m * STEP, hi = (m + 1) * STEP - 1;
                const int kmin = hi <= N / 2 ? lo : (lo > N / 2 ? N - hi : (lo < N - hi ? lo : N - hi));
                const int e = MIR ? Z::e_first_mirror(tid, r) : Z::e_first(tid, i, r);
                const int k = e <= N / 2 ? e : N - e;
                cx2<T> q;
                if (DMA && c > 0) {
                    q = ok ? pre[k * NL + l] : cx2<T>{};
                } else if constexpr (XT) {
                    const uint32_t off = ok ? (l * H + k) * static_cast<uint32_t>(sizeof(cx2<T>)) : 0xFFFFFFC0u;
                    q = ld_cx2<T, XPOL>(rsx, off);
                } else if constexpr (kNarrow) {
                    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                        const_cast<cx2<T>*>(AB + static_cast<size_t>(kmin) * row), 0,
                        static_cast<int>(STEP * row * sizeof(cx2<T>)), 0x00020000);
                    const uint32_t off =
                        ok ? ((k - kmin) * row + l) * static_cast<uint32_t>(sizeof(cx2<T>)) : 0xFFFFFFC0u;
                    q = ld_cx2<T, XPOL>(rs, off);
                } else {
                    q = AB[static_cast<size_t>(k) * row + (ok ? l : 0)];
                    if (!ok) q = cx2<T>{};
                }
                cx<T> va = q.a, vb = q.b;
                if (k == 0 || k == N / 2) {
                    va.y = T(0);
                    vb.y = T(0);
                }
                v[0][i][r] = e <= N / 2 ? cx<T>{va.x - vb.y, va.y + vb.x} : cx<T>{va.x + vb.y, vb.x - va.y};
            }
        if (DMA && c == 0) prefetch(1);
        if (DMA && c == 1) {
            __syncthreads();  // every wave has its pair-1 rows out of tile B
            prefetch(2);
        }
        if (DMA && c < 2)
            zfft<T, N, 1, true, LINES, false, MIR, false, true>(v, w, lds, twl, tid);
        else
            zfft<T, N, 1, true, LINES, PP, MIR>(v, w, lds, twl, tid, tile_b);
#pragma unroll
        for (int i = 0; i < TLS; ++i)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                u[c][i + r * TLS] = w[0][i][r].x * a.scale;
                s[c][i + r * TLS] = w[0][i][r].y * a.scale;
            }
        if constexpr (PARK) {
            if (c == 0) {
#pragma unroll
                for (int m = 0; m < E; ++m) park[m * NT + tid] = {u[0][m], s[0][m]};
            }
        }
    }
    if constexpr (VMAX) {
        // |u|^2 in double as the diagnostics' reduction forms it (k_reduce_phys);
        // lanes past the last line hold zeros
        double m = 0.0;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const double p = uval(0, i), q = u[1][i], r = u[2][i];
            m = fmax(m, p * p + q * q + r * r);
        }
        m = wave_max(m);
        const unsigned slot = (blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) % kVmaxSlots;
        if (tid % 64 == 0)
            atomicMax(a.vmax2 + slot * kVmaxStride, static_cast<unsigned long long>(__double_as_longlong(m)));
    }
    constexpr int PA[6] = {0, 0, 0, 1, 1, 2};
    constexpr int PB[6] = {0, 1, 2, 1, 2, 2};
    constexpr int ITEMS = NL * (N / 2 + 1);
    constexpr int IPT = (ITEMS + NT - 1) / NT;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        cx<T> v[1][T0][R0], w[1][TLS][RL];
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const int m = i + r * T0;
                v[0][i][r] = {uval(PA[p], m) * uval(PB[p], m), sval(PA[p], m) * sval(PB[p], m)};
            }
        zfft<T, N, 1, false, LINES, PP, false, MIRL>(v, w, lds, twl, tid, tile_b);
        if constexpr (MIRL) {
            // element j + r*STEP for r < RL/2 (and N/2 in group j = 0): rows
            // k = 0 .. N/2 once over the work-group
            constexpr int HR = RL / 2;
            const int j = Z::j_mirror(tid);
            const bool self0 = j == 0, selfh = j == Z::STEP / 2;
            cx2<T>* o = reinterpret_cast<cx2<T>*>(a.out[p] + zw) +
                        (XT ? xt_block + static_cast<size_t>(l) * H : static_cast<size_t>(z) * a.zrow + y0 + l);
            const size_t ostep = XT ? 1 : row;  // between rows k
#pragma unroll
            for (int r = 0; r <= HR; ++r) {
                const cx<T> zk = w[0][0][r % RL];
                cx<T> zm;
                if (r < HR) {
                    const cx<T> mine = w[0][0][RL - 1 - r];
                    cx<T> other;
                    other.x = __shfl_xor(mine.x, NL);
                    other.y = __shfl_xor(mine.y, NL);
                    zm = self0 ? w[0][0][(RL - r) % RL] : (selfh ? mine : other);
                } else {
                    zm = zk;  // k = N/2, group j = 0 only
                }
                cx2<T> q;
                q.a = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
                q.b = {T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
                if (ok && (r < HR || self0)) gst2<XPOL>(o + static_cast<size_t>(j + r * Z::STEP) * ostep, q);
            }
            continue;
        }
        __syncthreads();  // the last stage's reads of the tile are done
#pragma unroll
        for (int i = 0; i < TLS; ++i)
#pragma unroll
            for (int r = 0; r < RL; ++r) sep[Z::e_last(tid, i, r) * S + l] = w[0][i][r];
        __syncthreads();
        cx2<T>* o = reinterpret_cast<cx2<T>*>(a.out[p] + zw);
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
            const int idx = tid + i * NT;
            if (ITEMS % NT == 0 || idx < ITEMS) {
                const int l2 = idx % NL, k = idx / NL;
                if (y0 + l2 < a.ny) {
                    const cx<T> zk = sep[k * S + l2];
                    const cx<T> zm = sep[((N - k) % N) * S + l2];
                    cx2<T> q;
                    q.a = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
                    q.b = {T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
                    if constexpr (XT)
                        gst2<XPOL>(o + xt_block + static_cast<size_t>(l2) * H + k, q);
                    else
                        gst2<XPOL>(o + k * row + static_cast<size_t>(z) * a.zrow + y0 + l2, q);
                }
            }
        }
        // the next product's zfft synchronises before it writes the tile
    }
}

template <typename T, int N, int LINES = 0, bool VMAX = false, bool XT = false>
// The second bound is the minimum number of waves per SIMD: 512-thread tiles
// ask for 4, i.e. two work-groups per CU (16 waves), which caps the registers
// at 128 like the 1024-thread form.
__global__ void __launch_bounds__((ZTile<T, N, LINES>::NT), (ZTile<T, N, LINES>::kMinWaves))
    k_px_products_rd(PairXArgs<T> a) {
    using Z = ZTile<T, N, LINES>;
    constexpr bool kTwoTiles = z_lds<T, N, LINES>() == 2 * Z::LDS;
    // the two tiles as separate arrays, so that the compiler sees the
    // prefetch into the second one apart from the first one's accesses
    __shared__ cx<T> lds[Z::LDS];
    __shared__ cx<T> lds_b[kTwoTiles ? Z::LDS : 1];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    load_twiddles<T, N, Z::NT>(twl, a.tw, tid);
    // consecutive tiles of an XCD: neighbouring y blocks of one z, i.e. the
    // same translation pages of every k row
    px_products_rd_tile<T, N, LINES, VMAX, XT>(a, xcd_tile_id(), lds, twl, tid, kTwoTiles ? lds_b : nullptr);
}

// LDS-staged form: each transform stages its whole tile through LDS.
template <typename T, int N, int LINES = default_lines<T, N>()>
__global__ void __launch_bounds__((Tile<T, N, LINES>::NT)) k_px_products(PairXArgs<T> a) {
    using TL = Tile<T, N, LINES>;
    constexpr int NL = TL::NL, NT = TL::NT, E = TL::E, S = TL::STRIDE;
    __shared__ cx<T> buf[TL::LDS];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const TileId tile = xcd_tile_id();  // as in k_px_products_rd
    const int y0 = tile.x * NL;
    const int z = tile.y;
    const long long zw = tile.z * a.ms_w;
    load_twiddles<T, N, NT>(twl, a.tw, tid);
    T u[3][E], s[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        load_pair_x_lines<T, N, NL>(buf, a.ab[c] + zw, tid, y0, z, a.ny, a.krow, a.zrow);
        __syncthreads();
        fft_tile<T, N, NL, NT, true>(buf, twl, tid);
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int idx = tid + i * NT;
            const int x = idx % N, l = idx / N;
            cx<T> v = czero<T>();
            if ((TL::ELEMS % NT == 0 || idx < TL::ELEMS) && y0 + l < a.ny) v = buf[x * S + l];
            u[c][i] = v.x * a.scale;
            s[c][i] = v.y * a.scale;
        }
        __syncthreads();
    }
    products_forward_x<T, N, NL, E>(u, s, buf, twl, a.out, zw, tid, y0, z, a.ny, a.krow, a.zrow);
}

// ---------------------------------------------------------------------------
// forward z of the six product pairs + RHS completion + RK3 update (one tile;
// the twiddles are in twl already)
// ---------------------------------------------------------------------------
// stash (optional): LDS for the updated state (z_stash_at), for the next
// stage's inverse z in the same work-group.
template <typename T, int N, int LINES = 0>
__device__ __forceinline__ void pfwd_z_rk_tile(const PairFwdZArgs<T>& a, TileId tile, cx<T>* lds, const cx<T>* twl,
                                               int tid, cx<T>* stash = nullptr) {
    using Z = ZTile<T, N, LINES>;
    constexpr int NL = Z::NL, T0 = Z::T0, R0 = Z::R0, TLS = Z::TLAST, RL = Z::RL;
    constexpr int SL = TLS * RL;  // last-stage slots per lane
    constexpr int HS = SL / 2;    // modes owned by each lane of a line
    static_assert(SL % 2 == 0, "the two lanes of a line split the slots");
    const int l = tid % NL, f = l & 1;
    const int ky = tile.x * (NL / 2) + (l >> 1);
    const int kx = tile.y;
    const bool ok = ky < a.ny;
    const uint32_t kyc = ok ? ky : a.ny - 1;
    constexpr uint32_t CB = sizeof(cx<T>);
    const uint32_t ny = a.ny;
    const long long zw = tile.z * a.ms_w;
    cx<T>* const u3 = a.u3 + tile.z * a.ms_u;
    cx<T>* const g3 = a.g3 + tile.z * a.ms_u;
    const T* const coef = a.coef + tile.z * kCoefStride;
    const size_t plane = static_cast<size_t>(kx) * N * ny;
    const ZRow zr(kyc, a.kb > 0 ? a.kb : ny, N);
    const uint32_t p_lane = ((zr.base + (tid / NL) * zr.step) * 2 + f) * CB, p_row = zr.step * 2 * CB;
    const uint32_t p_bytes = N * ny * 2 * CB;
    const int my = signed_mode(kyc, a.n);
    const T fx = T(kx), fy = T(my);
    // element (kz) of owned mode h of this lane: slot h (f = 0) or h + HS (f = 1)
    auto owned_e = [&](int h) {
        const int lo = Z::e_last(tid, h / RL, h % RL), hi = Z::e_last(tid, (h + HS) / RL, (h + HS) % RL);
        return f ? hi : lo;
    };
    // per owned mode: signed kz, |k|^2, mask and the conjugate phase, looked up
    // once (not per product pair), and the stage coefficients, so that no
    // dependent global load sits between a transform and its accumulation
    // (four registers per mode: kz as T, the combine's factor 1/2 or 0 for a
    // masked mode, the conjugate phase; the integer |k|^2 and mask are formed
    // again for the update)
    T fzh[HS], halfh[HS];
    cx<T> phc[HS];
#pragma unroll
    for (int h = 0; h < HS; ++h) {
        const int mz = signed_mode(owned_e(h), a.n);
        fzh[h] = T(mz);
        halfh[h] = keep_mode(kx * kx + my * my + mz * mz, a.n) ? T(0.5) : T(0);
        const cx<T> ph = a.phase[kx + my + mz + a.n];
        phc[h] = {ph.x, -ph.y};
    }
    const T alpha = coef[2 * a.stage];
    const T beta = coef[2 * a.stage + 1];
    const T nu = coef[6];
    cx<T> r0[HS], r1[HS], r2[HS];
#pragma unroll
    for (int h = 0; h < HS; ++h) r0[h] = r1[h] = r2[h] = czero<T>();
    cx<T> v[1][T0][R0], w[1][TLS][RL], vn[T0][R0];
    auto load_pair = [&](int p, cx<T> (&dst)[T0][R0]) {
        const Buf<T, z_pair_load_policy<T, N>()> b(pick(a.in, p) + zw + plane * 2, p_bytes);
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r) dst[i][r] = b.ld(p_lane + (i * Z::STEP + r * (N / R0)) * p_row);
    };
    load_pair(0, vn);
    // products in the reference's accumulation order (00,01,02,11,12,22);
    // the next pair's loads are issued before the current pair's transform
#pragma unroll 1
    for (int p = 0; p < 6; ++p) {
        const int pa = p < 3 ? 0 : (p < 5 ? 1 : 2);
        const int pb = p < 3 ? p : (p < 5 ? p - 2 : 2);
#pragma unroll
        for (int i = 0; i < T0; ++i)
#pragma unroll
            for (int r = 0; r < R0; ++r) v[0][i][r] = vn[i][r];
        if (p < 5) load_pair(p + 1, vn);
        zfft<T, N, 1, false, LINES, z_pingpong<T, N, LINES>()>(v, w, lds, twl, tid);
        // lane f = 0 holds P_hat, f = 1 holds Q_hat of every slot; swap halves
#pragma unroll
        for (int h = 0; h < HS; ++h) {
            const cx<T> lo = w[0][h / RL][h % RL], hi = w[0][(h + HS) / RL][(h + HS) % RL];
            const cx<T> mine = f ? hi : lo;
            const cx<T> other = shfl_xor1(f ? lo : hi);
            const cx<T> P = f ? other : mine;
            const cx<T> Q = f ? mine : other;
            // 0.5 (P + Q conj(phi)), zero for a masked mode (its accumulators
            // are zeroed by the projection below whatever they hold)
            const cx<T> val = (P + cmul(Q, phc[h])) * halfh[h];
            const T fz = fzh[h];
            const T kb = pb == 0 ? fx : (pb == 1 ? fy : fz);
            const T ka = pa == 0 ? fx : (pa == 1 ? fy : fz);
            const cx<T> da = {kb * val.y, -kb * val.x};  // -i k_b v into r_a
            const cx<T> db = {ka * val.y, -ka * val.x};  // -i k_a v into r_b
            // branch-free: at most one of the two terms of each sum is
            // nonzero, so every accumulator receives exactly the terms (and
            // roundings) of the reference's order
            r0[h] = r0[h] + (pa == 0 ? da : czero<T>());
            r1[h] = r1[h] + ((pa == 1 ? da : czero<T>()) + (pb == 1 && pa != 1 ? db : czero<T>()));
            r2[h] = r2[h] + ((pa == 2 ? da : czero<T>()) + (pb == 2 && pa != 2 ? db : czero<T>()));
        }
    }
    // the stash overlays the transform tiles: all waves are past the last
    // transform's reads before any lane writes it
    if (stash) __syncthreads();
    if (!ok) return;
    const bool first = a.stage == 0;
    const Buf<T> U(u3 + plane * 3, N * ny * 3 * CB);
    const Buf<T> G(g3 + plane * 3, N * ny * 3 * CB);
    // all loads of the update first (the buffer accesses may alias as far as
    // the compiler knows, so loads placed after stores would wait for them)
    uint32_t off[HS];
    cx<T> u[HS][3], g[HS][3];
#pragma unroll
    for (int h = 0; h < HS; ++h) {
        off[h] = (zr.base + owned_e(h) * zr.step) * 3 * CB;
#pragma unroll
        for (int c = 0; c < 3; ++c) u[h][c] = U.ld(off[h] + c * CB);
    }
#pragma unroll
    for (int h = 0; h < HS; ++h)
#pragma unroll
        for (int c = 0; c < 3; ++c) g[h][c] = first ? czero<T>() : G.ld(off[h] + c * CB);
#pragma unroll
    for (int h = 0; h < HS; ++h) {
        const int mz = static_cast<int>(fzh[h]);
        const int k2 = kx * kx + my * my + mz * mz;
        const bool keep = keep_mode(k2, a.n);
        cx<T> rr[3] = {r0[h], r1[h], r2[h]};
        const T vk2 = nu * T(k2);
#pragma unroll
        for (int c = 0; c < 3; ++c) rr[c] = rr[c] - u[h][c] * vk2;
        project_or_zero(rr[0], rr[1], rr[2], kx, my, mz, k2, keep);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            g[h][c] = first ? rr[c] : g[h][c] * alpha + rr[c];
            u[h][c] = u[h][c] + g[h][c] * beta;
        }
        project_or_zero(u[h][0], u[h][1], u[h][2], kx, my, mz, k2, keep);
    }
    // g after the last stage is never read (the next step's first stage sets
    // g = RHS without reading it, impl.hpp:994-1000): stage 2 does not store it
    const bool keep_g = a.stage != 2;
#pragma unroll
    for (int h = 0; h < HS; ++h)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (keep_g) G.st(off[h] + c * CB, g[h][c]);
            U.st(off[h] + c * CB, u[h][c]);
        }
    if (stash) {
#pragma unroll
        for (int h = 0; h < HS; ++h)
#pragma unroll
            for (int c = 0; c < 3; ++c) stash[z_stash_at<N, NL>(c, owned_e(h), l >> 1)] = u[h][c];
    }
}

// The last pass of an RK stage: forward z + RHS completion + update of the
// tile's (kx, ky) lines, then the next stage's inverse z of the same lines
// (impl.hpp:979-1045 runs compute_rhs's backward transforms of u_hat at the
// start of every stage; here they follow the update that produced u_hat, so
// the state is not read from HBM again and one launch per stage goes away).
// The inverse half writes the velocity pairs into
// inv.out[0..2] (the first three product pair fields, whose values the
// forward half has consumed: every work-group reads its own lines only).
template <typename T, int N, int LINES = 0>
__global__ void __launch_bounds__((ZTile<T, N, LINES>::NT), (ZTile<T, N, LINES>::kMinWaves))
    k_pfwd_z_rk_inv(PairFwdZArgs<T> a, PairInvZArgs<T> inv) {
    using Z = ZTile<T, N, LINES>;
    __shared__ cx<T> lds[z_lds<T, N, LINES>()];
    __shared__ cx<T> twl[N];
    const int tid = threadIdx.x;
    const TileId tile = xcd_tile_id();
    load_twiddles<T, N, Z::NT>(twl, a.tw, tid);
    // the updated state goes to the next stage's inverse z through LDS when
    // 3 * N * NL/2 values fit in the kernel's LDS (the stash overlays the
    // transform tiles), otherwise it is read back from global memory (L2).
    // Work-group scope is enough for that read-back: the barrier's release /
    // acquire fences make every wave's state stores visible to the others
    // (one CU, one L1); a device-scope fence would write back the XCD's
    // whole L2 per work-group (measured: 6x slower pass).
    constexpr bool kStash = z_stash_fits<T, N, LINES>();
    pfwd_z_rk_tile<T, N, LINES>(a, tile, lds, twl, tid, kStash ? lds : nullptr);
    __syncthreads();
    pinv_z_tile<T, N, PZ_RHS, LINES>(inv, tile, lds, twl, tid, kStash ? lds : nullptr);
}

// max |k . u_hat| over the component-interleaved state (impl.hpp:1167-1170)
template <typename T>
// (state in ky blocks of kb; kb = n: the plain layout)
__global__ void __launch_bounds__(kReduceThreads) k_reduce_div3(const cx<T>* u3, int n, int kb, double* partial) {
    const long long count = static_cast<long long>(n / 2 + 1) * n * n;
    double m = 0.0;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        int kx, kz, ky;
        spectral_mode_of(i, n, kb, kx, kz, ky);
        const T fx = T(kx), fy = T(signed_mode(ky, n)), fz = T(signed_mode(kz, n));
        const cx<T> d = (u3[3 * i] * fx + u3[3 * i + 1] * fy) + u3[3 * i + 2] * fz;
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

}  // namespace sph

// Host-side dispatch of the pass kernels: one table of launchers per
// (storage type, transform length). The kernels themselves are compiled in
// passes_f32.hip / passes_f64.hip.
#pragma once

#include "pair_passes.hpp"

namespace sph {

// Variants of the tuned passes (inverse y, x, forward y, forward z + update +
// next inverse z) that the solver times on its own buffers, keeping the
// fastest (Group::autotune):
//   y passes: LDS-staged tiles of the default width, half and a quarter of
//             it (512^3 FP32 keeps the quarter: more, smaller work-groups per
//             CU hide the latency of the tile loads better);
//   x pass:   register-direct tiles of the default z-tile width, then half of
//             it, then the LDS-staged form at the default width, then
//             register-direct tiles of the default width in the alternative
//             radix order (z_radix_plan_alt; lengths that have one);
//   forward z: the default z-tile width, half of it (half: twice the
//             work-groups per CU where they fit), then the default width in
//             the alternative radix order (the standalone inverse z takes the
//             same width and order).
constexpr int kTunedPasses = 4;
constexpr int kMaxVariants = 4;
// the x-pass variant of the LDS-staged form and of the alternative radix
// order, the forward-z variant of the latter
constexpr int kXLdsVariant = 2;
constexpr int kXOrderVariant = 3;
constexpr int kZOrderVariant = 2;

template <typename T>
struct PassTable {
    int n;   // transform length
    int nl;  // lines per tile
    int nt;  // threads per work-group
    int variants[kTunedPasses];                // variants of each tuned pass
    int tile_lines[kTunedPasses][kMaxVariants];  // lines per tile of each variant
    // radix order of each variant (0: z_radix_plan, 1: z_radix_plan_alt; the
    // y passes always 0)
    int radix_order[kTunedPasses][kMaxVariants];
    void (*inv_z)(const InvZArgs<T>&, int nky, int nkx, hipStream_t);
    void (*lines)(cx<T>* data, long long nlines, bool inverse, const cx<T>* tw, hipStream_t);
    void (*inv_x)(const InvXArgs<T>&, int ny, int nz, hipStream_t);
    void (*fwd_x)(const FwdXArgs<T>&, int ny, int nz, hipStream_t);
    void (*fwd_z)(const FwdZArgs<T>&, int mode, int nky, int nkx, hipStream_t);
    // solver stage on pair-interleaved buffers (pinv_z, pfwd_z_rk: variant =
    // z-tile width, as the forward-z pass's variants)
    void (*pinv_z)(const PairInvZArgs<T>&, int mode, int nky, int nkx, int members, int variant, hipStream_t);
    void (*plines)(cx<T>* data, long long npairlines, long long member_stride, int members, bool inverse,
                   const cx<T>* tw, int variant, hipStream_t);
    void (*pinv_x)(const PairInvXArgs<T>&, int ny, int nz, int members, hipStream_t);
    // xt: the transposed work layout (register-direct variants 0, 1 and
    // kXOrderVariant only)
    void (*px_products)(const PairXArgs<T>&, int ny, int nz, int members, int variant, bool xt, hipStream_t);
    // the register-direct x pass of one cubic solve that also reduces max
    // |u|^2 into a.vmax2 (variants 0, 1 and kXOrderVariant; px_products_vmax)
    void (*px_products_vmax)(const PairXArgs<T>&, int variant, bool xt, hipStream_t);
    // forward z + update, then the next stage's inverse z (k_pfwd_z_rk_inv)
    void (*pfwd_z_rk)(const PairFwdZArgs<T>&, const PairInvZArgs<T>&, int nky, int nkx, int members, int variant,
                      hipStream_t);
    // transposed work layout between the y passes and the x pass (k_plines_t):
    // supported (xt), its y-pass tile widths (variants 0..2), and the pass
    // itself (line layout in and transposed out when inverse, else the other
    // way round; nfields pair fields field_stride apart, members
    // member_stride apart)
    bool xt;
    int xt_lines[kMaxVariants];
    void (*plines_t)(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
                     int members, bool inverse, const cx<T>* tw, int variant, hipStream_t);
    // ky-blocked spectral layout (k_plines_b): supported (zb), its ky block
    // (the default z tile's NL/2), and the y pass between it and the line
    // layout of the x pass (blocked in and line layout out when inverse;
    // variants 0..2: the line layout's y-tile widths)
    bool zb;
    int zb_kb;
    // (xs_kx, xs_z: complex values between kx lines and between z lines of
    // the line side)
    void (*plines_b)(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
                     long long xs_kx, long long xs_z, int members, bool inverse, const cx<T>* tw, int variant,
                     hipStream_t);
};

// nullptr when n is not a compiled length.
template <typename T>
const PassTable<T>* pass_table(int n);

// Defined in passes_vmax.hip, a translation unit of its own: compiled in
// the same module as the plain x pass, its kernels changed that pass's
// register allocation (a 12-byte spill per thread at 512^3, 0.13 GB more
// HBM writes per launch).
template <typename T, int N>
void px_products_vmax(const PairXArgs<T>& a, int variant, bool xt, hipStream_t s);

// The compiled transform lengths: powers of two and twice or four times them
// times 3 up to 1024 (SPH_LENGTHS_A), and lengths with a factor 5 and the
// rest of the 2^a 3 ones up to 768 (SPH_LENGTHS_B), compiled in translation
// units of their own (passes_f32.hip, passes_f32_b.hip and the FP64 pair).
#define SPH_LENGTHS_A(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(256) X(512) X(1024)
#define SPH_LENGTHS_B(X) X(80) X(160) X(192) X(320) X(384) X(640) X(768)
#define SPH_LENGTHS(X) SPH_LENGTHS_A(X) SPH_LENGTHS_B(X)

// The pass tables of one translation unit's lengths (part 0: A, part 1: B).
template <typename T, int PART>
const PassTable<T>* pass_table_part(int n);

template <typename T, int N>
struct Launch {
    using TL = Tile<T, N>;
    using ZT = ZTile<T, N>;
    static void inv_z(const InvZArgs<T>& a, int nky, int nkx, hipStream_t s) {
        const dim3 grid((nky + ZT::NL - 1) / ZT::NL, nkx);
        k_inv_z<T, N><<<grid, ZT::NT, 0, s>>>(a);
    }
    static void lines(cx<T>* data, long long nlines, bool inverse, const cx<T>* tw, hipStream_t s) {
        const dim3 grid(static_cast<unsigned>((nlines + TL::NL - 1) / TL::NL));
        if (inverse)
            k_lines<T, N, true><<<grid, TL::NT, 0, s>>>(data, nlines, tw);
        else
            k_lines<T, N, false><<<grid, TL::NT, 0, s>>>(data, nlines, tw);
    }
    static void inv_x(const InvXArgs<T>& a, int ny, int nz, hipStream_t s) {
        const dim3 grid((ny + TL::NL - 1) / TL::NL, nz);
        k_inv_x<T, N><<<grid, TL::NT, 0, s>>>(a);
    }
    static void fwd_x(const FwdXArgs<T>& a, int ny, int nz, hipStream_t s) {
        const dim3 grid((ny + TL::NL - 1) / TL::NL, nz);
        k_fwd_x<T, N><<<grid, TL::NT, 0, s>>>(a);
    }
    static void fwd_z(const FwdZArgs<T>& a, int mode, int nky, int nkx, hipStream_t s) {
        const dim3 grid((nky + ZT::NL - 1) / ZT::NL, nkx);
        if (mode == FZ_PROJECT)
            k_fwd_z<T, N, FZ_PROJECT><<<grid, ZT::NT, 0, s>>>(a);
        else
            k_fwd_z<T, N, FZ_PLAIN><<<grid, ZT::NT, 0, s>>>(a);
    }
    // z tiles: the default width and half of it when that still makes a
    // whole number of wavefronts
    static constexpr int Z0 = ZT::NL;
    static constexpr int Z1 = Z0 / 2 >= 2 && (Z0 / 2) * N / ZT::E >= 64 ? Z0 / 2 : Z0;
    // the alternative radix order at its own default width
    static constexpr bool kAlt = z_plan_has_alt(N);
    static constexpr int ZA = zlines(0, 1);
    static constexpr int ZA_LINES = ZTile<T, N, ZA>::NL;
    // the x pass's alternative-order variant: at 1024 in FP32 (8*16*8) at
    // half the default width, 4 lines of 256 threads (1024^3: 19.1-19.6
    // against 23.2 ms at the default width; the forward-z pass keeps the
    // default width: 32.3 against 38.1 ms)
    static constexpr int XA = N == 1024 && sizeof(T) == 4 ? zlines(ZT::NL / 2, 1) : ZA;
    static constexpr int XA_LINES = ZTile<T, N, XA>::NL;
    template <int L>
    static void pinv_z_l(const PairInvZArgs<T>& a, int mode, int nky, int nkx, int members, hipStream_t s) {
        using ZV = ZTile<T, N, L>;
        const dim3 grid((nky + ZV::NL / 2 - 1) / (ZV::NL / 2), nkx, members);
        if (mode == PZ_RHS)
            k_pinv_z<T, N, PZ_RHS, L><<<grid, ZV::NT, 0, s>>>(a);
        else if (mode == PZ_DIAG)
            k_pinv_z<T, N, PZ_DIAG, L><<<grid, ZV::NT, 0, s>>>(a);
        else
            k_pinv_z<T, N, PZ_VEL, L><<<grid, ZV::NT, 0, s>>>(a);
    }
    static void pinv_z(const PairInvZArgs<T>& a, int mode, int nky, int nkx, int members, int variant,
                       hipStream_t s) {
        if (variant == 1 && mode == PZ_RHS)
            pinv_z_l<Z1>(a, mode, nky, nkx, members, s);
        else if (kAlt && variant == kZOrderVariant && mode == PZ_RHS)
            pinv_z_l<ZA>(a, mode, nky, nkx, members, s);
        else
            pinv_z_l<Z0>(a, mode, nky, nkx, members, s);
    }
    static constexpr int L0 = TL::NL;
    static constexpr int L1 = L0 / 2 >= 2 ? L0 / 2 : L0;
    static constexpr int L2 = L1 / 2 >= 2 ? L1 / 2 : L1;
    template <int L>
    static void plines_l(cx<T>* data, long long npairlines, long long member_stride, int members, bool inverse,
                         const cx<T>* tw, hipStream_t s) {
        using TV = Tile<T, N, L>;
        const dim3 grid(static_cast<unsigned>((npairlines + L / 2 - 1) / (L / 2)), members);
        if (inverse)
            k_plines<T, N, true, L><<<grid, TV::NT, 0, s>>>(data, npairlines, member_stride, tw);
        else
            k_plines<T, N, false, L><<<grid, TV::NT, 0, s>>>(data, npairlines, member_stride, tw);
    }
    static void plines(cx<T>* data, long long npairlines, long long member_stride, int members, bool inverse,
                       const cx<T>* tw, int variant, hipStream_t s) {
        if (variant == 2)
            plines_l<L2>(data, npairlines, member_stride, members, inverse, tw, s);
        else if (variant == 1)
            plines_l<L1>(data, npairlines, member_stride, members, inverse, tw, s);
        else
            plines_l<L0>(data, npairlines, member_stride, members, inverse, tw, s);
    }
    static void pinv_x(const PairInvXArgs<T>& a, int ny, int nz, int members, hipStream_t s) {
        const dim3 grid((ny + TL::NL - 1) / TL::NL, nz, members);
        k_pinv_x<T, N><<<grid, TL::NT, 0, s>>>(a);
    }
    // register-direct x tiles: default z-tile width and half of it when that
    // still makes a whole number of wavefronts
    static constexpr int X0 = ZT::NL;
    static constexpr int X1 = X0 / 2 >= 2 && (X0 / 2) * N / ZT::E >= 64 ? X0 / 2 : X0;
    template <int L, bool XT = false>
    static void px_products_rd(const PairXArgs<T>& a, int ny, int nz, int members, hipStream_t s) {
        using ZV = ZTile<T, N, L>;
        const dim3 grid((ny + ZV::NL - 1) / ZV::NL, nz, members);
        k_px_products_rd<T, N, L, false, XT><<<grid, ZV::NT, 0, s>>>(a);
    }
    // the transposed work layout (k_plines_t, k_px_products_rd<..., XT>) at
    // the lengths where the x pass's rows lie far apart
    static constexpr bool kXT = N >= 256;
    // its y tiles: the default width, half and twice of it (twice: 128-byte
    // transposed rows at n = 1024 in FP32, where the default is 8 lines)
    static constexpr int T0L = TL::NL;
    static constexpr int T1L = T0L / 2 >= 2 ? T0L / 2 : T0L;
    static constexpr int T2L =
        (tile_lds<2 * T0L, sizeof(T) == 4 && 2 * T0L <= 8>(N) + N) * static_cast<int>(sizeof(cx<T>)) <= kLdsBytes
            ? 2 * T0L
            : T0L;
    template <int L>
    static void plines_t_l(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
                           int members, bool inverse, const cx<T>* tw, hipStream_t s) {
        using TV = Tile<T, N, L>;
        constexpr int H = N / 2 + 1;
        const dim3 grid((H + L / 2 - 1) / (L / 2), N, nfields * members);
        if (inverse)
            k_plines_t<T, N, true, L><<<grid, TV::NT, 0, s>>>(in, out, nfields, field_stride, member_stride, tw);
        else
            k_plines_t<T, N, false, L><<<grid, TV::NT, 0, s>>>(in, out, nfields, field_stride, member_stride, tw);
    }
    static void plines_t(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
                         int members, bool inverse, const cx<T>* tw, int variant, hipStream_t s) {
        if constexpr (kXT) {
            if (variant == 2)
                plines_t_l<T2L>(in, out, nfields, field_stride, member_stride, members, inverse, tw, s);
            else if (variant == 1)
                plines_t_l<T1L>(in, out, nfields, field_stride, member_stride, members, inverse, tw, s);
            else
                plines_t_l<T0L>(in, out, nfields, field_stride, member_stride, members, inverse, tw, s);
        }
    }
    // the ky-blocked spectral layout, at the lengths whose ky blocks and y
    // tiles divide them (the one-launch ensemble kernels assume the same
    // bound)
    static constexpr bool kZB = N >= 64;
    template <int L>
    static void plines_b_l(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
                           long long xs_kx, long long xs_z, int members, bool inverse, const cx<T>* tw,
                           hipStream_t s) {
        using TV = Tile<T, N, L>;
        constexpr long long H = N / 2 + 1;
        const dim3 grid(static_cast<unsigned>(H * N / (L / 2)), 1, nfields * members);
        if (inverse)
            k_plines_b<T, N, true, L><<<grid, TV::NT, 0, s>>>(in, out, nfields, field_stride, member_stride, xs_kx,
                                                              xs_z, tw);
        else
            k_plines_b<T, N, false, L><<<grid, TV::NT, 0, s>>>(in, out, nfields, field_stride, member_stride, xs_kx,
                                                               xs_z, tw);
    }
    static void plines_b(const cx<T>* in, cx<T>* out, int nfields, long long field_stride, long long member_stride,
                         long long xs_kx, long long xs_z, int members, bool inverse, const cx<T>* tw, int variant,
                         hipStream_t s) {
        if constexpr (kZB) {
            if (variant == 2)
                plines_b_l<L2>(in, out, nfields, field_stride, member_stride, xs_kx, xs_z, members, inverse, tw, s);
            else if (variant == 1)
                plines_b_l<L1>(in, out, nfields, field_stride, member_stride, xs_kx, xs_z, members, inverse, tw, s);
            else
                plines_b_l<L0>(in, out, nfields, field_stride, member_stride, xs_kx, xs_z, members, inverse, tw, s);
        }
    }
    static void px_products(const PairXArgs<T>& a, int ny, int nz, int members, int variant, bool xt,
                            hipStream_t s) {
        if constexpr (kXT) {
            if (xt) {
                if (variant == 1)
                    px_products_rd<X1, true>(a, ny, nz, members, s);
                else if (kAlt && variant == kXOrderVariant)
                    px_products_rd<XA, true>(a, ny, nz, members, s);
                else
                    px_products_rd<X0, true>(a, ny, nz, members, s);
                return;
            }
        }
        // the register-direct form addresses rows of at most N * N modes
        if (static_cast<long long>(ny) * nz > static_cast<long long>(N) * N) variant = 2;
        if (variant == kXLdsVariant) {
            const dim3 grid((ny + L0 - 1) / L0, nz, members);
            k_px_products<T, N, L0><<<grid, TL::NT, 0, s>>>(a);
        } else if (variant == 1) {
            px_products_rd<X1>(a, ny, nz, members, s);
        } else if (kAlt && variant == kXOrderVariant) {
            px_products_rd<XA>(a, ny, nz, members, s);
        } else {
            px_products_rd<X0>(a, ny, nz, members, s);
        }
    }
    template <int L>
    static void pfwd_z_rk_l(const PairFwdZArgs<T>& a, const PairInvZArgs<T>& inv, int nky, int nkx, int members,
                            hipStream_t s) {
        using ZV = ZTile<T, N, L>;
        const dim3 grid((nky + ZV::NL / 2 - 1) / (ZV::NL / 2), nkx, members);
        k_pfwd_z_rk_inv<T, N, L><<<grid, ZV::NT, 0, s>>>(a, inv);
    }
    static void pfwd_z_rk(const PairFwdZArgs<T>& a, const PairInvZArgs<T>& inv, int nky, int nkx, int members,
                          int variant, hipStream_t s) {
        if (variant == 1)
            pfwd_z_rk_l<Z1>(a, inv, nky, nkx, members, s);
        else if (kAlt && variant == kZOrderVariant)
            pfwd_z_rk_l<ZA>(a, inv, nky, nkx, members, s);
        else
            pfwd_z_rk_l<Z0>(a, inv, nky, nkx, members, s);
    }
    static constexpr PassTable<T> table{N,
                                        TL::NL,
                                        TL::NT,
                                        {3, kAlt ? 4 : 3, 3, kAlt ? 3 : (Z1 != Z0 ? 2 : 1)},
                                        {{L0, L1, L2, L2}, {X0, X1, L0, XA_LINES}, {L0, L1, L2, L2},
                                         {Z0, Z1, ZA_LINES, ZA_LINES}},
                                        {{0, 0, 0, 0}, {0, 0, 0, kAlt ? 1 : 0}, {0, 0, 0, 0},
                                         {0, 0, kAlt ? 1 : 0, kAlt ? 1 : 0}},
                                        &inv_z,
                                        &lines,
                                        &inv_x,
                                        &fwd_x,
                                        &fwd_z,
                                        &pinv_z,
                                        &plines,
                                        &pinv_x,
                                        &px_products,
                                        &px_products_vmax<T, N>,
                                        &pfwd_z_rk,
                                        kXT,
                                        {T0L, T1L, T2L, T2L},
                                        &plines_t,
                                        kZB,
                                        zblock_ky<T, N>(),
                                        &plines_b};
};

}  // namespace sph

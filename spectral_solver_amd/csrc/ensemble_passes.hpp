// One launch per pass for a whole ensemble of uneven grid sizes.
//
// An ensemble's members of equal size form a group; the per-group pass
// kernels (pair_passes.hpp) take a group per launch. Here one launch runs a
// pass for every group whose size is in kEnsSizes: its work-groups are all
// one wavefront (64 threads), each size using the tile width that makes its
// tile exactly that, and each work-group decodes its (group, tile) from its
// block index and runs that size's tile code:
//   z and x tiles   ZTile<T, N, 64 E / N>   (16: 16 lines, 32: 16, 48: 16,
//                                            64: 8, 96: 8, 128: 4)
//   y tiles         Tile<T, N, L> of 64 threads (L = 32, 16, 8, 8, 4, 4).
// A stage is then four launches however many sizes the ensemble holds, and
// the tiles of small grids fill the CUs beside those of large ones. The
// per-group arguments live in a device table (read with uniform loads);
// group boundaries and grids are kernel arguments.
#pragma once

#include <hip/hip_cooperative_groups.h>

#include "pair_passes.hpp"

namespace sph {

constexpr int kEnsThreads = 64;
constexpr int kEnsGroups = 6;  // at most one group per size below
#define SPH_ENS_SIZES(X) X(16) X(32) X(48) X(64) X(96) X(128)
constexpr int kEnsMaxN = 128;

constexpr bool ens_size(int n) {
    return n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128;
}

template <typename T, int N>
struct EnsTiles {
    static constexpr int ZL = kEnsThreads * ZTile<T, N>::E / N;  // z and x tile lines
    static constexpr int YL = 512 / N >= 32 ? 32 : (512 / N >= 16 ? 16 : (512 / N >= 8 ? 8 : 4));  // y tile lines
    static_assert(ZTile<T, N, ZL>::NT == kEnsThreads, "one-wavefront z tile");
    static_assert(Tile<T, N, YL>::NT == kEnsThreads, "one-wavefront y tile");
    static_assert(N % ZL == 0, "x tiles cover the y lines exactly");
    static constexpr bool SWZ = plines_swizzled<T, N>(YL);
    // LDS of the fused forward-z pass (the stash overlays the tiles), the x
    // pass (ping-pong where the x kernel uses it), the y pass
    static constexpr int LDS_Z = z_lds<T, N, ZL>();
    static constexpr int LDS_Y = tile_lds<YL, SWZ>(N);
    static constexpr int LDS = LDS_Z > LDS_Y ? LDS_Z : LDS_Y;
};

template <typename T>
constexpr int ens_lds() {
    int m = 0;
#define SPH_ENS_LDS(N) m = EnsTiles<T, N>::LDS > m ? EnsTiles<T, N>::LDS : m;
    SPH_ENS_SIZES(SPH_ENS_LDS)
#undef SPH_ENS_LDS
    return m;
}

// Everything a group's tiles need, for every pass (member 0; members add
// their strides). fwd[s] is the forward-z argument of RK stage s.
template <typename T>
struct EnsGroupArgs {
    int n;
    const cx<T>* tw;
    PairInvZArgs<T> inv;
    PairFwdZArgs<T> fwd[3];
    PairXArgs<T> x;
    cx<T>* lines;                 // pair fields of member 0 (y passes)
    long long npairlines[2];      // inverse y (3 fields), forward y (6 fields)
    long long member_stride;      // complex values between members' work buffers
    // ky-blocked layout (Group::use_layout(2)): the y passes move pairs
    // between lines (ky-blocked) and xlines (the x pass's fields, kx lines
    // xs_kx and z lines xs_z complex values apart), fields field_stride apart
    int blocked;
    cx<T>* xlines;
    long long xs_kx, xs_z, field_stride;
};

// The launch: group q owns logical blocks first[q] .. first[q+1]-1, laid out
// as its own grid (gx, gy, members) x fastest.
struct EnsGrid {
    int groups;
    unsigned total;  // blocks of the launch (first[groups])
    unsigned first[kEnsGroups + 1];
    unsigned gx[kEnsGroups];
    unsigned gy[kEnsGroups];
};

// Logical block of this work-group. Work-groups are dealt to the 8 XCDs
// round-robin; runs of kEnsRun consecutive logical tiles go to one XCD
// (neighbouring tiles share cache lines of the state rows), the runs
// themselves round-robin (each XCD gets tiles of every group).
constexpr unsigned kEnsRun = 8;
__device__ __forceinline__ unsigned ens_block(unsigned total) {
    constexpr unsigned kXcds = 8, kChunk = kXcds * kEnsRun;
    const unsigned p = blockIdx.x;
    const unsigned full = total / kChunk * kChunk;
    if (p >= full) return p;
    const unsigned xcd = p % kXcds, i = p / kXcds;  // i-th block of this XCD
    return (i / kEnsRun) * kChunk + xcd * kEnsRun + i % kEnsRun;
}

template <typename T, int N, int PASS>
__device__ __forceinline__ void ens_tile(const EnsGroupArgs<T>& a, TileId t, int stage, cx<T>* lds, cx<T>* twl,
                                         int tid) {
    constexpr int ZL = EnsTiles<T, N>::ZL, YL = EnsTiles<T, N>::YL;
    load_twiddles<T, N, kEnsThreads>(twl, a.tw, tid);
    if constexpr (PASS == 0) {
        pinv_z_tile<T, N, PZ_RHS, ZL>(a.inv, t, lds, twl, tid);
    } else if constexpr (PASS == 1 || PASS == 3) {
        constexpr bool INV = PASS == 1;
        if (a.blocked) {
            // the ky-blocked layout exists from n = 64 on (pass_table.hpp)
            if constexpr (N >= 64) {
                static_assert(N % (YL / 2) == 0 && N % zblock_ky<T, N>() == 0, "blocked y tiles divide the lines");
                // tile t.x of the group's pair lines, field by field; t.y the member
                const long long per_field = static_cast<long long>(N / 2 + 1) * N / (YL / 2);
                const int field = static_cast<int>(t.x / per_field);
                const long long g0 = (t.x % per_field) * (YL / 2);
                const long long fo = field * a.field_stride + static_cast<long long>(t.y) * a.member_stride;
                const cx<T>* src = (INV ? a.lines : a.xlines) + fo;
                cx<T>* dst = (INV ? a.xlines : a.lines) + fo;
                plines_b_tile<T, N, INV, YL>(src, dst, a.xs_kx, a.xs_z, static_cast<int>(g0 / N),
                                             static_cast<int>(g0 % N), lds, twl, tid);
            }
        } else {
            plines_tile<T, N, INV, YL>(a.lines, a.npairlines[INV ? 0 : 1], a.member_stride, t, lds, twl, tid);
        }
    } else if constexpr (PASS == 2) {
        px_products_rd_tile<T, N, ZL>(a.x, t, lds, twl, tid);
    } else {
        constexpr bool kStash = z_stash_fits<T, N, ZL>();
        pfwd_z_rk_tile<T, N, ZL>(a.fwd[stage], t, lds, twl, tid, kStash ? lds : nullptr);
        __syncthreads();
        pinv_z_tile<T, N, PZ_RHS, ZL>(a.inv, t, lds, twl, tid, kStash ? lds : nullptr);
    }
}

// PASS: 0 inverse z, 1 inverse y, 2 x (inverse x + products + forward x),
// 3 forward y, 4 forward z + update + the next stage's inverse z. Two waves
// per SIMD at least (a 256-register budget: the 96^3 x tile needs 189, and
// held to 128 the forward-z tiles spill).
// Logical block b of a launch over `grid` -> (group q, its tile t), and that
// tile's pass PASS.
template <typename T, int PASS>
__device__ __forceinline__ void ens_run_block(const EnsGroupArgs<T>* __restrict__ table, const EnsGrid& grid,
                                              unsigned b, int stage, cx<T>* lds, cx<T>* twl, int tid) {
    int q = 0;
    unsigned first = grid.first[0], gx = grid.gx[0], gy = grid.gy[0];
#pragma unroll
    for (int i = 1; i < kEnsGroups; ++i)
        if (i < grid.groups && b >= grid.first[i]) {
            q = i;
            first = grid.first[i];
            gx = grid.gx[i];
            gy = grid.gy[i];
        }
    const unsigned l = b - first;
    const TileId t{l % gx, (l / gx) % gy, l / (gx * gy)};
    const EnsGroupArgs<T>& a = table[q];
    switch (a.n) {
#define SPH_ENS_CASE(N)                                  \
    case N:                                              \
        ens_tile<T, N, PASS>(a, t, stage, lds, twl, tid); \
        break;
        SPH_ENS_SIZES(SPH_ENS_CASE)
#undef SPH_ENS_CASE
        default:
            break;
    }
}

template <typename T, int PASS>
__global__ void __launch_bounds__(kEnsThreads, 2)
    k_ens_pass(const EnsGroupArgs<T>* __restrict__ table, EnsGrid grid, int stage) {
    __shared__ cx<T> lds[ens_lds<T>()];
    __shared__ cx<T> twl[kEnsMaxN];
    ens_run_block<T, PASS>(table, grid, ens_block(grid.total), stage, lds, twl, threadIdx.x);
}

// One launch per RK stage: passes 1-4 of every fused group in one
// cooperative launch of resident work-groups, each walking the tiles of a
// pass and meeting the others at a grid-wide barrier before the next pass
// (cooperative_groups' grid sync: the launch is refused, not left waiting,
// when the grid does not fit the device at once). Work-group w runs on XCD
// w % 8 (the dispatcher's round robin) and takes tiles of that XCD's
// contiguous eighth of each pass, as ens_block arranges for the per-pass
// launches. gridDim.x is a multiple of 8.
struct EnsStageGrids {
    EnsGrid pass[4];  // passes 1..4
};
template <typename T, int PASS>
__device__ __forceinline__ void ens_stage_pass(const EnsGroupArgs<T>* __restrict__ table, const EnsGrid& grid,
                                               int stage, cx<T>* lds, cx<T>* twl, int tid) {
    constexpr unsigned kXcds = 8;
    const unsigned per_xcd = (grid.total + kXcds - 1) / kXcds;
    const unsigned xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds, wn = gridDim.x / kXcds;
    const unsigned lo = xcd * per_xcd;
    const unsigned hi = lo + per_xcd < grid.total ? lo + per_xcd : grid.total;
    for (unsigned b = lo + wi; b < hi; b += wn) {
        ens_run_block<T, PASS>(table, grid, b, stage, lds, twl, tid);
        __syncthreads();  // the next tile rewrites the LDS tiles and twiddles
    }
}
// (two waves per SIMD like the per-pass kernels: the four passes' tile code
// in one kernel then spills 476 bytes per lane; held to one wave per SIMD it
// fits in 256 + 138 registers and runs slower, 12.7 against 12.1 ms per
// 64-member step)
template <typename T>
__global__ void __launch_bounds__(kEnsThreads, 2)
    k_ens_stage(const EnsGroupArgs<T>* __restrict__ table, EnsStageGrids grids, int stage) {
    __shared__ cx<T> lds[ens_lds<T>()];
    __shared__ cx<T> twl[kEnsMaxN];
    const int tid = threadIdx.x;
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    ens_stage_pass<T, 1>(table, grids.pass[0], stage, lds, twl, tid);
    grid.sync();
    ens_stage_pass<T, 2>(table, grids.pass[1], stage, lds, twl, tid);
    grid.sync();
    ens_stage_pass<T, 3>(table, grids.pass[2], stage, lds, twl, tid);
    grid.sync();
    ens_stage_pass<T, 4>(table, grids.pass[3], stage, lds, twl, tid);
}

// Host: the launch grid of pass `pass` for groups of sizes n[q] with
// members[q] members each (tiles as above).
template <typename T>
EnsGrid ens_grid(int pass, int groups, const int* n, const int* members) {
    EnsGrid g{};
    g.groups = groups;
    unsigned first = 0;
    for (int q = 0; q < groups; ++q) {
        const int N = n[q], H = N / 2 + 1;
        int zl = 0, yl = 0;
        switch (N) {
#define SPH_ENS_DIMS(M)             \
    case M:                         \
        zl = EnsTiles<T, M>::ZL;    \
        yl = EnsTiles<T, M>::YL;    \
        break;
            SPH_ENS_SIZES(SPH_ENS_DIMS)
#undef SPH_ENS_DIMS
            default:
                throw std::runtime_error("size " + std::to_string(N) + " is not fused");
        }
        unsigned gx = 0, gy = 0, gz = static_cast<unsigned>(members[q]);
        if (pass == 0 || pass == 4) {  // z tiles: ky blocks x kx planes
            gx = (N + zl / 2 - 1) / (zl / 2);
            gy = H;
        } else if (pass == 2) {  // x tiles: y blocks x z planes
            gx = N / zl;
            gy = N;
        } else {  // y tiles: pair-line blocks (3 or 6 fields), members as the second dimension
            const long long lines = (pass == 1 ? 3LL : 6LL) * H * N;
            gx = static_cast<unsigned>((lines + yl / 2 - 1) / (yl / 2));
            gy = gz;
            gz = 1;
        }
        g.first[q] = first;
        g.gx[q] = gx;
        g.gy[q] = gy;
        first += gx * gy * gz;
    }
    g.first[groups] = first;
    g.total = first;
    return g;
}

// Launches pass `pass` (stage `stage` for pass 4) over the table's groups.
template <typename T>
void ens_launch(int pass, const EnsGroupArgs<T>* table, const EnsGrid& grid, int stage, hipStream_t s);

// Resident work-groups of k_ens_stage on this device (a multiple of 8; 0 if
// cooperative launches are not supported), and its launch for RK stage
// `stage` over passes 1-4 with that many work-groups.
template <typename T>
int ens_stage_blocks();
template <typename T>
void ens_launch_stage(const EnsGroupArgs<T>* table, const EnsStageGrids& grids, int blocks, int stage, hipStream_t s);

}  // namespace sph

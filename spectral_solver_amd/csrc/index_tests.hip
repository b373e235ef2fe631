// Host checks of the index maps the pass kernels rely on: every LDS layout
// and every permutation of butterflies or slots must be a bijection, or a
// kernel would silently overwrite or skip values. Built by build.py, run by
// tests/test_index_maps.py (no GPU needed).
#include <cmath>
#include <cstdio>
#include <vector>

#include "pass_table.hpp"

namespace {

int failures = 0;

void expect(bool ok, const char* what, int a, int b) {
    if (!ok) {
        ++failures;
        std::printf("FAIL %s (%d, %d)\n", what, a, b);
    }
}

// tile_at<NL, true>: (row, line) -> slot is a bijection onto [0, N*NL)
template <int NL>
void swizzled_tile(int n) {
    std::vector<int> seen(static_cast<std::size_t>(n) * NL, 0);
    for (int row = 0; row < n; ++row)
        for (int line = 0; line < NL; ++line) {
            const int a = sph::tile_at<NL, true>(row, line);
            expect(a >= 0 && a < n * NL, "tile_at range", row, line);
            if (a >= 0 && a < n * NL) ++seen[a];
        }
    for (int a = 0; a < n * NL; ++a) expect(seen[a] == 1, "tile_at bijection", n, a);
}

// z_stash_at<N, NL>: (component, element, ky) -> slot is a bijection onto
// [0, 3*N*NL/2) and fits in the kernel's LDS when the kernel uses it
template <typename T, int N>
void stash() {
    constexpr int NL = sph::ZTile<T, N>::NL;
    std::vector<int> seen(3 * N * (NL / 2), 0);
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < N; ++e)
            for (int k = 0; k < NL / 2; ++k) {
                const int a = sph::z_stash_at<N, NL>(c, e, k);
                expect(a >= 0 && a < 3 * N * (NL / 2), "stash range", e, k);
                if (a >= 0 && a < 3 * N * (NL / 2)) ++seen[a];
            }
    for (int a = 0; a < 3 * N * (NL / 2); ++a) expect(seen[a] == 1, "stash bijection", N, a);
    if (sph::z_stash_fits<T, N>()) expect(3 * N * (NL / 2) <= sph::z_lds<T, N>(), "stash fits", N, NL);
    expect(sph::z_lds<T, N>() * static_cast<int>(sizeof(sph::cx<T>)) + N * static_cast<int>(sizeof(sph::cx<T>)) <=
               sph::kLdsBytes,
           "z tile LDS", N, NL);
}

// ZTile::j_mirror: a bijection of the first-stage butterflies in which the
// groups of j and STEP - j are neighbours (p, p + 1 with p even) in one
// wavefront, and element e of one is element N - e of the other
template <typename T, int N, int LINES>
void mirror() {
    using Z = sph::ZTile<T, N, LINES>;
    if constexpr (Z::kMirrorable) {
        std::vector<int> seen(Z::STEP, 0);
        for (int p = 0; p < Z::STEP; ++p) {
            const int j = Z::j_mirror(p * Z::NL);
            expect(j >= 0 && j < Z::STEP, "mirror range", N, p);
            if (j >= 0 && j < Z::STEP) ++seen[j];
            if (p >= 2 && p % 2 == 0) {
                const int jn = Z::j_mirror((p + 1) * Z::NL);
                expect(j + jn == Z::STEP, "mirror partner", N, p);
                expect((p * Z::NL) / 64 == ((p + 1) * Z::NL) / 64, "mirror same wavefront", N, p);
                for (int r = 0; r < Z::R0; ++r)
                    expect(Z::e_first_mirror(p * Z::NL, r) + Z::e_first_mirror((p + 1) * Z::NL, Z::R0 - 1 - r) == N,
                           "mirror elements", N, r);
            }
        }
        for (int j = 0; j < Z::STEP; ++j) expect(seen[j] == 1, "mirror bijection", N, j);
    }
}

// ZTile::e_last_mirror: for lane group p, the partner of output slot r
// (element N - e) is slot RL-1-r of group p ^ 1 (p >= 2), own slot
// (RL - r) % RL (p = 0) or own slot RL-1-r (p = 1); partners share a
// wavefront; the rows k = j + r*STEP, r < RL/2, plus N/2 from group 0 cover
// 0 .. N/2 once (the x pass's lane-exchange pair separation)
template <typename T, int N, int LINES>
void mirror_last() {
    using Z = sph::ZTile<T, N, LINES>;
    if constexpr (Z::kMirrorableLast) {
        constexpr int RL = Z::RL, NL = Z::NL;
        std::vector<int> rows(N / 2 + 1, 0);
        for (int p = 0; p < Z::STEP; ++p) {
            const int tid = p * NL;
            for (int r = 0; r < RL; ++r) {
                const int e = Z::e_last_mirror(tid, r);
                int partner;
                if (p == 0)
                    partner = Z::e_last_mirror(tid, (RL - r) % RL);
                else if (p == 1)
                    partner = Z::e_last_mirror(tid, RL - 1 - r);
                else
                    partner = Z::e_last_mirror((p ^ 1) * NL, RL - 1 - r);
                expect((e + partner) % N == 0, "last mirror partner", N, p * RL + r);
                if (p >= 2) expect(tid / 64 == ((p ^ 1) * NL) / 64, "last mirror same wavefront", N, p);
                if (r < RL / 2 || (p == 0 && r == RL / 2)) {
                    expect(e >= 0 && e <= N / 2, "last mirror row range", N, e);
                    if (e >= 0 && e <= N / 2) ++rows[e];
                }
            }
        }
        for (int k = 0; k <= N / 2; ++k) expect(rows[k] == 1, "last mirror rows", N, k);
    }
}

// the first-stage write order r ^ s for every lane-group value s <= SWM is a
// permutation of the R0 slots (zfft, fft_stage)
void xor_orders() {
    for (int r0 : {2, 4, 8})
        for (int s = 0; s < r0; ++s) {
            std::vector<int> seen(r0, 0);
            for (int r = 0; r < r0; ++r) ++seen[r ^ s];
            for (int r = 0; r < r0; ++r) expect(seen[r] == 1, "xor order", r0, s);
        }
}

// the register DFTs against the defining sum (long double twiddles) on
// seeded inputs, both directions
template <int R>
void register_dft() {
    unsigned long long state = 12345 + R;
    auto rnd = [&]() {
        state = state * 6364136223846793005ULL + 1442695040888963407ULL;
        return static_cast<double>(state >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    };
    for (int inv = 0; inv < 2; ++inv) {
        sph::cx<double> v[R], x[R];
        for (int i = 0; i < R; ++i) x[i] = v[i] = {rnd(), rnd()};
        if (inv)
            sph::dft<R, true>(v);
        else
            sph::dft<R, false>(v);
        double worst = 0.0;
        for (int k = 0; k < R; ++k) {
            long double re = 0, im = 0;
            for (int j = 0; j < R; ++j) {
                const long double a = (inv ? 2.0L : -2.0L) * 3.14159265358979323846264338327950288L * j * k / R;
                re += x[j].x * cosl(a) - x[j].y * sinl(a);
                im += x[j].x * sinl(a) + x[j].y * cosl(a);
            }
            const double e = std::fabs(static_cast<double>(re) - v[k].x) + std::fabs(static_cast<double>(im) - v[k].y);
            worst = e > worst ? e : worst;
        }
        expect(worst < 1e-14, "register dft against the defining sum", R, inv);
    }
}

// The ky-blocked spectral layout: spectral_mode_index is a bijection of
// the n/2+1 kx planes of an n^3 grid for every ky block kb that divides n (kb
// = n: the plain (kx*n + kz)*n + ky), spectral_mode_of inverts it, and the z
// tiles' ZRow addresses the same modes.
void blocked_modes(int n, int kb) {
    const int H = n / 2 + 1;
    const long long total = static_cast<long long>(H) * n * n;
    std::vector<int> seen(static_cast<std::size_t>(total), 0);
    for (int kx = 0; kx < H; ++kx)
        for (int kz = 0; kz < n; ++kz)
            for (int ky = 0; ky < n; ++ky) {
                const long long i = sph::spectral_mode_index(kx, kz, ky, n, kb);
                expect(i >= 0 && i < total, "blocked mode range", n, kb);
                if (i < 0 || i >= total) continue;
                ++seen[static_cast<std::size_t>(i)];
                int x, z, y;
                sph::spectral_mode_of(i, n, kb, x, z, y);
                expect(x == kx && z == kz && y == ky, "blocked mode inverse", n, kb);
                const sph::ZRow zr(ky, kb, n);
                expect(static_cast<long long>(kx) * n * n + zr.base + static_cast<long long>(kz) * zr.step == i,
                       "z-tile row of the blocked layout", n, kb);
                if (kb == n)
                    expect(i == (static_cast<long long>(kx) * n + kz) * n + ky, "plain layout", n, kz);
            }
    for (long long i = 0; i < total; ++i) expect(seen[static_cast<std::size_t>(i)] == 1, "blocked mode bijection", n, kb);
}

// k_plines_b's blocked side: the q-th value of a tile of NL lines (NL/2 pair
// lines of consecutive kz) is element ky of line `line`, each exactly once,
// at the offset of (ky, pair line line/2, member line&1) in the layout
// relative to the tile's first value; consecutive q are consecutive offsets
// within runs of NL*KB values (coalesced).
template <int N, int NL, int KB>
void plines_b_map(int z0) {
    std::vector<int> seen(static_cast<std::size_t>(N) * NL, 0);
    const long long base = sph::spectral_mode_index(0, z0, 0, N, KB) * 2;
    long long prev = -1;
    for (int q = 0; q < N * NL; ++q) {
        int ky = -1, line = -1;
        const long long o = sph::plines_b_slot<N, NL, KB>(q, ky, line);
        expect(ky >= 0 && ky < N && line >= 0 && line < NL, "plines_b slot range", q, N);
        if (ky < 0 || ky >= N || line < 0 || line >= NL) continue;
        ++seen[static_cast<std::size_t>(ky) * NL + line];
        const long long want = sph::spectral_mode_index(0, z0 + line / 2, ky, N, KB) * 2 + (line & 1) - base;
        expect(o == want, "plines_b slot offset", q, N);
        if (q % (NL * KB) != 0) expect(o == prev + 1, "plines_b run contiguous", q, N);
        prev = o;
    }
    for (int a = 0; a < N * NL; ++a) expect(seen[static_cast<std::size_t>(a)] == 1, "plines_b slot bijection", N, a);
}

}  // namespace

int main() {
    for (int n : {64, 96, 128, 256}) {
        blocked_modes(n, n);
        for (int kb : {1, 2, 4, 8, 16})
            if (n % kb == 0) blocked_modes(n, kb);
    }
    // the blocked y tiles the library instantiates: KB = the default z
    // tile's NL/2, NL = 16, 8, 4, and a tile in the middle of the kz range
    plines_b_map<512, 16, sph::zblock_ky<float, 512>()>(8);
    plines_b_map<512, 8, sph::zblock_ky<float, 512>()>(500);
    plines_b_map<512, 4, sph::zblock_ky<float, 512>()>(2);
    plines_b_map<1024, 8, sph::zblock_ky<float, 1024>()>(1016);
    plines_b_map<1024, 4, sph::zblock_ky<float, 1024>()>(4);
    plines_b_map<96, 4, sph::zblock_ky<float, 96>()>(90);
    plines_b_map<96, 16, sph::zblock_ky<float, 96>()>(8);
    plines_b_map<512, 8, sph::zblock_ky<double, 512>()>(16);
    plines_b_map<64, 32, sph::zblock_ky<float, 64>()>(16);
    for (int n : {8, 12, 16, 24, 32, 48, 64, 96, 128, 256, 512, 1024}) {
        swizzled_tile<2>(n);
        swizzled_tile<4>(n);
        swizzled_tile<8>(n);
    }
    stash<float, 256>();
    stash<float, 512>();
    stash<float, 1024>();
    stash<double, 256>();
    stash<double, 512>();
    stash<double, 1024>();
    mirror<float, 32, 0>();
    mirror<float, 64, 0>();
    mirror<float, 256, 0>();
    mirror<float, 512, 0>();
    mirror<float, 512, 8>();
    mirror<float, 1024, 0>();
    mirror<double, 512, 0>();
    mirror<double, 256, 0>();
    mirror_last<float, 32, 0>();
    mirror_last<float, 64, 0>();
    mirror_last<float, 128, 0>();
    mirror_last<float, 256, 0>();
    mirror_last<float, 512, 0>();
    mirror_last<float, 512, 8>();
    mirror_last<float, 1024, 0>();
    mirror_last<float, 1024, 4>();
    mirror_last<double, 1024, 0>();
    mirror_last<double, 64, 0>();
    mirror_last<double, 128, 0>();
    mirror_last<double, 256, 0>();
    mirror_last<double, 512, 0>();
    xor_orders();
    register_dft<2>();
    register_dft<3>();
    register_dft<4>();
    register_dft<5>();
    register_dft<6>();
    register_dft<8>();
    register_dft<16>();
    if (failures) {
        std::printf("%d index-map checks failed\n", failures);
        return 1;
    }
    std::printf("all index-map checks passed\n");
    return 0;
}

// Device building blocks of the batched 1D FFT used by every pass kernel.
//
// A work-group owns a "tile" of NL lines of length N held in LDS with layout
// [element][line] (line fastest; rows padded to NL+1 complex values so that
// both the line-fast stage accesses and the element-fast global staging are
// free of systematic bank conflicts, or, for narrow FP32 tiles, unpadded and
// rotated: tile_at, ztile_stride). Each Stockham stage of radix R reads R values
// per butterfly task into registers, applies twiddles and a register DFT_R,
// then writes them back in self-sorting (Stockham) order:
//   x_r = a[j + r*N/R] * w_{Ns*R}^{r*(j mod Ns)},  y = DFT_R(x),
//   a'[(j - j mod Ns)*R + (j mod Ns) + r*Ns] = y_r.
// Butterfly tasks are numbered line-fast, so a wavefront touches consecutive
// LDS addresses in every stage. Twiddles come from a per-length table
// e^{-2*pi*i*m/N}, computed on the host in long double and rounded once to
// the storage type; inverse transforms use the conjugate.
#pragma once

#include <hip/hip_runtime.h>

namespace sph {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x;
    T y;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {s * a.x, s * a.y}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> czero() { return {T(0), T(0)}; }

// ---------------------------------------------------------------------------
// Radix plan: powers of two as radix-8 stages (a trailing 8*2 becomes 4*4),
// then 3s and 5s.
// ---------------------------------------------------------------------------
struct RadixPlan {
    int r[16];
    int count;
};

constexpr RadixPlan radix_plan(int n) {
    RadixPlan p{};
    int twos = 0;
    while (n % 2 == 0) {
        n /= 2;
        ++twos;
    }
    int eights = twos / 3;
    const int rem = twos % 3;
    int fours = 0, pairs = 0;
    if (rem == 1 && eights > 0) {
        eights -= 1;
        fours = 2;
    } else if (rem == 1) {
        pairs = 1;
    } else if (rem == 2) {
        fours = 1;
    }
    for (int i = 0; i < eights; ++i) p.r[p.count++] = 8;
    for (int i = 0; i < fours; ++i) p.r[p.count++] = 4;
    for (int i = 0; i < pairs; ++i) p.r[p.count++] = 2;
    while (n % 3 == 0) {
        n /= 3;
        p.r[p.count++] = 3;
    }
    while (n % 5 == 0) {
        n /= 5;
        p.r[p.count++] = 5;
    }
    if (n != 1) p.count = -1;  // unsupported factor
    return p;
}

constexpr bool length_supported(int n) { return n >= 2 && radix_plan(n).count > 0; }

// Plan of the register-direct z passes, where every thread holds lcm(radices)
// elements per field: a radix-8 stage next to a radix-3 one would make that 24,
// so lengths with a factor 3 take radix-4 stages and, for an odd power of two,
// one radix-6 stage (2*3) in place of a 2 and a 3: 12 elements per thread and
// the fewest stages (96 = 4*4*6, 24 = 4*6, 48 = 4*4*3, 12 = 4*3, 384 =
// 4*4*4*6). Lengths with a factor 5 likewise take radix-4 stages, a radix-2
// one for an odd power of two, then the 5s: 20 elements per thread (80 =
// 4*4*5, 160 = 4*4*2*5, 640 = 4*4*4*2*5) instead of 40 next to a radix 8.
// Powers of two with two or more radix-8 stages put one of them last (1024 =
// 8*4*4*8, 256 = 8*4*8): the same stages, and the last one then holds one
// task per thread (ZTile::kMirrorableLast).
//
// order = 1 is the alternative stage order the autotuner times beside it
// (z_radix_plan_alt).
constexpr RadixPlan z_radix_plan_alt(int n);
constexpr RadixPlan z_radix_plan(int n, int order = 0) {
    if (order == 1) return z_radix_plan_alt(n);
    const RadixPlan p = radix_plan(n);
    if (p.count < 0) return p;
    bool has3 = false, has5 = false;
    int eights = 0;
    for (int s = 0; s < p.count; ++s) {
        has3 |= p.r[s] == 3;
        has5 |= p.r[s] == 5;
        eights += p.r[s] == 8;
    }
    if (!has3 && !has5) {
        RadixPlan q = p;
        if (eights >= 2 && p.r[0] == 8 && p.r[1] == 8 && p.r[p.count - 1] != 8) {
            for (int s = 1; s < p.count - 1; ++s) q.r[s] = p.r[s + 1];
            q.r[p.count - 1] = 8;
        }
        return q;
    }
    int twos = 0, threes = 0, fives = 0;
    while (n % 2 == 0) {
        n /= 2;
        ++twos;
    }
    while (n % 3 == 0) {
        n /= 3;
        ++threes;
    }
    while (n % 5 == 0) {
        n /= 5;
        ++fives;
    }
    RadixPlan q{};
    for (int i = 0; i < twos / 2; ++i) q.r[q.count++] = 4;
    if (twos % 2) {
        if (threes > 0) {
            q.r[q.count++] = 6;
            --threes;
        } else {
            q.r[q.count++] = 2;
        }
    }
    for (int i = 0; i < threes; ++i) q.r[q.count++] = 3;
    for (int i = 0; i < fives; ++i) q.r[q.count++] = 5;
    return q;
}

// Alternative stage order of the z-type passes, timed by the autotuner
// against z_radix_plan(n):
//   1024: 8*16*8, three stages instead of four, with a radix-16 middle stage
//     (16 values per thread: the tile runs one 512-thread work-group per CU
//     with 256 registers);
//   other powers of two: the radix-8 stages last and radix-4 stages first
//     (256 = 4*8*8, 128 = 4*4*8, 32 = 4*8); a plan of three or more radix-8
//     stages only (512 = 8*8*8) instead takes its first two as 4*4*4 (512 =
//     4*4*4*8): one stage more, a radix-4 first stage;
//   lengths with a factor 3 or 5: the default plan's last stage first
//     (96 = 6*4*4, 48 = 3*4*4, 384 = 6*4*4*4, 80 = 5*4*4).
// Where this gives the default plan again (64 = 8*8, 16 = 4*4, 8), the
// length has no alternative (z_plan_has_alt).
constexpr RadixPlan z_radix_plan_alt(int n) {
    const RadixPlan p = z_radix_plan(n, 0);
    if (p.count < 2) return p;
    if (n == 1024) {
        // three stages with a radix-16 middle: 16 values per thread, twice
        // the registers of the default tile (ZTile::kMinWaves)
        RadixPlan q{};
        q.r[q.count++] = 8;
        q.r[q.count++] = 16;
        q.r[q.count++] = 8;
        return q;
    }
    bool pow2 = true;
    int eights = 0;
    for (int s = 0; s < p.count; ++s) {
        pow2 &= p.r[s] == 2 || p.r[s] == 4 || p.r[s] == 8;
        eights += p.r[s] == 8;
    }
    RadixPlan q{};
    if (pow2 && eights == p.count && eights >= 3) {
        q.r[q.count++] = 4;
        q.r[q.count++] = 4;
        q.r[q.count++] = 4;
        for (int s = 2; s < p.count; ++s) q.r[q.count++] = 8;
    } else if (pow2) {
        for (int s = 0; s < p.count; ++s)
            if (p.r[s] != 8) q.r[q.count++] = p.r[s];
        for (int s = 0; s < eights; ++s) q.r[q.count++] = 8;
    } else {
        q.r[q.count++] = p.r[p.count - 1];
        for (int s = 0; s < p.count - 1; ++s) q.r[q.count++] = p.r[s];
    }
    return q;
}

constexpr bool z_plan_has_alt(int n) {
    const RadixPlan a = z_radix_plan(n, 0), b = z_radix_plan(n, 1);
    if (a.count != b.count) return true;
    for (int s = 0; s < a.count; ++s)
        if (a.r[s] != b.r[s]) return true;
    return false;
}

// Template argument LINES of the z-type tiles and kernels: the tile width
// (0 = the default) plus kOrderUnit times the radix order (0: z_radix_plan,
// 1: z_radix_plan_alt).
constexpr int kOrderUnit = 1024;
constexpr int zlines(int lines, int order) { return lines + order * kOrderUnit; }

// ---------------------------------------------------------------------------
// Tile geometry per (storage type, length).
// ---------------------------------------------------------------------------
constexpr int round_up(int v, int m) { return (v + m - 1) / m * m; }
constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Default lines per tile: one 128-byte row segment of complex values per
// element. The pair passes also run narrower tiles (see PassTable).
template <typename T, int N>
constexpr int default_lines() {
    return sizeof(T) == 4 ? (N >= 1024 ? 8 : 16) : (N >= 1024 ? 4 : 8);
}

template <typename T, int N, int NL_ = default_lines<T, N>()>
struct Tile {
    static constexpr int NL = NL_;
    static_assert(NL >= 2 && NL % 2 == 0, "tiles hold an even number of lines");
    static constexpr int STRIDE = NL + 1;
    static constexpr int ELEMS = NL * N;
    static constexpr int NT = clampi(round_up(ELEMS / 8, 64), 64, 1024);
    static constexpr int E = (ELEMS + NT - 1) / NT;  // elements per thread
    static constexpr int LDS = STRIDE * N;           // complex values in the tile buffer
};

// ---------------------------------------------------------------------------
// Register DFTs. rot() multiplies by -i (forward) or +i (inverse).
// ---------------------------------------------------------------------------
template <bool INV, typename T>
__host__ __device__ __forceinline__ cx<T> rot(cx<T> a) {
    if constexpr (INV) {
        return {-a.y, a.x};
    } else {
        return {a.y, -a.x};
    }
}

template <int R, bool INV, typename T>
__host__ __device__ __forceinline__ void dft(cx<T>* v) {
    if constexpr (R == 2) {
        const cx<T> a = v[0];
        v[0] = a + v[1];
        v[1] = a - v[1];
    } else if constexpr (R == 4) {
        const cx<T> a = v[0] + v[2], b = v[0] - v[2];
        const cx<T> c = v[1] + v[3], d = rot<INV>(v[1] - v[3]);
        v[0] = a + c;
        v[2] = a - c;
        v[1] = b + d;
        v[3] = b - d;
    } else if constexpr (R == 8) {
        const T h = T(0.707106781186547524400844362104849039L);
        const cx<T> a0 = v[0] + v[4], a1 = v[0] - v[4];
        const cx<T> a2 = v[2] + v[6], a3 = rot<INV>(v[2] - v[6]);
        const cx<T> a4 = v[1] + v[5], a5 = v[1] - v[5];
        const cx<T> a6 = v[3] + v[7], a7 = rot<INV>(v[3] - v[7]);
        const cx<T> e0 = a0 + a2, e2 = a0 - a2, e1 = a1 + a3, e3 = a1 - a3;
        const cx<T> o0 = a4 + a6, o2 = a4 - a6, o1 = a5 + a7, o3 = a5 - a7;
        const cx<T> t1 = (o1 + rot<INV>(o1)) * h;
        const cx<T> t2 = rot<INV>(o2);
        const cx<T> t3 = (rot<INV>(o3) - o3) * h;
        v[0] = e0 + o0;
        v[4] = e0 - o0;
        v[1] = e1 + t1;
        v[5] = e1 - t1;
        v[2] = e2 + t2;
        v[6] = e2 - t2;
        v[3] = e3 + t3;
        v[7] = e3 - t3;
    } else if constexpr (R == 3) {
        const T s = T(0.866025403784438646763723170752936183L);
        const cx<T> t = v[1] + v[2];
        const cx<T> m = v[0] - t * T(0.5);
        const cx<T> d = rot<INV>((v[1] - v[2]) * s);
        v[0] = v[0] + t;
        v[1] = m + d;
        v[2] = m - d;
    } else if constexpr (R == 6) {
        // 6 = 2 * 3: radix-3 DFTs of the even and odd elements, then radix-2
        // combinations with the twiddles w6^k, w6 = e^{-+2 pi i/6}
        const T c = T(0.5);
        const T sn = T(0.866025403784438646763723170752936183L);  // sin(pi/3)
        cx<T> a3[3] = {v[0], v[2], v[4]};
        cx<T> b3[3] = {v[1], v[3], v[5]};
        dft<3, INV>(a3);
        dft<3, INV>(b3);
        // w6 = (1/2, -+sin(pi/3)), w6^2 = (-1/2, -+sin(pi/3))
        const cx<T> w1 = INV ? cx<T>{c, sn} : cx<T>{c, -sn};
        const cx<T> w2 = INV ? cx<T>{-c, sn} : cx<T>{-c, -sn};
        const cx<T> t1 = cmul(b3[1], w1), t2 = cmul(b3[2], w2);
        v[0] = a3[0] + b3[0];
        v[3] = a3[0] - b3[0];
        v[1] = a3[1] + t1;
        v[4] = a3[1] - t1;
        v[2] = a3[2] + t2;
        v[5] = a3[2] - t2;
    } else if constexpr (R == 16) {
        // 16 = 4 x 4: DFT4 over x[m + 4 q] for each m, twiddles w16^(m k1),
        // DFT4 over m: y[k1 + 4 k2] = sum_m w4^(m k2) w16^(m k1) DFT4_m[k1]
        const T c1 = T(0.923879532511286756128183189396788933L);  // cos(pi/8)
        const T s1 = T(0.382683432365089771728459984030398866L);  // sin(pi/8)
        const T h = T(0.707106781186547524400844362104849039L);
        cx<T> a[4][4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            cx<T> q[4] = {v[m], v[m + 4], v[m + 8], v[m + 12]};
            dft<4, INV>(q);
#pragma unroll
            for (int k = 0; k < 4; ++k) a[m][k] = q[k];
        }
        // w16^j = (cos(pi j/8), -+sin(pi j/8)) for j = m k1 in {1, 2, 3, 4, 6, 9}
        auto tw = [&](cx<T> x, T c, T sn) -> cx<T> {
            const cx<T> w = INV ? cx<T>{c, sn} : cx<T>{c, -sn};
            return cmul(x, w);
        };
        a[1][1] = tw(a[1][1], c1, s1);
        a[1][2] = tw(a[1][2], h, h);
        a[1][3] = tw(a[1][3], s1, c1);
        a[2][1] = tw(a[2][1], h, h);
        a[2][2] = rot<INV>(a[2][2]);
        a[2][3] = tw(a[2][3], -h, h);
        a[3][1] = tw(a[3][1], s1, c1);
        a[3][2] = tw(a[3][2], -h, h);
        a[3][3] = tw(a[3][3], -c1, -s1);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cx<T> q[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            dft<4, INV>(q);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = q[k2];
        }
    } else if constexpr (R == 5) {
        // y_k = sum_q x_q w^{qk}, w = e^{-+2 pi i/5}
        const T c1 = T(0.309016994374947424102293417182819059L);   // cos(2pi/5)
        const T c2 = T(-0.809016994374947424102293417182819059L);  // cos(4pi/5)
        const T s1 = T(0.951056516295153572116439333379382143L);   // sin(2pi/5)
        const T s2 = T(0.587785252292473129168705954639072769L);   // sin(4pi/5)
        const cx<T> p1 = v[1] + v[4], m1 = v[1] - v[4];
        const cx<T> p2 = v[2] + v[3], m2 = v[2] - v[3];
        const cx<T> a1 = v[0] + p1 * c1 + p2 * c2;
        const cx<T> a2 = v[0] + p1 * c2 + p2 * c1;
        const cx<T> b1 = rot<INV>(m1 * s1 + m2 * s2);
        const cx<T> b2 = rot<INV>(m1 * s2 - m2 * s1);
        v[0] = v[0] + p1 + p2;
        v[1] = a1 + b1;
        v[4] = a1 - b1;
        v[2] = a2 + b2;
        v[3] = a2 - b2;
    }
}

// ---------------------------------------------------------------------------
// Tile layout in LDS: element `row` of line `line`.
//   padded:   row * (NL + 1) + line
//   swizzled: row * NL + ((line + 2 * (row / (16 / NL))) mod NL), FP32 with
//             NL = 2, 4 or 8 lines only.
// The LDS serves 64-bit accesses 16 lanes at a time over 32 four-byte banks.
// An unpadded FP32 row covers 2*NL banks, so rows r and r + 16/NL share
// banks; the rotation moves the second onto other banks. That makes the
// contiguous global <-> LDS copies of a pair-line tile (16 lanes: 8
// consecutive rows x the 2 lines of a pair) conflict-free, and the stage
// reads (16 lanes: 16/NL consecutive rows x NL lines) stay so. The first
// stage's writes (rows j*R + r for consecutive j) additionally go out in the
// order r ^ (j mod 16/NL) (fft_stage).
// ---------------------------------------------------------------------------
template <int NL, bool SWZ>
__host__ __device__ __forceinline__ int tile_at(int row, int line) {
    if constexpr (SWZ) {
        static_assert(NL == 2 || NL == 4 || NL == 8, "swizzled rows hold 2, 4 or 8 lines");
        constexpr unsigned RB = 16 / NL;
        const unsigned r = static_cast<unsigned>(row), l = static_cast<unsigned>(line);
        return static_cast<int>(r * NL + ((l + 2 * (r / RB)) & (NL - 1)));
    } else {
        return row * (NL + 1) + line;
    }
}

template <int NL, bool SWZ>
constexpr int tile_lds(int n) {
    return SWZ ? NL * n : (NL + 1) * n;
}

// ---------------------------------------------------------------------------
// One in-place Stockham stage over the whole tile.
// ---------------------------------------------------------------------------
template <typename T, int N, int NL, int NT, int R, int NS, bool INV, bool SWZ = false>
__device__ __forceinline__ void fft_stage(cx<T>* buf, const cx<T>* twl, int tid) {
    constexpr int M = N / R;
    constexpr int TASKS = NL * M;
    constexpr int TPT = (TASKS + NT - 1) / NT;
    cx<T> v[TPT][R];
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int t = tid + i * NT;
        if (TASKS % NT == 0 || t < TASKS) {
            const int line = t % NL;
            const int j = t / NL;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = buf[tile_at<NL, SWZ>(j + r * M, line)];
            if constexpr (NS > 1) {
                const int k = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cx<T> w = twl[r * k * (N / (NS * R))];
                    if constexpr (INV) w.y = -w.y;
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            dft<R, INV>(v[i]);
        }
    }
    __syncthreads();
    // first stage of a swizzled tile: lane j writes element r ^ (j & SWM) in
    // write r (R a power of two)
    constexpr bool kPow2 = (R & (R - 1)) == 0;
    constexpr int SWM = SWZ && NS == 1 && kPow2 ? ((16 / NL) < R ? 16 / NL : R) - 1 : 0;
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int t = tid + i * NT;
        if (TASKS % NT == 0 || t < TASKS) {
            const int line = t % NL;
            const int j = t / NL;
            const int k = j % NS;
            const int base = (j - k) * R + k;
            if constexpr (SWM == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r) buf[tile_at<NL, SWZ>(base + r * NS, line)] = v[i][r];
            } else {
                const int sw = j & SWM;
#pragma unroll
                for (int b = 1; b <= SWM; b <<= 1) {
                    const bool take = (sw & b) != 0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (!(r & b)) {
                            const cx<T> a0 = v[i][r], a1 = v[i][r | b];
                            v[i][r] = take ? a1 : a0;
                            v[i][r | b] = take ? a0 : a1;
                        }
                }
#pragma unroll
                for (int r = 0; r < R; ++r) buf[tile_at<NL, SWZ>(base + (r ^ sw) * NS, line)] = v[i][r];
            }
        }
    }
    __syncthreads();
}

template <typename T, int N, int NL, int NT, bool INV, int S = 0, int NS = 1, bool SWZ = false>
__device__ __forceinline__ void fft_tile(cx<T>* buf, const cx<T>* twl, int tid) {
    constexpr RadixPlan P = radix_plan(N);
    if constexpr (S < P.count) {
        constexpr int R = P.r[S];
        fft_stage<T, N, NL, NT, R, NS, INV, SWZ>(buf, twl, tid);
        fft_tile<T, N, NL, NT, INV, S + 1, NS * R, SWZ>(buf, twl, tid);
    }
}

// ---------------------------------------------------------------------------
// Register-direct tile transform for strided (z) passes.
//
// NF fields of NL lines each. Thread tid owns line l = tid % NL of every field;
// in the first stage it holds elements j + r*N/R0 (j = tid/NL + i*NT/NL), which
// the caller loads straight from global memory (for fixed r a wavefront reads
// NL consecutive values of NT/NL/... rows: whole 128-byte segments). The first
// stage runs in registers, the middle stages go through LDS, and the last
// stage leaves element j + r*N/RL of line l in registers for the caller's
// post-operation and coalesced store.
//
// LDS rows (one element of every line) and bank conflicts. The LDS serves
// 64-bit accesses 16 lanes at a time over 32 four-byte banks. An FP32 row of
// NL <= 8 lines covers 2*NL banks, so 16/NL consecutive rows fill the banks
// once: stored unpadded, every access whose 16 lanes touch 16/NL consecutive
// rows is conflict-free (the middle and last stages, and the x pass's own
// accesses). The first-stage write is the exception: its lanes write rows
// j*R0 + r for the butterflies j of neighbouring lane groups, a multiple of R0
// apart, which share banks; there lane group p = tid/NL writes its values in
// the order r ^ (p mod 16/NL), so that a write instruction's rows land in
// distinct banks (also with the mirrored butterfly order, j_mirror). Wider
// FP32 rows and FP64 rows keep one element of padding per row.
// ---------------------------------------------------------------------------
template <typename T>
constexpr int ztile_stride(int nl) {
    return sizeof(T) == 4 && nl <= 8 ? nl : nl + 1;
}

template <typename T, int N, int LINES_ = 0>  // zlines(width or 0 for the default, radix order)
struct ZTile {
    static constexpr int ORD = LINES_ / kOrderUnit;
    static constexpr int NL_ = LINES_ % kOrderUnit;
    static constexpr RadixPlan P = z_radix_plan(N, ORD);
    static constexpr int lcm_radix() {
        int l = 1;
        for (int s = 0; s < P.count; ++s) {
            int a = l, b = P.r[s];
            while (b) {
                const int t = a % b;
                a = b;
                b = t;
            }
            l = l / a * P.r[s];
        }
        return l;
    }
    static constexpr int pick_nl() {
        // two field tiles must fit in LDS: FP64 at N = 1024 takes narrow tiles
        int nl = sizeof(T) == 4 ? (N >= 1024 ? 8 : 16) : (N >= 1024 ? 2 : 8);
        while (nl * N / lcm_radix() < 64) nl *= 2;
        return nl;
    }
    static constexpr int NL = NL_ ? NL_ : pick_nl();
    static constexpr int E = lcm_radix();  // elements per thread and field
    static constexpr int NT = NL * N / E;
    // minimum waves per SIMD of the kernels on this tile (their launch
    // bounds): 512-thread tiles ask for 4 (two work-groups per CU, 128
    // registers), tiles of 16 or more values per thread for 2 (one
    // work-group per CU, 256 registers), others for 1
    static constexpr int kMinWaves = E >= 16 ? 2 : (NT == 512 ? 4 : 1);
    static constexpr int S = ztile_stride<T>(NL);  // LDS row stride
    static constexpr int LDS = S * N;          // complex values per field tile
    static constexpr int R0 = P.r[0];
    static constexpr int RL = P.r[P.count - 1];
    static constexpr int T0 = E / R0;          // first-stage tasks per thread
    static constexpr int TLAST = E / RL;       // last-stage tasks per thread
    static constexpr int STEP = NT / NL;       // j step between a thread's tasks
    static_assert(NT <= 1024 && NT % 64 == 0 && NT % NL == 0, "bad z tile");
    static_assert(E % R0 == 0 && E % RL == 0, "bad z tile");
    // element index of first-stage slot (i, r) / last-stage slot (i, r)
    __host__ __device__ static int e_first(int tid, int i, int r) { return tid / NL + i * STEP + r * (N / R0); }
    __host__ __device__ static int e_last(int tid, int i, int r) { return tid / NL + i * STEP + r * (N / RL); }
    // First-stage butterflies in mirrored order (zfft's MIRROR): lane group
    // p = tid / NL takes butterfly j = 0, STEP/2, 1, STEP-1, 2, STEP-2, ...
    // for p = 0, 1, 2, 3, ..., so that the groups of j and STEP - j are
    // neighbours in one wavefront. With one task per thread and N/R0 = STEP,
    // element e = j + r*STEP of group j and element N - e of group STEP - j
    // (slot R0-1-r) are then loaded by the same wavefront: the x pass's
    // Hermitian extension reads each spectral row twice, and the second read
    // now comes from the same wavefront's recent access.
    static constexpr bool kMirrorable = T0 == 1 && N / R0 == STEP && STEP % 2 == 0 && 64 / NL >= 2;
    __host__ __device__ static int j_mirror(int tid) {
        const int p = tid / NL;
        const int m = p >> 1;
        return p < 2 ? p * (STEP / 2) : ((p & 1) ? STEP - m : m);
    }
    __host__ __device__ static int e_first_mirror(int tid, int r) { return j_mirror(tid) + r * (N / R0); }
    // Last-stage butterflies in the same order (zfft's MIRROR_LAST): with one
    // task per thread and N/RL = STEP, output element e = j + r*STEP of lane
    // group p >= 2 and element N - e (slot RL-1-r) of group p ^ 1 are one
    // lane exchange apart (lane ^ NL); groups 0 and 1 hold j = 0 and STEP/2,
    // whose elements N - e are their own slots (RL - r) % RL and RL-1-r.
    static constexpr bool kMirrorableLast = TLAST == 1 && N / RL == STEP && STEP % 2 == 0 && 64 / NL >= 2;
    __host__ __device__ static int e_last_mirror(int tid, int r) { return j_mirror(tid) + r * (N / RL); }
};

// Work-group barrier of the z transforms. LB: for LDS only -- this wave's
// LDS accesses complete, then the hardware barrier -- so that global-to-LDS
// loads of another buffer (a prefetch) stay in flight across it; a
// __syncthreads() would also wait for them (its release fence counts them).
template <bool LB>
__device__ __forceinline__ void zbarrier() {
    if constexpr (LB) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt left at their maxima
        __builtin_amdgcn_s_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    } else {
        __syncthreads();
    }
}

// One middle Stockham stage over NF field tiles. In place (src == dst): read
// all, sync, write all, sync. Ping-pong (src != dst): read src, write dst,
// sync; the reads of src and writes of dst need no barrier between them.
template <typename T, int N, int NL, int NT, int NF, int R, int NS, bool INV, bool INPLACE, bool LB = false>
__device__ __forceinline__ void zstage_mid(const cx<T>* src, cx<T>* dst, const cx<T>* twl, int tid) {
    constexpr int M = N / R;
    constexpr int TASKS = NL * M;
    static_assert(TASKS % NT == 0, "middle stage must split evenly");
    constexpr int TPT = TASKS / NT;
    constexpr int S = ztile_stride<T>(NL);
    cx<T> v[NF][TPT][R];
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int t = tid + i * NT;
        const int line = t % NL, j = t / NL, k = j % NS;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[f][i][r] = src[f * (S * N) + (j + r * M) * S + line];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cx<T> w = twl[r * k * (N / (NS * R))];
                if constexpr (INV) w.y = -w.y;
                v[f][i][r] = cmul(v[f][i][r], w);
            }
            dft<R, INV>(v[f][i]);
        }
    }
    if constexpr (INPLACE) zbarrier<LB>();
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int t = tid + i * NT;
        const int line = t % NL, j = t / NL, k = j % NS;
        const int base = (j - k) * R + k;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < R; ++r) dst[f * (S * N) + (base + r * NS) * S + line] = v[f][i][r];
    }
    zbarrier<LB>();
}

// stages S .. count-2 of the plan of ZTile<T, N, LINES>; NS is the product
// of the radices before stage S. The data starts in a; with PP it alternates
// between a and b. Returns the tile holding the result.
template <typename T, int N, int LINES, int NF, bool INV, int S, int NS, bool PP, bool LB = false>
__device__ __forceinline__ cx<T>* zstages_mid(cx<T>* a, cx<T>* b, const cx<T>* twl, int tid) {
    using Z = ZTile<T, N, LINES>;
    constexpr RadixPlan P = Z::P;
    if constexpr (S <= P.count - 2) {
        cx<T>* const dst = PP ? b : a;
        zstage_mid<T, N, Z::NL, Z::NT, NF, P.r[S], NS, INV, !PP, LB>(a, dst, twl, tid);
        return zstages_mid<T, N, LINES, NF, INV, S + 1, NS * P.r[S], PP, LB>(dst, a, twl, tid);
    } else {
        return a;
    }
}

// Middle stages of the z plan (ping-pong tiles alternate once per middle
// stage; the first-stage write goes to tile A).
template <typename T, int N, int LINES = 0>
constexpr int z_mid_stages() {
    constexpr int c = ZTile<T, N, LINES>::P.count;
    return c > 2 ? c - 2 : 0;
}

constexpr int kLdsBytes = 160 * 1024;  // LDS per work-group on gfx950

// First and middle stages of zfft for plans of two or more stages: v (the
// first-stage inputs) is consumed and the last stage's input tile is
// returned (its writes are complete: the barrier of the stage that wrote it
// has passed). Template arguments as zfft.
template <typename T, int N, int NF, bool INV, int LINES = 0, bool PP = false, bool MIRROR = false, bool LB = false,
          int A0, int B0>
__device__ __forceinline__ const cx<T>* zfft_head(cx<T> (&v)[NF][A0][B0], cx<T>* lds, const cx<T>* twl, int tid,
                                                  cx<T>* lds_b = nullptr) {
    using Z = ZTile<T, N, LINES>;
    static_assert(A0 == Z::T0 && B0 == Z::R0, "slot arrays must match the tile");
    static_assert(!MIRROR || Z::kMirrorable, "mirrored first stage needs one task per thread");
    static_assert(Z::P.count >= 2, "one-stage plans have no middle");
    constexpr int NL = Z::NL, NT = Z::NT, S = Z::S, R0 = Z::R0;
    const int l = tid % NL;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < Z::T0; ++i) dft<R0, INV>(v[f][i]);
    // previous readers of tile A are done (ping-pong with an odd number of
    // middle stages: the last reads of A were the middle stage's, before
    // the barrier that ended it)
    if constexpr (!PP || z_mid_stages<T, N, LINES>() % 2 == 0) zbarrier<LB>();
    // unpadded rows: lane group p = tid / NL writes element r ^ (p & SWM) in
    // write r
    constexpr bool kPow2 = (R0 & (R0 - 1)) == 0;
    constexpr int SWM = S == NL && kPow2 ? (16 / NL < R0 ? 16 / NL : R0) - 1 : 0;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < Z::T0; ++i) {
            const int j = MIRROR ? Z::j_mirror(tid) : tid / NL + i * Z::STEP;
            cx<T>* const row0 = lds + f * Z::LDS + j * R0 * S + l;
            if constexpr (SWM == 0) {
#pragma unroll
                for (int r = 0; r < R0; ++r) row0[r * S] = v[f][i][r];
            } else {
                const int sw = (tid / NL) & SWM;  // lane group parity (also in mirrored order)
                cx<T> u[R0];
#pragma unroll
                for (int r = 0; r < R0; ++r) u[r] = v[f][i][r];
                // u[r] = v[r ^ sw]: one conditional exchange per bit
#pragma unroll
                for (int b = 1; b <= SWM; b <<= 1) {
                    const bool take = (sw & b) != 0;
#pragma unroll
                    for (int r = 0; r < R0; ++r)
                        if (!(r & b)) {
                            const cx<T> a0 = u[r], a1 = u[r | b];
                            u[r] = take ? a1 : a0;
                            u[r | b] = take ? a0 : a1;
                        }
                }
#pragma unroll
                for (int r = 0; r < R0; ++r) row0[(r ^ sw) * S] = u[r];
            }
        }
    zbarrier<LB>();
    // stages 1 .. count-2 (NS after stage 0 is R0)
    return zstages_mid<T, N, LINES, NF, INV, 1, R0, PP, LB>(lds, lds_b ? lds_b : lds + NF * Z::LDS, twl, tid);
}

// v: first-stage inputs (consumed), w: last-stage outputs. LINES: tile width
// (0 = the default ZTile width). PP: lds holds two tiles (2 * NF * Z::LDS)
// and the middle stages alternate between them, which drops the barrier
// between a stage's reads and writes; with an odd number of middle stages the
// last stage reads tile B, so the next transform's first-stage write to tile
// A needs no barrier either. Between PP transforms a caller may not touch lds.
// MIRROR / MIRROR_LAST: first / last-stage butterflies in ZTile::j_mirror
// order (w then holds the elements e_last_mirror). LB: LDS-only barriers
// (zbarrier). lds_b: tile B of PP where it is not the one after tile A.
template <typename T, int N, int NF, bool INV, int LINES = 0, bool PP = false, bool MIRROR = false,
          bool MIRROR_LAST = false, bool LB = false, int A0, int B0, int A1, int B1>
__device__ __forceinline__ void zfft(cx<T> (&v)[NF][A0][B0], cx<T> (&w)[NF][A1][B1], cx<T>* lds, const cx<T>* twl,
                                     int tid, cx<T>* lds_b = nullptr) {
    using Z = ZTile<T, N, LINES>;
    static_assert(A0 == Z::T0 && B0 == Z::R0 && A1 == Z::TLAST && B1 == Z::RL, "slot arrays must match the tile");
    static_assert(!MIRROR_LAST || Z::kMirrorableLast, "mirrored last stage needs one task per thread");
    constexpr int NL = Z::NL, S = Z::S, R0 = Z::R0, RL = Z::RL;
    if constexpr (Z::P.count == 1) {
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int i = 0; i < Z::T0; ++i) {
                dft<R0, INV>(v[f][i]);
#pragma unroll
                for (int r = 0; r < R0; ++r) w[f][i][r] = v[f][i][r];
            }
    } else {
        const int l = tid % NL;
        const cx<T>* const last = zfft_head<T, N, NF, INV, LINES, PP, MIRROR, LB>(v, lds, twl, tid, lds_b);
        constexpr int NSL = N / RL;
#pragma unroll
        for (int i = 0; i < Z::TLAST; ++i) {
            const int j = MIRROR_LAST ? Z::j_mirror(tid) : tid / NL + i * Z::STEP;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
#pragma unroll
                for (int r = 0; r < RL; ++r) w[f][i][r] = last[f * Z::LDS + (j + r * NSL) * S + l];
#pragma unroll
                for (int r = 1; r < RL; ++r) {
                    cx<T> tw = twl[r * j];
                    if constexpr (INV) tw.y = -tw.y;
                    w[f][i][r] = cmul(w[f][i][r], tw);
                }
                dft<RL, INV>(w[f][i]);
            }
        }
    }
}

template <typename T, int N, int NT>
__device__ __forceinline__ void load_twiddles(cx<T>* twl, const cx<T>* __restrict__ tw, int tid) {
#pragma unroll
    for (int i = tid; i < N; i += NT) twl[i] = tw[i];
}

}  // namespace sph

// The register-direct x pass with the velocity-maximum reduction (the
// adaptive time step), for float and double storage at every compiled length.
// Its own translation unit: next to the plain x pass in one module it changed
// that pass's register allocation.
#include "pass_table.hpp"

namespace sph {

template <typename T, int N, int L, bool XT>
void px_products_vmax_l(const PairXArgs<T>& a, hipStream_t s) {
    using ZV = ZTile<T, N, L>;
    const dim3 grid((a.ny + ZV::NL - 1) / ZV::NL, a.nz, 1);
    k_px_products_rd<T, N, L, true, XT><<<grid, ZV::NT, 0, s>>>(a);
}

template <typename T, int N>
void px_products_vmax(const PairXArgs<T>& a, int variant, bool xt, hipStream_t s) {
    using LN = Launch<T, N>;
    constexpr int L = LN::X0, L1 = LN::X1, LA = LN::XA;
    if constexpr (LN::kXT) {
        if (xt) {
            if (variant == 1)
                px_products_vmax_l<T, N, L1, true>(a, s);
            else if (LN::kAlt && variant == kXOrderVariant)
                px_products_vmax_l<T, N, LA, true>(a, s);
            else
                px_products_vmax_l<T, N, L, true>(a, s);
            return;
        }
    }
    if (variant == 1)
        px_products_vmax_l<T, N, L1, false>(a, s);
    else if (LN::kAlt && variant == kXOrderVariant)
        px_products_vmax_l<T, N, LA, false>(a, s);
    else
        px_products_vmax_l<T, N, L, false>(a, s);
}

#define SPH_VMAX(N)                                                                                 \
    template void px_products_vmax<float, N>(const PairXArgs<float>&, int, bool, hipStream_t);   \
    template void px_products_vmax<double, N>(const PairXArgs<double>&, int, bool, hipStream_t);
SPH_LENGTHS(SPH_VMAX)
#undef SPH_VMAX

}  // namespace sph

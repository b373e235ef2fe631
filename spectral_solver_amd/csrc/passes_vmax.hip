// The register-direct x pass with the velocity-maximum reduction (the
// adaptive time step), for float and double storage at every compiled length.
// Its own translation unit: next to the plain x pass in one module it changed
// that pass's register allocation.
#include "pass_table.hpp"

namespace sph {

template <typename T, int N>
void px_products_vmax(const PairXArgs<T>& a, int variant, bool xt, hipStream_t s) {
    constexpr int L = Launch<T, N>::X0;
    constexpr int L1 = Launch<T, N>::X1;
    const dim3 grid((a.ny + (variant == 1 ? L1 : L) - 1) / (variant == 1 ? L1 : L), a.nz, 1);
    if constexpr (Launch<T, N>::kXT) {
        if (xt) {
            if (variant == 1)
                k_px_products_rd<T, N, L1, true, true><<<grid, ZTile<T, N, L1>::NT, 0, s>>>(a);
            else
                k_px_products_rd<T, N, L, true, true><<<grid, ZTile<T, N, L>::NT, 0, s>>>(a);
            return;
        }
    }
    if (variant == 1)
        k_px_products_rd<T, N, L1, true><<<grid, ZTile<T, N, L1>::NT, 0, s>>>(a);
    else
        k_px_products_rd<T, N, L, true><<<grid, ZTile<T, N, L>::NT, 0, s>>>(a);
}

#define SPH_VMAX(N)                                                                                 \
    template void px_products_vmax<float, N>(const PairXArgs<float>&, int, bool, hipStream_t);   \
    template void px_products_vmax<double, N>(const PairXArgs<double>&, int, bool, hipStream_t);
SPH_LENGTHS(SPH_VMAX)
#undef SPH_VMAX

}  // namespace sph

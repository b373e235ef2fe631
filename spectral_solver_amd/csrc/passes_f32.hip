// Instantiates the pass kernels for float storage at the lengths of
// SPH_LENGTHS_A (those of SPH_LENGTHS_B: passes_f32_b.hip) and the lookup
// over both.
#include "pass_table.hpp"

namespace sph {

template <>
const PassTable<float>* pass_table_part<float, 0>(int n) {
    switch (n) {
#define SPH_CASE(N) \
    case N:         \
        return &Launch<float, N>::table;
        SPH_LENGTHS_A(SPH_CASE)
#undef SPH_CASE
        default:
            return nullptr;
    }
}

template <>
const PassTable<float>* pass_table<float>(int n) {
    const PassTable<float>* t = pass_table_part<float, 0>(n);
    return t ? t : pass_table_part<float, 1>(n);
}

}  // namespace sph

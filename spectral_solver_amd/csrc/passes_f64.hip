// Instantiates the pass kernels for double storage at the lengths of
// SPH_LENGTHS_A (those of SPH_LENGTHS_B: passes_f64_b.hip) and the lookup
// over both.
#include "pass_table.hpp"

namespace sph {

template <>
const PassTable<double>* pass_table_part<double, 0>(int n) {
    switch (n) {
#define SPH_CASE(N) \
    case N:         \
        return &Launch<double, N>::table;
        SPH_LENGTHS_A(SPH_CASE)
#undef SPH_CASE
        default:
            return nullptr;
    }
}

template <>
const PassTable<double>* pass_table<double>(int n) {
    const PassTable<double>* t = pass_table_part<double, 0>(n);
    return t ? t : pass_table_part<double, 1>(n);
}

}  // namespace sph

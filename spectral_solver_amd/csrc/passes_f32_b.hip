// Instantiates the pass kernels for float storage at the lengths of
// SPH_LENGTHS_B (a translation unit of its own, compiled beside the others).
#include "pass_table.hpp"

namespace sph {

template <>
const PassTable<float>* pass_table_part<float, 1>(int n) {
    switch (n) {
#define SPH_CASE(N) \
    case N:         \
        return &Launch<float, N>::table;
        SPH_LENGTHS_B(SPH_CASE)
#undef SPH_CASE
        default:
            return nullptr;
    }
}

}  // namespace sph

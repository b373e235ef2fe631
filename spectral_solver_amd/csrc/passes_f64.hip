// Instantiates the pass kernels for double storage at every compiled length.
#include "pass_table.hpp"

namespace sph {

template <>
const PassTable<double>* pass_table<double>(int n) {
    switch (n) {
#define SPH_CASE(N) \
    case N:         \
        return &Launch<double, N>::table;
        SPH_LENGTHS(SPH_CASE)
#undef SPH_CASE
        default:
            return nullptr;
    }
}

}  // namespace sph

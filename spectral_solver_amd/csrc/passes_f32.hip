// Instantiates the pass kernels for float storage at every compiled length.
#include "pass_table.hpp"

namespace sph {

template <>
const PassTable<float>* pass_table<float>(int n) {
    switch (n) {
#define SPH_CASE(N) \
    case N:         \
        return &Launch<float, N>::table;
        SPH_LENGTHS(SPH_CASE)
#undef SPH_CASE
        default:
            return nullptr;
    }
}

}  // namespace sph

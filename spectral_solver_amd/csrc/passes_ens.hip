// The one-launch-per-pass ensemble kernels (ensemble_passes.hpp) for float
// and double storage.
#include "ensemble_passes.hpp"

namespace sph {

template <typename T>
void ens_launch(int pass, const EnsGroupArgs<T>* table, const EnsGrid& grid, int stage, hipStream_t s) {
    if (grid.total == 0) return;
    const dim3 blocks(grid.total), threads(kEnsThreads);
    switch (pass) {
        case 0:
            k_ens_pass<T, 0><<<blocks, threads, 0, s>>>(table, grid, stage);
            break;
        case 1:
            k_ens_pass<T, 1><<<blocks, threads, 0, s>>>(table, grid, stage);
            break;
        case 2:
            k_ens_pass<T, 2><<<blocks, threads, 0, s>>>(table, grid, stage);
            break;
        case 3:
            k_ens_pass<T, 3><<<blocks, threads, 0, s>>>(table, grid, stage);
            break;
        default:
            k_ens_pass<T, 4><<<blocks, threads, 0, s>>>(table, grid, stage);
            break;
    }
}

template void ens_launch<float>(int, const EnsGroupArgs<float>*, const EnsGrid&, int, hipStream_t);
template void ens_launch<double>(int, const EnsGroupArgs<double>*, const EnsGrid&, int, hipStream_t);

}  // namespace sph

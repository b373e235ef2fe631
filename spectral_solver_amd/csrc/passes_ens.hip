// The one-launch-per-pass and one-launch-per-stage ensemble kernels
// (ensemble_passes.hpp) for float and double storage.
#include "common.hpp"
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

template <typename T>
int ens_stage_blocks() {
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_ens_stage<T>),
                                                     kEnsThreads, 0) != hipSuccess)
        return 0;
    return per_cu * cus / 8 * 8;
}

template <typename T>
void ens_launch_stage(const EnsGroupArgs<T>* table, const EnsStageGrids& grids, int blocks, int stage, hipStream_t s) {
    if (blocks < 8 || blocks % 8 != 0) throw std::runtime_error("one launch per stage: no resident grid");
    EnsStageGrids g = grids;
    int st = stage;
    void* args[] = {&table, &g, &st};
    hip_check(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_ens_stage<T>), dim3(blocks),
                                         dim3(kEnsThreads), args, 0, s),
              "hipLaunchCooperativeKernel (one launch per stage)");
}

template int ens_stage_blocks<float>();
template int ens_stage_blocks<double>();
template void ens_launch_stage<float>(const EnsGroupArgs<float>*, const EnsStageGrids&, int, int, hipStream_t);
template void ens_launch_stage<double>(const EnsGroupArgs<double>*, const EnsStageGrids&, int, int, hipStream_t);

}  // namespace sph

// kbg_sweep_f32b.hip -- FP32 fused 1D sweep, two distributions, KFVS.
#include "kbg_sweep.h"

namespace kbg {

hipError_t launch_sweep_f32_2kfvs(int wg_waves, const Solve1D* d_solves, const WorkItem* d_items, const Pack* d_packs,
                                  int npacks, int par, double* d_rpart, int nstep, hipStream_t st) {
    return launch_sweep_t<float, 2, KFVS>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
}

}  // namespace kbg

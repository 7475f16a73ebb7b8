// kbg_sweep_f32.hip -- FP32-storage instances of the fused 1D sweep (kbg_sweep.h) and the launcher.
#include "kbg_sweep.h"

namespace kbg {

hipError_t launch_sweep1d_f64(int nf, int flux, int nw, const Solve1D* d_solves, const WorkItem* d_items, int nitems,
                              int par, double* d_rpart, hipStream_t st);

hipError_t launch_sweep1d(bool fp64, int nf, int flux, int nw, const Solve1D* d_solves, const WorkItem* d_items,
                          int nitems, int par, double* d_rpart, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    if (nw < 1 || nw > kSweepMaxWaves) return hipErrorInvalidValue;
    if (fp64) return launch_sweep1d_f64(nf, flux, nw, d_solves, d_items, nitems, par, d_rpart, st);
    if (nf == 1 && flux == KFVS) return launch_sweep_t<float, 1, KFVS>(nw, d_solves, d_items, nitems, par, d_rpart, st);
    if (nf == 1 && flux == KCU) return launch_sweep_t<float, 1, KCU>(nw, d_solves, d_items, nitems, par, d_rpart, st);
    if (nf == 2 && flux == KFVS) return launch_sweep_t<float, 2, KFVS>(nw, d_solves, d_items, nitems, par, d_rpart, st);
    if (nf == 2 && flux == KCU) return launch_sweep_t<float, 2, KCU>(nw, d_solves, d_items, nitems, par, d_rpart, st);
    if (nf == 2 && flux == UGKS) return launch_sweep_t<float, 2, UGKS>(nw, d_solves, d_items, nitems, par, d_rpart, st);
    return hipErrorInvalidValue;
}

}  // namespace kbg

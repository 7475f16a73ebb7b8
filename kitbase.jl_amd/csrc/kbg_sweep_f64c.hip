// kbg_sweep_f64c.hip -- FP64 fused 1D sweep, KCU and UGKS fluxes.
#include "kbg_sweep.h"

namespace kbg {

hipError_t launch_sweep_f64_kinetic(int nf, int flux, int wg_waves, const Solve1D* d_solves, const WorkItem* d_items,
                                    const Pack* d_packs, int npacks, int par, double* d_rpart, int nstep, hipStream_t st) {
    if (nf == 1 && flux == KCU) return launch_sweep_t<double, 1, KCU>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    if (nf == 2 && flux == KCU) return launch_sweep_t<double, 2, KCU>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    if (nf == 2 && flux == UGKS) return launch_sweep_t<double, 2, UGKS>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    return hipErrorInvalidValue;
}

}  // namespace kbg

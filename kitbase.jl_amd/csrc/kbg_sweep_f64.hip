// kbg_sweep_f64.hip -- FP64 instantiations of the fused 1D sweep (KFVS); KCU and UGKS are compiled in
// kbg_sweep_f64c.hip so the translation units build in parallel.
#include "kbg_sweep.h"

namespace kbg {

hipError_t launch_sweep_f64_kinetic(int nf, int flux, int wg_waves, const Solve1D* d_solves, const WorkItem* d_items,
                                    const Pack* d_packs, int npacks, int par, double* d_rpart, int nstep, hipStream_t st);

hipError_t launch_sweep1d_f64(int nf, int flux, int wg_waves, const Solve1D* d_solves, const WorkItem* d_items,
                              const Pack* d_packs, int npacks, int par, double* d_rpart, int nstep, hipStream_t st) {
    if (nf == 1 && flux == KFVS) return launch_sweep_t<double, 1, KFVS>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    if (nf == 2 && flux == KFVS) return launch_sweep_t<double, 2, KFVS>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    return launch_sweep_f64_kinetic(nf, flux, wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
}

}  // namespace kbg

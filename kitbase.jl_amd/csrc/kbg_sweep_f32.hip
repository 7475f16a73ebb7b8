// kbg_sweep_f32.hip -- FP32 instantiations of the fused 1D sweep (one distribution, KFVS) and the
// dispatcher; the other flux / distribution combinations are compiled in kbg_sweep_f32b.hip and
// kbg_sweep_f32c.hip so the translation units build in parallel.
#include "kbg_sweep.h"

namespace kbg {

hipError_t launch_sweep1d_f64(int nf, int flux, int wg_waves, const Solve1D* d_solves, const WorkItem* d_items,
                              const Pack* d_packs, int npacks, int par, double* d_rpart, int nstep, hipStream_t st);
hipError_t launch_sweep_f32_2kfvs(int wg_waves, const Solve1D* d_solves, const WorkItem* d_items, const Pack* d_packs,
                                  int npacks, int par, double* d_rpart, int nstep, hipStream_t st);
hipError_t launch_sweep_f32_kinetic(int nf, int flux, int wg_waves, const Solve1D* d_solves, const WorkItem* d_items,
                                    const Pack* d_packs, int npacks, int par, double* d_rpart, int nstep, hipStream_t st);

hipError_t launch_sweep1d(bool fp64, int nf, int flux, int wg_waves, const Solve1D* d_solves, const WorkItem* d_items,
                          const Pack* d_packs, int npacks, int par, double* d_rpart, int nstep, hipStream_t st) {
    if (npacks <= 0) return hipSuccess;
    if (wg_waves < 1 || wg_waves > kSweepMaxWaves) return hipErrorInvalidValue;
    if (fp64) return launch_sweep1d_f64(nf, flux, wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    if (nf == 1 && flux == KFVS) return launch_sweep_t<float, 1, KFVS>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    if (nf == 2 && flux == KFVS) return launch_sweep_f32_2kfvs(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
    return launch_sweep_f32_kinetic(nf, flux, wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
}

}  // namespace kbg

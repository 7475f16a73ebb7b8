// kbg_sweep_f32b.hip -- FP32 fused 1D sweep, two distributions, KFVS; the conversions of the FP32
// sweep's stored departures.
#include "kbg_sweep.h"

namespace kbg {

hipError_t launch_sweep_f32_2kfvs(int wg_waves, const Solve1D* d_solves, const WorkItem* d_items, const Pack* d_packs,
                                  int npacks, int par, double* d_rpart, int nstep, hipStream_t st) {
    return launch_sweep_t<float, 2, KFVS>(wg_waves, d_solves, d_items, d_packs, npacks, par, d_rpart, nstep, st);
}

// Conversions of an FP32 swept solve's f buffer between the stored departure d = f - M~(prim) and f
// (kbg_sweep.h maxw_coef / maxw_at): one thread per (storage cell, distribution row).
__global__ __launch_bounds__(256) void k_delta1d(const Solve1D* __restrict__ pS, double* __restrict__ stage, int mode, int buf) {
    const Solve1D& S = *pS;
    const int N = S.nx + 2 * S.ng;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;  // q nv + v
    if (k >= N) return;
    const int q = row >= S.nv ? 1 : 0, v = row - q * S.nv;
    double p[3];
    conserve_prim<3>(S.w[buf] + (size_t)k * 3, S.gam, p);  // the w buffer paired with f[buf]
    const MaxwCoef mc = maxw_coef(p, S.K);
    const VChunk<float>* vt = (const VChunk<float>*)S.vtab;
    float M = maxw_at(vt[v / kVChunk], mc.A, mc.nl2, mc.U, v % kVChunk);
    if (q == 1) M = maxw_b(M, mc.kb);
    const size_t o = (size_t)row * S.np + k;
    float* f = (float*)S.f[buf];
    if (mode == 0) f[o] = float(stage[o] - double(M));
    else if (mode == 1) stage[o] = double(M) + double(f[o]);
    else if (mode == 2) f[o] = M + f[o];
    else f[o] = f[o] - M;
}

hipError_t launch_delta1d(const Solve1D& S, const Solve1D* d_S, double* stage, int mode, int buf, hipStream_t st) {
    const int N = S.nx + 2 * S.ng;
    if (!S.delta || N <= 0) return hipSuccess;
    if (mode < 0 || mode > 3 || ((mode == 0 || mode == 1) && !stage)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_delta1d, dim3((N + 255) / 256, S.nf * S.nv), dim3(256), 0, st, d_S, stage, mode, buf);
    return hipGetLastError();
}

}  // namespace kbg

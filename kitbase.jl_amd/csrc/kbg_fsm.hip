// kbg_fsm.hip -- KitBase's fast spectral Boltzmann collision operator (boltzmann_fft!,
// theory_fsm.jl:271-314) for the 1D1F3V step!(:fsm) (step_1f.jl:122-157), on gfx950.
//
// Per step, for the interior cells c0 .. c0 + ncell - 1 of a solve (FP64 throughout):
//   k_fsm_load   f (storage precision) -> complex spec[cell][n3]
//   hipFFT       batched 3D backward transforms (bfft, unnormalised), one per cell
//   k_fsm_modes  f_spec = fftshift(spec) / n3, then per cell the 2 M(M-1) + 2 products
//                f_spec phi_d, f_spec psi_d (d < M(M-1)), f_spec phipsi, f_spec -> work
//   hipFFT       batched 3D forward transforms of all the products in place
//   k_fsm_q      Q = 4 pi^2 / Kn / M^2 Re(sum_d F[f_spec phi_d] F[f_spec psi_d]
//                                      - F[f_spec phipsi] F[f_spec])
// The kernel modes phi, psi, phipsi (kernel_mode, theory_fsm.jl:24-90) come from the host with the
// descriptor.  Arrays are [w][v][u] with u fastest (KitBase's column-major f[i, j, k]).  The update
// kernel then adds dt Q to the transport update (k_update1d).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include "kbg_device.h"
#include "kbg_types.h"

namespace kbg {

namespace {

__global__ __launch_bounds__(256) void k_fsm_load_f32(const float* __restrict__ f, long fcs, int k0, int ncell, int n3,
                                                      double2* __restrict__ spec) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    spec[i] = make_double2(double(f[(long)(k0 + ci) * fcs + q]), 0.0);
}

__global__ __launch_bounds__(256) void k_fsm_load_f64(const double* __restrict__ f, long fcs, int k0, int ncell, int n3,
                                                      double2* __restrict__ spec) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    spec[i] = make_double2(f[(long)(k0 + ci) * fcs + q], 0.0);
}

// fftshift: element (i, j, k) of the shifted array is element ((i + nu/2) % nu, ...) of the input
// for even sizes; in general the input index whose shift lands on (i, j, k), as Julia's
// circshift(x, n ÷ 2) and numpy's roll(x, n // 2) place it
__global__ __launch_bounds__(256) void k_fsm_modes(const double2* __restrict__ spec, double2* __restrict__ work,
                                                   const double* __restrict__ phi, const double* __restrict__ psi,
                                                   const double* __restrict__ phipsi, int nu, int nv, int nw, int nd,
                                                   int ncell) {
    const int n3 = nu * nv * nw;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    const int iu = q % nu, jv = (q / nu) % nv, kw = q / (nu * nv);
    const int su = (iu - nu / 2 + nu) % nu, sv = (jv - nv / 2 + nv) % nv, sw = (kw - nw / 2 + nw) % nw;
    const double2 x = spec[(size_t)ci * n3 + su + nu * (sv + nv * sw)];
    const double inv = 1.0 / double(n3);
    const double sr = x.x * inv, si = x.y * inv;
    double2* __restrict__ out = work + (size_t)ci * (2 * nd + 2) * n3 + q;
    for (int d = 0; d < nd; d++) {
        const double a = phi[(size_t)d * n3 + q], b = psi[(size_t)d * n3 + q];
        out[(size_t)(2 * d) * n3] = make_double2(sr * a, si * a);
        out[(size_t)(2 * d + 1) * n3] = make_double2(sr * b, si * b);
    }
    const double c = phipsi[q];
    out[(size_t)(2 * nd) * n3] = make_double2(sr * c, si * c);
    out[(size_t)(2 * nd + 1) * n3] = make_double2(sr, si);
}

__global__ __launch_bounds__(256) void k_fsm_q(const double2* __restrict__ work, double* __restrict__ Q, int n3, int nd,
                                               int ncell, double scale) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    const double2* __restrict__ in = work + (size_t)ci * (2 * nd + 2) * n3 + q;
    double re = 0.0;
    for (int d = 0; d < nd; d++) {
        const double2 a = in[(size_t)(2 * d) * n3], b = in[(size_t)(2 * d + 1) * n3];
        re += a.x * b.x - a.y * b.y;
    }
    const double2 a = in[(size_t)(2 * nd) * n3], b = in[(size_t)(2 * nd + 1) * n3];
    re -= a.x * b.x - a.y * b.y;
    Q[i] = scale * re;
}

unsigned blocks_of(size_t n) { return unsigned((n + 255) / 256); }

}  // namespace

hipError_t fsm_collision(const FsmPlan& P, bool fp64, const void* f, long fcs, int k0, hipStream_t st) {
    if (P.ncell <= 0) return hipSuccess;
    const int n3 = P.nu * P.nv * P.nw, nd = P.M * (P.M - 1);
    const size_t cells = (size_t)P.ncell * n3;
    if (fp64) hipLaunchKernelGGL(k_fsm_load_f64, dim3(blocks_of(cells)), dim3(256), 0, st, (const double*)f, fcs, k0, P.ncell, n3, P.spec);
    else hipLaunchKernelGGL(k_fsm_load_f32, dim3(blocks_of(cells)), dim3(256), 0, st, (const float*)f, fcs, k0, P.ncell, n3, P.spec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (hipfftSetStream(P.plan_b, st) != HIPFFT_SUCCESS || hipfftSetStream(P.plan_f, st) != HIPFFT_SUCCESS) return hipErrorUnknown;
    if (hipfftExecZ2Z(P.plan_b, (hipfftDoubleComplex*)P.spec, (hipfftDoubleComplex*)P.spec, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    hipLaunchKernelGGL(k_fsm_modes, dim3(blocks_of(cells)), dim3(256), 0, st, P.spec, P.work, P.phi, P.psi, P.phipsi, P.nu, P.nv,
                       P.nw, nd, P.ncell);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (hipfftExecZ2Z(P.plan_f, (hipfftDoubleComplex*)P.work, (hipfftDoubleComplex*)P.work, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    const double scale = 4.0 * kPi * kPi / P.kn / (double(P.M) * P.M);
    hipLaunchKernelGGL(k_fsm_q, dim3(blocks_of(cells)), dim3(256), 0, st, P.work, P.Q, n3, nd, P.ncell, scale);
    return hipGetLastError();
}

hipError_t fsm_plan_create(FsmPlan& P) {
    const int n3 = P.nu * P.nv * P.nw, nd = P.M * (P.M - 1);
    int n[3] = {P.nw, P.nv, P.nu};
    if (hipfftPlanMany(&P.plan_b, 3, n, nullptr, 1, n3, nullptr, 1, n3, HIPFFT_Z2Z, P.ncell) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    if (hipfftPlanMany(&P.plan_f, 3, n, nullptr, 1, n3, nullptr, 1, n3, HIPFFT_Z2Z, P.ncell * (2 * nd + 2)) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    P.planned = true;
    return hipSuccess;
}

void fsm_plan_destroy(FsmPlan& P) {
    if (!P.planned) return;
    hipfftDestroy(P.plan_b);
    hipfftDestroy(P.plan_f);
    P.planned = false;
}

}  // namespace kbg

// kbg_fsm.hip -- KitBase's fast spectral Boltzmann collision operator (boltzmann_fft!,
// theory_fsm.jl:271-314) for the 1D1F3V step!(:fsm) (step_1f.jl:122-157), on gfx950.
//
// Per step, for the interior cells c0 .. c0 + ncell - 1 of a solve:
//   k_fsm_load   f (storage precision) -> complex spec[cell][n3]
//   hipFFT       batched 3D backward transforms (bfft, unnormalised), one per cell
//   k_fsm_modes  f_spec = fftshift(spec) / n3, then per cell the 2 M(M-1) + 2 products
//                f_spec phi_d, f_spec psi_d (d < M(M-1)), f_spec phipsi, f_spec -> work
//   hipFFT       batched 3D forward transforms of all the products in place
//   k_fsm_q      Q = 4 pi^2 / Kn / M^2 Re(sum_d F[f_spec phi_d] F[f_spec psi_d]
//                                      - F[f_spec phipsi] F[f_spec])   (sums in FP64)
// The transforms, the spectra and the kernel modes are FP64 for FP64 storage and FP32 for FP32
// storage (R below).  hipFFT's forward pass streams the 2 M(M-1) + 2 product arrays per cell through
// HBM about a dozen times; on cubic grids of side 8, 12, 16, 18 or 24 the forward transforms are the
// library's own pencil transforms instead (kbg_fft.h): FP32 entirely in LDS (k_fsm_pair), FP64
// written once and read once (k_fsm_slab64, k_fsm_q).  Q is FP64 in all of them.  The kernel modes phi,
// psi, phipsi (kernel_mode, theory_fsm.jl:24-90) come from the host with the descriptor.  Arrays are
// [w][v][u] with u fastest (KitBase's column-major f[i, j, k]).  The update kernel then adds dt Q to
// the transport update (k_update1d).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include "kbg_device.h"
#include "kbg_fft.h"
#include "kbg_types.h"

namespace kbg {

namespace {

template <class R> struct Cx;
template <> struct Cx<double> { using type = double2; };
template <> struct Cx<float> { using type = float2; };
template <class R> using cx_t = typename Cx<R>::type;

template <class T, class R>
__global__ __launch_bounds__(256) void k_fsm_load(const T* __restrict__ f, long fcs, int k0, int ncell, int n3,
                                                  cx_t<R>* __restrict__ spec) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    cx_t<R> z;
    z.x = R(f[(long)(k0 + ci) * fcs + q]);
    z.y = R(0);
    spec[i] = z;
}

// fftshift: element (i, j, k) of the shifted array is element ((i + nu/2) % nu, ...) of the input
// for even sizes; in general the input index whose shift lands on (i, j, k), as Julia's
// circshift(x, n ÷ 2) and numpy's roll(x, n // 2) place it
template <class R>
__global__ __launch_bounds__(256) void k_fsm_modes(const cx_t<R>* __restrict__ spec, cx_t<R>* __restrict__ work,
                                                   const R* __restrict__ phi, const R* __restrict__ psi,
                                                   const R* __restrict__ phipsi, int nu, int nv, int nw, int nd, int ncell) {
    const int n3 = nu * nv * nw;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    const int iu = q % nu, jv = (q / nu) % nv, kw = q / (nu * nv);
    const int su = (iu - nu / 2 + nu) % nu, sv = (jv - nv / 2 + nv) % nv, sw = (kw - nw / 2 + nw) % nw;
    const cx_t<R> x = spec[(size_t)ci * n3 + su + nu * (sv + nv * sw)];
    const double inv = 1.0 / double(n3);
    const R sr = R(double(x.x) * inv), si = R(double(x.y) * inv);
    cx_t<R>* __restrict__ out = work + (size_t)ci * (2 * nd + 2) * n3 + q;
    cx_t<R> z;
    for (int d = 0; d < nd; d++) {
        const R a = phi[(size_t)d * n3 + q], b = psi[(size_t)d * n3 + q];
        z.x = sr * a;
        z.y = si * a;
        out[(size_t)(2 * d) * n3] = z;
        z.x = sr * b;
        z.y = si * b;
        out[(size_t)(2 * d + 1) * n3] = z;
    }
    const R c = phipsi[q];
    z.x = sr * c;
    z.y = si * c;
    out[(size_t)(2 * nd) * n3] = z;
    z.x = sr;
    z.y = si;
    out[(size_t)(2 * nd + 1) * n3] = z;
}

template <class R>
__global__ __launch_bounds__(256) void k_fsm_q(const cx_t<R>* __restrict__ work, double* __restrict__ Q, int n3, int nd,
                                               int ncell, double scale) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    const cx_t<R>* __restrict__ in = work + (size_t)ci * (2 * nd + 2) * n3 + q;
    double re = 0.0;
    for (int d = 0; d < nd; d++) {
        const cx_t<R> a = in[(size_t)(2 * d) * n3], b = in[(size_t)(2 * d + 1) * n3];
        re += double(a.x) * double(b.x) - double(a.y) * double(b.y);
    }
    const cx_t<R> a = in[(size_t)(2 * nd) * n3], b = in[(size_t)(2 * nd + 1) * n3];
    re -= double(a.x) * double(b.x) - double(a.y) * double(b.y);
    Q[i] = scale * re;
}

// ---- FP32 storage on a cubic grid of N^3 velocities (N from fsm_cube_size): the forward transforms
// in LDS, no product array.  One workgroup per (cell, pair): pair p < nd transforms f_spec phi_p and
// f_spec psi_p, pair nd transforms f_spec phipsi and f_spec; the product Re(A B) of the two
// transforms is written per pair (part[cell][pair][n3]) and k_fsm_sum adds the pairs in a fixed
// order.  Each transform, with thread t < N^2 at (ta, tb) = (t % N, t / N):
//   1. the thread loads its w-pencil (u, v) = (ta, tb) of f_spec times the mode straight from global
//      memory (f_spec = fftshift(bfft(f)) / N^3), transforms it along w in registers and stores it
//      into an LDS array [w][v][u] whose rows are padded to an odd length (N | 1, so the pencils a
//      wavefront reads fall in different banks);
//   2. transforms the u-pencil (v, w) = (ta, tb) in LDS;
//   3. transforms the v-pencil (u, w) = (ta, tb) and keeps it in registers, where the second
//      transform's pencil meets the first.
KBG_D int fsm_xcd_item() {
    const int G = gridDim.x, bx = blockIdx.x;
    const int x8 = bx & 7, q8 = G >> 3, r8 = G & 7;
    return x8 * q8 + (x8 < r8 ? x8 : r8) + (bx >> 3);
}

template <int N> constexpr int fsm_pair_threads() { return (N * N + 63) / 64 * 64; }
// the first transform's kept v-pencil: its first KEEP values wait in LDS behind the array (registers
// for the rest), which keeps the kernel within the register budget of 576-thread workgroups
constexpr int kFsmKeep = 8;
template <int N> constexpr int fsm_pair_lds() { return N * N * (N | 1) * 8 + kFsmKeep * N * N * 8; }

template <int N>
__global__ __launch_bounds__(fsm_pair_threads<N>()) void k_fsm_pair(const float2* __restrict__ spec, const float* __restrict__ phi,
                                                                  const float* __restrict__ psi, const float* __restrict__ phipsi,
                                                                  const float2* __restrict__ tw, float* __restrict__ part, int nd,
                                                                  int ncell) {
    constexpr int N2 = N * N, N3 = N2 * N, P = N | 1, H = N / 2;
    extern __shared__ float2 lds[];
    const int np = nd + 1;
    const int item = fsm_xcd_item();
    if (item >= ncell * np) return;
    const int ci = item / np, p = item % np;
    const int t = threadIdx.x;
    const bool act = t < N2;
    const int ta = t % N, tb = t / N;
    // the thread's column of the unshifted spectrum: fftshift places element (ta, tb, k) at
    // ((ta + N - N/2) % N, (tb + N - N/2) % N, (k + N - N/2) % N)
    const float2* __restrict__ col = spec + (size_t)ci * N3 + (ta + N - H) % N + N * ((tb + N - H) % N);
    const float inv = 1.0f / float(N3);
    float2* keep = lds + N2 * P;  // [kFsmKeep][N2]: the first values of each thread's kept pencil
    float2 A[N];
    for (int h = 0; h < 2; h++) {
        const float* __restrict__ mode = p < nd ? (h ? psi : phi) + (size_t)p * N3 : (h ? nullptr : phipsi);
        if (h) __syncthreads();  // every pencil of the first transform has left the array
        float2 x[N], X[N];
        if (act) {  // 1. along w, from global memory
#pragma unroll
            for (int k = 0; k < N; k++) x[k] = col[N2 * ((k + N - H) % N)];
            const float* __restrict__ mc = mode + ta + N * tb;
#pragma unroll
            for (int k = 0; k < N; k++) {
                const float m = (mode ? mc[N2 * k] : 1.0f) * inv;
                x[k] = make_float2(x[k].x * m, x[k].y * m);
            }
            fft_dit<N, N, 1, float2>(x, X, tw);
            float2* r = lds + ta + P * tb;
#pragma unroll
            for (int k = 0; k < N; k++) r[P * N * k] = X[k];
        }
        __syncthreads();
        if (act) {  // 2. along u: pencil (v, w) = (ta, tb)
            float2* r = lds + P * (ta + N * tb);
#pragma unroll
            for (int k = 0; k < N; k++) x[k] = r[k];
            fft_dit<N, N, 1, float2>(x, X, tw);
#pragma unroll
            for (int k = 0; k < N; k++) r[k] = X[k];
        }
        __syncthreads();
        if (act) {  // 3. along v: pencil (u, w) = (ta, tb), kept in registers
            const float2* r = lds + ta + P * N * tb;
#pragma unroll
            for (int k = 0; k < N; k++) x[k] = r[P * k];
            fft_dit<N, N, 1, float2>(x, X, tw);
            if (h == 0) {
#pragma unroll
                for (int k = 0; k < N; k++) {
                    if (k < kFsmKeep) keep[k * N2 + t] = X[k];  // this thread's slots only
                    else A[k] = X[k];
                }
            } else {
                float* __restrict__ o = part + (size_t)item * N3 + ta + N2 * tb;
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const float2 a = k < kFsmKeep ? keep[k * N2 + t] : A[k];
                    o[N * k] = a.x * X[k].x - a.y * X[k].y;
                }
            }
        }
    }
}

// ---- FP64 storage on a cubic grid: the transforms of the products f_spec phi_d, f_spec psi_d,
// f_spec phipsi, f_spec in the layout of k_fsm_modes (work[cell][2 M(M-1) + 2][n3]), each written
// once by k_fsm_slab64, then k_fsm_q.  One workgroup per slab of S w-planes of one product array:
//   1. every w-pencil (u, v) of the array: f_spec times the mode straight from the backward
//      transform's output (the fftshift is an index map), transformed along w in registers; the S
//      frequencies of the slab go to LDS (rows padded to an odd length).  The N / S slabs of an array
//      each redo these transforms: arithmetic from L2-resident inputs instead of a pass over HBM;
//   2. one thread per u-pencil of the slab, in LDS;
//   3. one thread per v-pencil, whose transform goes straight to the array.
template <int N> constexpr int fsm_slab() { return N % 8 == 0 ? 8 : 6; }
template <int N> constexpr int fsm_slab_threads() { return (fsm_slab<N>() * N + 63) / 64 * 64; }
template <int N> constexpr int fsm_slab_lds() { return fsm_slab<N>() * N * (N | 1) * 16; }

template <int N>
__global__ __launch_bounds__(fsm_slab_threads<N>()) void k_fsm_slab64(const double2* __restrict__ spec, const double* __restrict__ phi,
                                                                    const double* __restrict__ psi, const double* __restrict__ phipsi,
                                                                    const double2* __restrict__ tw, double2* __restrict__ work, int nd,
                                                                    int ncell) {
    constexpr int N2 = N * N, N3 = N2 * N, P = N | 1, H = N / 2, S = fsm_slab<N>(), NS = N / S;
    constexpr int NTH = fsm_slab_threads<N>();
    extern __shared__ double2 slab[];
    const int na = 2 * nd + 2;
    const int item = fsm_xcd_item();
    if (item >= ncell * na * NS) return;
    const int sl = item % NS, a = (item / NS) % na, ci = item / (NS * na);
    const int p = a >> 1, h = a & 1;
    const double* __restrict__ mode = p < nd ? (h ? psi : phi) + (size_t)p * N3 : (h ? nullptr : phipsi);
    const double inv = 1.0 / double(N3);
    const int t = threadIdx.x;
    double2 x[N], X[N];
    // 1. along w, every pencil of the array; the slab's planes kept
    for (int pen = t; pen < N2; pen += NTH) {
        const int ta = pen % N, tb = pen / N;
        const double2* __restrict__ col = spec + (size_t)ci * N3 + (ta + N - H) % N + N * ((tb + N - H) % N);
#pragma unroll
        for (int k = 0; k < N; k++) {
            const double2 z = col[N2 * ((k + N - H) % N)];
            const double m = mode ? mode[pen + N2 * k] : 1.0;
            x[k] = make_double2((z.x * inv) * m, (z.y * inv) * m);
        }
        fft_dit<N, N, 1, double2>(x, X, tw);
#pragma unroll
        for (int k = 0; k < N; k++)  // compile-time register indices, a uniform branch per plane
            if (k / S == sl) slab[ta + P * (tb + N * (k % S))] = X[k];
    }
    __syncthreads();
    const bool act = t < S * N;
    const int ta = t % N, tb = t / N;
    if (act) {  // 2. along u: pencil (v, w) = (ta, tb)
        double2* r = slab + P * (ta + N * tb);
#pragma unroll
        for (int k = 0; k < N; k++) x[k] = r[k];
        fft_dit<N, N, 1, double2>(x, X, tw);
#pragma unroll
        for (int k = 0; k < N; k++) r[k] = X[k];
    }
    __syncthreads();
    if (act) {  // 3. along v: pencil (u, w) = (ta, tb), to the array
        const double2* r = slab + ta + P * N * tb;
#pragma unroll
        for (int k = 0; k < N; k++) x[k] = r[P * k];
        fft_dit<N, N, 1, double2>(x, X, tw);
        double2* __restrict__ o = work + ((size_t)ci * na + a) * N3 + ta + N2 * (sl * S + tb);
#pragma unroll
        for (int k = 0; k < N; k++) o[N * k] = X[k];
    }
}

// Q = scale (sum_{p < nd} part_p - part_nd), pairs in order
__global__ __launch_bounds__(256) void k_fsm_sum(const float* __restrict__ part, double* __restrict__ Q, int n3, int nd,
                                                 int ncell, double scale) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)ncell * n3) return;
    const int ci = int(i / n3), q = int(i % n3);
    const float* __restrict__ pp = part + (size_t)ci * (nd + 1) * n3 + q;
    double re = 0.0;
    for (int d = 0; d < nd; d++) re += double(pp[(size_t)d * n3]);
    re -= double(pp[(size_t)nd * n3]);
    Q[i] = scale * re;
}

template <int N> hipError_t launch_pair(const FsmPlan& P, hipStream_t st) {
    const int nd = P.M * (P.M - 1);
    hipLaunchKernelGGL(k_fsm_pair<N>, dim3(P.ncell * (nd + 1)), dim3(fsm_pair_threads<N>()), fsm_pair_lds<N>(), st,
                       (const float2*)P.spec, (const float*)P.phi, (const float*)P.psi, (const float*)P.phipsi,
                       (const float2*)P.tw, P.part, nd, P.ncell);
    return hipGetLastError();
}

template <int N> hipError_t launch_cube64(const FsmPlan& P, hipStream_t st) {
    const int nd = P.M * (P.M - 1), na = 2 * nd + 2;
    const unsigned blocks = unsigned(P.ncell * na * (N / fsm_slab<N>()));
    hipLaunchKernelGGL(k_fsm_slab64<N>, dim3(blocks), dim3(fsm_slab_threads<N>()), fsm_slab_lds<N>(), st, (const double2*)P.spec,
                       (const double*)P.phi, (const double*)P.psi, (const double*)P.phipsi, (const double2*)P.tw,
                       (double2*)P.work, nd, P.ncell);
    return hipGetLastError();
}

template <int N> hipError_t cube_attr() {
    hipError_t e = hipFuncSetAttribute((const void*)k_fsm_pair<N>, hipFuncAttributeMaxDynamicSharedMemorySize, fsm_pair_lds<N>());
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_fsm_slab64<N>, hipFuncAttributeMaxDynamicSharedMemorySize, fsm_slab_lds<N>());
}

unsigned blocks_of(size_t n) { return unsigned((n + 255) / 256); }

hipfftResult exec(const FsmPlan& P, hipfftHandle_t* plan, void* data, int dir) {
    if (P.f32) return hipfftExecC2C(plan, (hipfftComplex*)data, (hipfftComplex*)data, dir);
    return hipfftExecZ2Z(plan, (hipfftDoubleComplex*)data, (hipfftDoubleComplex*)data, dir);
}

template <class R> hipError_t fsm_run(const FsmPlan& P, bool fp64, const void* f, long fcs, int k0, hipStream_t st) {
    const int n3 = P.nu * P.nv * P.nw, nd = P.M * (P.M - 1);
    const size_t cells = (size_t)P.ncell * n3;
    const dim3 g(blocks_of(cells)), b(256);
    cx_t<R>* spec = (cx_t<R>*)P.spec;
    cx_t<R>* work = (cx_t<R>*)P.work;
    if (fp64) hipLaunchKernelGGL((k_fsm_load<double, R>), g, b, 0, st, (const double*)f, fcs, k0, P.ncell, n3, spec);
    else hipLaunchKernelGGL((k_fsm_load<float, R>), g, b, 0, st, (const float*)f, fcs, k0, P.ncell, n3, spec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (hipfftSetStream(P.plan_b, st) != HIPFFT_SUCCESS) return hipErrorUnknown;
    if (exec(P, P.plan_b, spec, HIPFFT_BACKWARD) != HIPFFT_SUCCESS) return hipErrorUnknown;
    const double scale = 4.0 * kPi * kPi / P.kn / (double(P.M) * P.M);
    if constexpr (sizeof(R) == 4) {
        if (P.cube_n) {
            switch (P.cube_n) {
            case 8: e = launch_pair<8>(P, st); break;
            case 12: e = launch_pair<12>(P, st); break;
            case 16: e = launch_pair<16>(P, st); break;
            case 18: e = launch_pair<18>(P, st); break;
            default: e = launch_pair<24>(P, st); break;
            }
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_fsm_sum, g, b, 0, st, P.part, P.Q, n3, nd, P.ncell, scale);
            return hipGetLastError();
        }
    }
    if (P.cube_n) {
        switch (P.cube_n) {
        case 8: e = launch_cube64<8>(P, st); break;
        case 12: e = launch_cube64<12>(P, st); break;
        case 16: e = launch_cube64<16>(P, st); break;
        case 18: e = launch_cube64<18>(P, st); break;
        default: e = launch_cube64<24>(P, st); break;
        }
        if (e != hipSuccess) return e;
    } else {
        if (hipfftSetStream(P.plan_f, st) != HIPFFT_SUCCESS) return hipErrorUnknown;
        hipLaunchKernelGGL(k_fsm_modes<R>, g, b, 0, st, spec, work, (const R*)P.phi, (const R*)P.psi, (const R*)P.phipsi, P.nu,
                           P.nv, P.nw, nd, P.ncell);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (exec(P, P.plan_f, work, HIPFFT_FORWARD) != HIPFFT_SUCCESS) return hipErrorUnknown;
    }
    hipLaunchKernelGGL(k_fsm_q<R>, g, b, 0, st, work, P.Q, n3, nd, P.ncell, scale);
    return hipGetLastError();
}

}  // namespace

int fsm_cube_size(int nu, int nv, int nw) {
    if (nu != nv || nu != nw) return 0;
    for (int n : {8, 12, 16, 18, 24})
        if (n == nu) return n;
    return 0;
}

hipError_t fsm_collision(const FsmPlan& P, bool fp64, const void* f, long fcs, int k0, hipStream_t st) {
    if (P.ncell <= 0) return hipSuccess;
    return P.f32 ? fsm_run<float>(P, fp64, f, fcs, k0, st) : fsm_run<double>(P, fp64, f, fcs, k0, st);
}

hipError_t fsm_plan_create(FsmPlan& P) {
    const int n3 = P.nu * P.nv * P.nw, nd = P.M * (P.M - 1);
    int n[3] = {P.nw, P.nv, P.nu};
    const hipfftType t = P.f32 ? HIPFFT_C2C : HIPFFT_Z2Z;
    if (P.cube_n) {
        hipError_t e;
        switch (P.cube_n) {
        case 8: e = cube_attr<8>(); break;
        case 12: e = cube_attr<12>(); break;
        case 16: e = cube_attr<16>(); break;
        case 18: e = cube_attr<18>(); break;
        default: e = cube_attr<24>(); break;
        }
        if (e != hipSuccess) return e;
    }
    if (hipfftPlanMany(&P.plan_b, 3, n, nullptr, 1, n3, nullptr, 1, n3, t, P.ncell) != HIPFFT_SUCCESS) return hipErrorUnknown;
    if (!P.cube_n &&
        hipfftPlanMany(&P.plan_f, 3, n, nullptr, 1, n3, nullptr, 1, n3, t, P.ncell * (2 * nd + 2)) != HIPFFT_SUCCESS) {
        hipfftDestroy(P.plan_b);
        return hipErrorUnknown;
    }
    P.planned = true;
    return hipSuccess;
}

void fsm_plan_destroy(FsmPlan& P) {
    if (!P.planned) return;
    hipfftDestroy(P.plan_b);
    if (!P.cube_n) hipfftDestroy(P.plan_f);
    P.planned = false;
}

}  // namespace kbg

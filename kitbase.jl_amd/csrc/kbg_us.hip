// kbg_us.hip -- the unstructured-mesh DVM on triangles (2D1F2V and 2D2F2V, KFVS) for gfx950.
//
//   k_uface   one workgroup per face, threads striding the velocity grid: KitBase's first-order
//             KFVS flux between the face's two cells (flux_kfvs.jl:615-697, called with the face's
//             normal and tangential velocities, solver_evolution.jl:872-1035) or the diffuse Maxwell
//             wall of a boundary face whose cell is cellType 2 (flux_boundary.jl:160-220), the face
//             moments reduced in FP64 and turned to the global frame (geo_general.jl:6-26).
//   k_ucell   one workgroup per stepped cell (cellType 0 or 2): the triangle step! (step_1f.jl:228-285,
//             step_2f.jl:256-326) -- wavefront 0 runs the FP64 chain (w -= sum dirc fw / area,
//             conserve_prim, tau, Shakhov heat flux), every thread relaxes its velocities.
//   k_uavg    one workgroup, the cellType-3 cells in index order: w and f the mean of the two
//             neighbours (update_boundary.jl:520-590).
//
// The wall is taken with rot = 1: the face normal of a boundary face points into the mesh, so the
// velocities with vn >= 0 come from the wall (KitBase's unstructured evolve! passes no `rot`, which
// its Maxwell-wall methods require).
#include <algorithm>

#include "kbg_device.h"
#include "kbg_types.h"

namespace kbg {

namespace {

constexpr int kUVpt = 4;  // velocities per thread
constexpr int kUMaxWaves = 16;

// heaviside (math_general.jl:13): H(0) = 1
template <class T> KBG_D T heav(T x) { return x >= T(0) ? T(1) : T(0); }
// exp in the storage precision
KBG_D double uexp(double x) { return exp(x); }
KBG_D float uexp(float x) { return expf(x); }

template <class T> __global__ __launch_bounds__(1024) void k_uface(const SolveUS* __restrict__ pS) {
    __shared__ double red[kUMaxWaves * 8];
    const SolveUS& S = *pS;
    const int f = blockIdx.x;
    if (f >= S.nface) return;
    const int kind = S.fkind[f];
    if (kind == 2) return;
    const int nv = S.nv, nf = S.nf, RS = nf * nv, tid = threadIdx.x, NT = blockDim.x;
    const double n1 = S.fgeo[f * 3], n2 = S.fgeo[f * 3 + 1], len = S.fgeo[f * 3 + 2], dt = S.dt;
    const T n1T = T(n1), n2T = T(n2);
    const T* __restrict__ U = (const T*)S.u;
    const T* __restrict__ V = (const T*)S.v;
    const T* __restrict__ WT = (const T*)S.wt;
    T* __restrict__ ff = (T*)S.ff + (size_t)f * RS;
    double fwl[4];
    if (kind == 0) {
        const T* __restrict__ fa = (const T*)S.f + (size_t)S.fc[2 * f] * RS;
        const T* __restrict__ fb = (const T*)S.f + (size_t)S.fc[2 * f + 1] * RS;
        const T dtT = T(dt), lenT = T(len);
        double s[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < kUVpt; k++) {
            const int iv = tid + NT * k;
            if (iv >= nv) continue;
            const T u = U[iv] * n1T + V[iv] * n2T, v = V[iv] * n1T - U[iv] * n2T, wq = WT[iv];
            const T d = heav(u);
            // (dt u f - dt^2 u^2 sf / 2) len with sf = 0 (no reconstruct! on these cells)
            const T h = fa[iv] * d + fb[iv] * (T(1) - d);
            const T fh = (dtT * u * h) * lenT;
            ff[iv] = fh;
            s[0] += double(wq * fh);
            s[1] += double(u * wq * fh);
            s[2] += double(v * wq * fh);
            s[3] += double((u * u + v * v) * wq * fh);
            if (nf == 2) {
                const T b = fa[nv + iv] * d + fb[nv + iv] * (T(1) - d);
                const T fbv = (dtT * u * b) * lenT;
                ff[nv + iv] = fbv;
                s[4] += double(wq * fbv);
            }
        }
        block_allreduce<5>(s, red);
        fwl[0] = s[0];
        fwl[1] = s[1];
        fwl[2] = s[2];
        fwl[3] = nf == 2 ? 0.5 * (s[3] + s[4]) : 0.5 * s[3];
    } else {
        // Maxwell wall (flux_boundary.jl:160-220), the wall state in the face's frame (geo_general.jl:130-150)
        const int c = S.fc[2 * f] >= 0 ? S.fc[2 * f] : S.fc[2 * f + 1];
        const T* __restrict__ fc = (const T*)S.f + (size_t)c * RS;
        const double* wp = S.fwall + (size_t)f * 4;
        const double bU = wp[1] * n1 + wp[2] * n2, bV = wp[2] * n1 - wp[1] * n2, bl = wp[3];
        const T Uw = T(bU), Vw = T(bV), lw = T(bl);
        double s2[2] = {0, 0};
        for (int k = 0; k < kUVpt; k++) {
            const int iv = tid + NT * k;
            if (iv >= nv) continue;
            const T u = U[iv] * n1T + V[iv] * n2T, v = V[iv] * n1T - U[iv] * n2T, wq = WT[iv];
            const T d = heav(u), cu = u - Uw, cv = v - Vw;
            s2[0] += double(wq * u * fc[iv] * (T(1) - d));
            s2[1] += double(wq * u * uexp(-lw * (cu * cu + cv * cv)) * d);
        }
        block_allreduce<2>(s2, red);
        const double SG = s2[1] * (bl / kPi);
        const double rho = -s2[0] / SG;
        const T A = T(rho * (bl / kPi)), kbw = T(S.K / (2.0 * bl)), sc = T(len * dt);
        double s[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < kUVpt; k++) {
            const int iv = tid + NT * k;
            if (iv >= nv) continue;
            const T u = U[iv] * n1T + V[iv] * n2T, v = V[iv] * n1T - U[iv] * n2T, wq = WT[iv];
            const T d = heav(u), cu = u - Uw, cv = v - Vw;
            const T H = A * uexp(-lw * (cu * cu + cv * cv));
            const T hW = H * d + fc[iv] * (T(1) - d);
            s[0] += double(wq * u * hW);
            s[1] += double(wq * (u * u) * hW);
            s[2] += double(wq * v * (hW * u));
            s[3] += double(wq * u * (hW * (u * u + v * v)));
            ff[iv] = u * hW * sc;
            if (nf == 2) {
                const T bW = (H * kbw) * d + fc[nv + iv] * (T(1) - d);
                s[4] += double(wq * u * bW);
                ff[nv + iv] = u * bW * sc;
            }
        }
        block_allreduce<5>(s, red);
        const double ld = len * dt;
        fwl[0] = s[0] * ld;
        fwl[1] = s[1] * ld;
        fwl[2] = s[2] * ld;
        fwl[3] = nf == 2 ? (0.5 * s[3] + 0.5 * s[4]) * ld : 0.5 * s[3] * ld;
    }
    if (tid == 0) {
        double* fw = S.fw + (size_t)f * 4;
        fw[0] = fwl[0];
        fw[1] = fwl[1] * n1 - fwl[2] * n2;
        fw[2] = fwl[1] * n2 + fwl[2] * n1;
        fw[3] = fwl[3];
    }
}

enum { U_RT = 0, U_INV, U_RDS, U_A, U_L, U_U, U_V, U_KB, U_PREU, U_PREV, U_AO, U_LO, U_UO, U_VO, U_KBO, U_N };

template <class T> __global__ __launch_bounds__(1024) void k_ucell(const SolveUS* __restrict__ pS, int rows) {
    __shared__ double red[kUMaxWaves * 4];
    __shared__ double cst[U_N];
    const SolveUS& S = *pS;
    const int c = blockIdx.x;
    if (c >= S.ncell) return;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int ct = S.ctype[c];
    double* rrow = rows ? S.rcell + (size_t)c * 8 : nullptr;
    if (ct != 0 && ct != 2) {
        if (rrow && tid < 8) rrow[tid] = 0.0;
        return;
    }
    const int nv = S.nv, nf = S.nf, RS = nf * nv;
    T* __restrict__ h = (T*)S.f + (size_t)c * RS;
    T* __restrict__ bd = h + nv;
    const T* __restrict__ U = (const T*)S.u;
    const T* __restrict__ V = (const T*)S.v;
    const T* __restrict__ WT = (const T*)S.wt;
    const double K = S.K, dt = S.dt;
    const bool shak = nf == 2 && S.coll == SHAKHOV;
    double po[4];
#pragma unroll
    for (int m = 0; m < 4; m++) po[m] = S.prim[(size_t)c * 4 + m];
    double q[4] = {0, 0, 0, 0};
    if (shak) {
        // heat flux of the old h, b (theory_moments.jl:567-578)
        const T Uo = T(po[1]), Vo = T(po[2]);
        for (int k = 0; k < kUVpt; k++) {
            const int iv = tid + NT * k;
            if (iv >= nv) continue;
            const T cu = U[iv] - Uo, cv = V[iv] - Vo, c2 = cu * cu + cv * cv, wq = WT[iv];
            q[0] += double(wq * cu * c2 * h[iv]);
            q[1] += double(wq * cu * bd[iv]);
            q[2] += double(wq * cv * c2 * h[iv]);
            q[3] += double(wq * cv * bd[iv]);
        }
        block_allreduce<4>(q, red);
    }
    const int* cf = S.cf + (size_t)c * 3;
    const double* dr = S.dirc + (size_t)c * 3;
    if (tid < 64) {
        const double ds = S.area[c];
        double wo[4], wn[4], pn[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            wo[m] = S.w[(size_t)c * 4 + m];
            wn[m] = wo[m] - (S.fw[(size_t)cf[0] * 4 + m] * dr[0] + S.fw[(size_t)cf[1] * 4 + m] * dr[1] +
                             S.fw[(size_t)cf[2] * 4 + m] * dr[2]) / ds;
        }
        conserve_prim<4>(wn, S.gam, pn);
        const double tau = vhs_tau(pn[0], pn[3], S.mu, S.om);
        const double r = dt / tau;
        if (tid == 0) {
            flag_health(S.health, prim_health<4>(pn));
            cst[U_RT] = r;
            cst[U_INV] = 1.0 / (1.0 + r);
            cst[U_RDS] = 1.0 / ds;
            cst[U_A] = pn[0] * (pn[3] / kPi);
            cst[U_L] = pn[3];
            cst[U_U] = pn[1];
            cst[U_V] = pn[2];
            cst[U_KB] = K / (2.0 * pn[3]);
            if (shak) {
                const double q1 = 0.5 * (q[0] + q[1]), q2 = 0.5 * (q[2] + q[3]);
                const double pre = 0.8 * (1.0 - S.Pr) * (po[3] * po[3]) / po[0];
                cst[U_PREU] = pre * q1;
                cst[U_PREV] = pre * q2;
                cst[U_AO] = po[0] * (po[3] / kPi);
                cst[U_LO] = po[3];
                cst[U_UO] = po[1];
                cst[U_VO] = po[2];
                cst[U_KBO] = K / (2.0 * po[3]);
            }
        }
        if (tid < 4) {
            const int m = tid;
            S.w[(size_t)c * 4 + m] = wn[m];
            S.prim[(size_t)c * 4 + m] = pn[m];
            if (rrow) {
                rrow[m] = (wn[m] - wo[m]) * (wn[m] - wo[m]);
                rrow[4 + m] = fabs(wn[m]);
            }
        }
    }
    __syncthreads();
    const T rT = T(cst[U_RT]), inv = T(cst[U_INV]), rds = T(cst[U_RDS]);
    const T An = T(cst[U_A]), ln = T(cst[U_L]), Un = T(cst[U_U]), Vn = T(cst[U_V]), kbn = T(cst[U_KB]);
    T preu = 0, prev = 0, Ao = 0, lo = 0, Uo = 0, Vo = 0, kbo = 0;
    if (shak) {
        preu = T(cst[U_PREU]);
        prev = T(cst[U_PREV]);
        Ao = T(cst[U_AO]);
        lo = T(cst[U_LO]);
        Uo = T(cst[U_UO]);
        Vo = T(cst[U_VO]);
        kbo = T(cst[U_KBO]);
    }
    const T d0 = T(dr[0]), d1 = T(dr[1]), d2 = T(dr[2]);
    const T* F1 = (const T*)S.ff + (size_t)cf[0] * RS;
    const T* F2 = (const T*)S.ff + (size_t)cf[1] * RS;
    const T* F3 = (const T*)S.ff + (size_t)cf[2] * RS;
    for (int k = 0; k < kUVpt; k++) {
        const int iv = tid + NT * k;
        if (iv >= nv) continue;
        const T u = U[iv], v = V[iv];
        const T cu = u - Un, cv = v - Vn;
        T MH = An * uexp(-ln * (cu * cu + cv * cv));
        T MB = MH * kbn;
        if (shak) {
            const T ou = u - Uo, ov = v - Vo, c2 = ou * ou + ov * ov;
            const T Mo = Ao * uexp(-lo * c2);
            const T gq = preu * ou + prev * ov;
            MH += gq * (T(2) * lo * c2 + T(K) - T(5)) * Mo;
            MB += gq * (T(2) * lo * c2 + T(K) - T(3)) * (Mo * kbo);
        }
        h[iv] = (h[iv] - (F1[iv] * d0 + F2[iv] * d1 + F3[iv] * d2) * rds + rT * MH) * inv;
        if (nf == 2) {
            const int o = nv + iv;
            bd[iv] = (bd[iv] - (F1[o] * d0 + F2[o] * d1 + F3[o] * d2) * rds + rT * MB) * inv;
        }
    }
}

template <class T> __global__ __launch_bounds__(1024) void k_uavg(const SolveUS* __restrict__ pS) {
    const SolveUS& S = *pS;
    const int RS = S.nf * S.nv, tid = threadIdx.x, NT = blockDim.x;
    for (int a = 0; a < S.navg; a++) {
        const int c = S.avg[a];
        int id[2] = {-1, -1}, k = 0;
        for (int j = 0; j < 3; j++) {
            const int n = S.nb[(size_t)c * 3 + j];
            if (n != -1 && k < 2) id[k++] = n;
        }
        if (k == 2) {
            if (tid < 4) S.w[(size_t)c * 4 + tid] = 0.5 * (S.w[(size_t)id[0] * 4 + tid] + S.w[(size_t)id[1] * 4 + tid]);
            T* f = (T*)S.f;
            for (int i = tid; i < RS; i += NT) f[(size_t)c * RS + i] = T(0.5) * (f[(size_t)id[0] * RS + i] + f[(size_t)id[1] * RS + i]);
            __syncthreads();
            if (tid == 0) conserve_prim<4>(S.w + (size_t)c * 4, S.gam, S.prim + (size_t)c * 4);
        }
        __syncthreads();
    }
}

// residual of update! (solver_update.jl:462-466) from the per-cell rows [(w_new - w_old)^2 | |w_new|]
__global__ __launch_bounds__(256) void k_uresid(const double* __restrict__ rows, int nrows, double* __restrict__ out) {
    __shared__ double part[256 * 8];
    for (int col = 0; col < 8; col++) {
        double s = 0.0;
        for (int r = threadIdx.x; r < nrows; r += 256) s += rows[(size_t)r * 8 + col];
        part[threadIdx.x * 8 + col] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int m = threadIdx.x;
        double a = 0.0, b = 0.0;
        for (int t = 0; t < 256; t++) {
            a += part[t * 8 + m];
            b += part[t * 8 + 4 + m];
        }
        out[m] = sqrt(a * double(nrows)) / (b + 1.e-7);
    }
}

// timestep for unstructured control volumes (solver_main.jl:400-426): tmax = max over the cells of
// ((u1 + c)^2 + (v1 + c)^2)^(1/2) / dmin, c the sound speed (theory_macro.jl:421), dmin the cell
// centre's least distance to its edges
__global__ __launch_bounds__(256) void k_ucfl(const double* __restrict__ prim, const double* __restrict__ dmin, int n, double u1,
                                              double v1, double gam, double* __restrict__ out) {
    __shared__ double part[256];
    double t = 0.0;
    for (int c = threadIdx.x; c < n; c += 256) {
        const double sos = sqrt(0.5 * gam / prim[(size_t)c * 4 + 3]);
        const double a = u1 + sos, b = v1 + sos;
        t = fmax(t, sqrt(a * a + b * b) / dmin[c]);
    }
    part[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < 256; k++) m = fmax(m, part[k]);
        out[0] = m;
    }
}

}  // namespace

hipError_t launch_uresid(const double* d_rows, int nrows, double* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_uresid, dim3(1), dim3(256), 0, st, d_rows, nrows, d_out);
    return hipGetLastError();
}

hipError_t launch_ucfl(const double* d_prim, const double* d_dmin, int n, double u1, double v1, double gam, double* d_out,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_ucfl, dim3(1), dim3(256), 0, st, d_prim, d_dmin, n, u1, v1, gam, d_out);
    return hipGetLastError();
}

hipError_t launch_uface(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, hipStream_t st) {
    if (S.nface <= 0) return hipSuccess;
    if (threads < 64 || threads > 1024 || threads % 64 || threads * kUVpt < S.nv) return hipErrorInvalidValue;
    if (fp64) hipLaunchKernelGGL(k_uface<double>, dim3(S.nface), dim3(threads), 0, st, d_S);
    else hipLaunchKernelGGL(k_uface<float>, dim3(S.nface), dim3(threads), 0, st, d_S);
    return hipGetLastError();
}

hipError_t launch_ucell(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, int rows, hipStream_t st) {
    if (S.ncell <= 0) return hipSuccess;
    if (threads < 64 || threads > 1024 || threads % 64 || threads * kUVpt < S.nv) return hipErrorInvalidValue;
    if (fp64) hipLaunchKernelGGL(k_ucell<double>, dim3(S.ncell), dim3(threads), 0, st, d_S, rows);
    else hipLaunchKernelGGL(k_ucell<float>, dim3(S.ncell), dim3(threads), 0, st, d_S, rows);
    return hipGetLastError();
}

hipError_t launch_uavg(bool fp64, const SolveUS& S, const SolveUS* d_S, hipStream_t st) {
    if (S.navg <= 0) return hipSuccess;
    if (fp64) hipLaunchKernelGGL(k_uavg<double>, dim3(1), dim3(256), 0, st, d_S);
    else hipLaunchKernelGGL(k_uavg<float>, dim3(1), dim3(256), 0, st, d_S);
    return hipGetLastError();
}

}  // namespace kbg

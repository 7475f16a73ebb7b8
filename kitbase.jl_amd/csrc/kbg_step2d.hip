// kbg_step2d.hip -- the fused 2D2F2V time step for gfx950 (structured quadrilateral mesh), batched
// over solves.
//
// One solve!-loop iteration (reconstruct! + evolve! + Maxwell walls + update!) in three launches per
// group of solves of one precision and flux:
//
//   k_face2d   one thread per interior face: the FP64 interface chain of the KCU / UGKS flux
//              (reconstructed conserved states, Gaussian moments, pdf slopes, time coefficients)
//              written as a 32-double record.  KFVS needs none.  The phase-level evolve! reads the
//              same records.
//   k_flux2d   one workgroup per face, threads striding the velocity grid: the upwind cell's slope
//              rebuilt per velocity from three neighbouring distributions (no slope array), the face
//              flux ff written once per face, its moment sums reduced in FP64 to fw.  Maxwell walls in
//              the same launch.
//   k_cell2d   one workgroup per cell: the FP64 cell chain on one wavefront, broadcast through LDS,
//              and the relaxation update of f in place.
//
// Each face flux is evaluated once (the flux is the VALU-heavy part of the step: about 100 FP32
// operations per face and velocity for UGKS); f is read by the faces through L2 and written once by
// the cells.  Cell workgroups are 64 ceil(nv / 256) threads, 4 velocities per thread; face
// workgroups 64 ceil(nv / 512) threads, 8 velocities per thread (more velocities share the per-face
// set-up); grids of up to 4096 (FP32) or 2048 (FP64) velocities take this step.
//
// Reference: evolve! solver_evolution.jl:608-679 with flux_kfvs.jl:98-135 / :654-697,
// flux_kcu.jl:149-191 / :898-967, flux_ugks.jl:121-253; reconstruct! solver_reconstruction.jl:475-620
// (reconstruct2! reconstruction.jl:48-50 on boundary rows); walls evolve_boundary.jl:52-178 ->
// flux_boundary.jl:190-220; update! solver_update.jl:193-246 + update_boundary.jl:172-306;
// step! step_2f.jl:179-249.
#include <algorithm>

#include "kbg_device.h"
#include "kbg_types.h"

namespace kbg {

namespace {

enum { SK_NONE = 0, SK_LIM = 1, SK_RIGHT = 2, SK_LEFT = 3 };

// slope kind of a cell at position p of n along a direction, and the reciprocal centre spacings
// 1 / (0.5 (d_{c-1} + d_c)), 1 / (0.5 (d_c + d_{c+1})) that k_recon2d uses
template <class T> struct SlopeC {
    int kind;
    T rL, rR;
};
template <class T> KBG_D SlopeC<T> slope_c(int interp2, const double* d, int cell, int p, int n, int step) {
    SlopeC<T> r{SK_NONE, T(0), T(0)};
    if (!interp2) return r;
    if (p > 1) r.rL = T(1.0 / (0.5 * (d[cell - step] + d[cell])));
    if (p < n) r.rR = T(1.0 / (0.5 * (d[cell] + d[cell + step])));
    r.kind = p == 1 ? SK_RIGHT : (p == n ? SK_LEFT : SK_LIM);
    return r;
}

// per-velocity slope from the three distributions around the cell (one-sided on boundary rows)
template <class T> KBG_D T slope_v(int kind, int lim, T cm, T cc, T cp, T rL, T rR) {
    const T sL = (cc - cm) * rL, sR = (cp - cc) * rR;
    const T l = limf<T>(lim, sL, sR);
    return kind == SK_LIM ? l : (kind == SK_RIGHT ? sR : (kind == SK_LEFT ? sL : T(0)));
}

// FP64 slope of the conserved variables of cell c along dir (k_recon2d's sw)
KBG_D void wslope(const Solve2D& S, const double* w, int c, int dir, double* sw) {
    const int nx = S.nx, ny = S.ny;
    const int p = dir == 0 ? c % nx + 1 : c / nx + 1;
    const int n = dir == 0 ? nx : ny;
    const int step = dir == 0 ? 1 : nx;
    const double* d = dir == 0 ? S.dx : S.dy;
    if (!S.interp2) {
        for (int m = 0; m < 4; m++) sw[m] = 0.0;
    } else if (p == 1 || p == n) {
        const int a = p == 1 ? c : c - step, e = p == 1 ? c + step : c;
        const double D = 0.5 * (d[a] + d[e]);
        for (int m = 0; m < 4; m++) sw[m] = (w[(size_t)e * 4 + m] - w[(size_t)a * 4 + m]) / D;
    } else {
        const int a = c - step, e = c + step;
        const double dL = 0.5 * (d[a] + d[c]), dR = 0.5 * (d[c] + d[e]);
        for (int m = 0; m < 4; m++)
            sw[m] = limiter<double>(S.limiter, (w[(size_t)c * 4 + m] - w[(size_t)a * 4 + m]) / dL,
                                    (w[(size_t)e * 4 + m] - w[(size_t)c * 4 + m]) / dR);
    }
}

// index of the solve that owns global item g, from the prefix array start[0..n] (start[0] = 0)
KBG_D int owner(const int* start, int n, int g) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (g >= start[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

}  // namespace

// ---- face coefficients ---------------------------------------------------------------------------

// Interior faces of each solve, x-faces (i = 2..nx) then y-faces (j = 2..ny).  INLINE_SW: the
// conserved slopes are rebuilt from w (fused step); otherwise they are read from the stored sw of
// reconstruct! (phase-level evolve!).
template <int FLUX, bool INLINE_SW>
__global__ __launch_bounds__(64) void k_face2d(const Solve2D* __restrict__ solves, const int* __restrict__ fstart, int nsolves,
                                               int nfaces1) {
    const int g = blockIdx.x * 64 + threadIdx.x;
    int si = 0, b = g;
    if (fstart) {
        if (g >= fstart[nsolves]) return;
        si = owner(fstart, nsolves, g);
        b = g - fstart[si];
    } else if (g >= nfaces1) {
        return;
    }
    const Solve2D& S = solves[si];
    const int nx = S.nx, ny = S.ny;
    const int nxf = (nx - 1) * ny;
    int dir, fi, cL, cR, pa;
    if (b < nxf) {
        dir = 0;
        const int i = b % (nx - 1) + 2, j = b / (nx - 1) + 1;
        fi = (i - 1) + (nx + 1) * (j - 1);
        cL = (i - 2) + nx * (j - 1);
        cR = cL + 1;
        pa = i - 1;
    } else {
        b -= nxf;
        dir = 1;
        const int i = b % nx + 1, j = b / nx + 2;
        fi = (i - 1) + nx * (j - 1);
        cL = (i - 1) + nx * (j - 2);
        cR = cL + nx;
        pa = j - 1;
    }
    const double* geo = (dir == 0 ? S.xgeo : S.ygeo) + (size_t)fi * 3;
    const double n1 = geo[0], n2 = geo[1];
    const double* d = dir == 0 ? S.dx : S.dy;
    const double dxL = 0.5 * d[cL], dxR = 0.5 * d[cR];
    const double* w = S.w;
    double swL[4], swR[4];
    if (INLINE_SW) {
        wslope(S, w, cL, dir, swL);
        wslope(S, w, cR, dir, swR);
    } else {
        for (int m = 0; m < 4; m++) {
            swL[m] = S.sw[(size_t)cL * 8 + dir * 4 + m];
            swR[m] = S.sw[(size_t)cR * 8 + dir * 4 + m];
        }
    }
    double wl[4], wr[4], WL[4], WR[4];
    for (int m = 0; m < 4; m++) {
        wl[m] = w[(size_t)cL * 4 + m] + dxL * swL[m];
        wr[m] = w[(size_t)cR * 4 + m] - dxR * swR[m];
    }
    local_frame4(wl, n1, n2, WL);
    local_frame4(wr, n1, n2, WR);
    double* rec = (dir == 0 ? S.xcoef : S.ycoef) + (size_t)fi * kFaceCoef;
    float* c32 = S.xc32 ? (dir == 0 ? S.xc32 : S.yc32) + (size_t)fi * kFaceC32 : nullptr;
    const double len = geo[2];
    if constexpr (FLUX == KCU) {
        KcuFace<4> F;
        kcu_face<4>(WL, WR, S.K, S.gam, S.mu, S.om, S.dt, F);
        *(KcuFace<4>*)rec = F;
        if (c32) {
            // the values k_flux2d would form from the record, with the same expressions
            for (int m = 0; m < kFaceC32; m++) c32[m] = 0.f;
            c32[0] = float(F.prim[0] * (F.prim[3] / kPi));
            c32[1] = float(F.prim[3]);
            c32[2] = float(F.prim[1]);
            c32[3] = float(F.prim[2]);
            c32[4] = float(S.K / (2.0 * F.prim[3]));
            c32[6] = float(F.Mt1 * len);
            c32[7] = float(F.Mt2 * len);
        }
    } else {
        UgksFace2 F;
        ugks_face_2d(WL, WR, S.K, S.gam, S.mu, S.om, S.dt, dxL, dxR, F);
        *(UgksFace2*)rec = F;
        if (c32) {
            const double kbd = S.K / (2.0 * F.prim[3]);
            c32[0] = float(F.prim[0] * (F.prim[3] / kPi));
            c32[1] = float(F.prim[3]);
            c32[2] = float(F.prim[1]);
            c32[3] = float(F.prim[2]);
            c32[4] = float(kbd);
            for (int m = 0; m < 5; m++) c32[5 + m] = float(F.Mt[m] * len);
            for (int m = 0; m < 4; m++) {
                c32[10 + m] = float(F.aL[m]);
                c32[14 + m] = float(F.aR[m]);
                c32[18 + m] = float(F.aT[m]);
            }
            c32[22] = float(0.5 * F.aL[3] * kbd);
            c32[23] = float(0.5 * F.aL[3] * F.Mxi2);
            c32[24] = float(0.5 * F.aR[3] * kbd);
            c32[25] = float(0.5 * F.aR[3] * F.Mxi2);
            c32[26] = float(0.5 * F.aT[3] * kbd);
            c32[27] = float(0.5 * F.aT[3] * F.Mxi2);
        }
    }
    if (c32) {
        // the two cells' slope coefficients and half widths, as k_flux2d's slope_c forms them
        const int n = dir == 0 ? nx : ny, step = dir == 0 ? 1 : nx;
        const SlopeC<float> sa = slope_c<float>(S.interp2, d, cL, pa, n, step);
        const SlopeC<float> sb = slope_c<float>(S.interp2, d, cR, pa + 1, n, step);
        c32[28] = sa.rL;
        c32[29] = sa.rR;
        c32[30] = sb.rL;
        c32[31] = sb.rR;
        c32[32] = float(0.5 * d[cL]);
        c32[33] = float(0.5 * d[cR]);
    }
}

hipError_t launch_face2d(bool inline_sw, int flux, const Solve2D* d_solves, const int* d_fstart, int nsolves, int nfaces,
                         hipStream_t st) {
    if (flux == KFVS || nfaces <= 0) return hipSuccess;
    const dim3 g((nfaces + 63) / 64), b(64);
    if (flux == KCU) {
        if (inline_sw) hipLaunchKernelGGL((k_face2d<KCU, true>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
        else hipLaunchKernelGGL((k_face2d<KCU, false>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
    } else {
        if (inline_sw) hipLaunchKernelGGL((k_face2d<UGKS, true>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
        else hipLaunchKernelGGL((k_face2d<UGKS, false>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
    }
    return hipGetLastError();
}

// ---- the fused step: face fluxes with inline slopes, then the cell update ------------------------

constexpr int kMaxWaves2 = 16;  // wavefronts per workgroup at most (1024 threads)
// velocities per thread of the face-flux kernel (a multiple of 4: its per-face set-up is shared by
// more velocities than the cell kernel's 4)
#ifndef KBG_FLUX2D_VPT32
#define KBG_FLUX2D_VPT32 8
#endif
#ifndef KBG_FLUX2D_VPT64
#define KBG_FLUX2D_VPT64 8
#endif
template <class T> constexpr int flux2d_vpt() { return sizeof(T) == 4 ? KBG_FLUX2D_VPT32 : KBG_FLUX2D_VPT64; }
template <class T> constexpr int flux2d_max_threads() { return step2d_max_threads<T>() * step2d_vpt<T>() / flux2d_vpt<T>(); }
// threads of the face-flux workgroup for a group whose cell workgroups have `threads` (4 velocities
// per thread, 64 ceil(nv / 256)): 64 ceil(threads / (16 VPT)), which covers the same velocities
constexpr int flux2d_threads(bool fp64, int threads) {
    return (threads + 16 * (fp64 ? KBG_FLUX2D_VPT64 : KBG_FLUX2D_VPT32) - 1) / (16 * (fp64 ? KBG_FLUX2D_VPT64 : KBG_FLUX2D_VPT32)) * 64;
}

// XCD-aware order: the workgroups of one XCD (ids congruent mod 8) take a contiguous run of items,
// so neighbouring faces and cells are read through the same L2
KBG_D int xcd_order() {
    const int G = gridDim.x, bx = blockIdx.x;
    const int x8 = bx & 7, q8 = G >> 3, r8 = G & 7;
    return x8 * q8 + (x8 < r8 ? x8 : r8) + (bx >> 3);
}

// Faces of every solve, x-faces (i = 1..nx+1, j = 1..ny) then y-faces (i = 1..nx, j = 1..ny+1); one
// workgroup per face, threads striding the velocity grid (VPT per thread).  An interior face builds
// the slope of its upwind cell per velocity from the three distributions around it (reconstruct!),
// the face values f_L + dx/2 s_L or f_R - dx/2 s_R, the flux ff (written to the face buffer) and its
// moment sums; the sums go to FP64 (wavefront reduce-scatter, then LDS in wavefront order) and
// wavefront 0 assembles fw in the global frame.  A wall face (bc = Maxwell) takes the wall density
// from zero mass flux in a first pass and writes the wall flux in a second.  A :fix boundary face
// keeps its stored flux.
template <class T, int FLUX>
__global__ __launch_bounds__(flux2d_max_threads<T>()) void k_flux2d(const Solve2D* __restrict__ solves, const int* __restrict__ fstart,
                                                             int nsolves, int nfaces1) {
    constexpr int VPT = flux2d_vpt<T>();
    constexpr int E = sizeof(T);
    __shared__ double red[kMaxWaves2 * 16];
    const int g = xcd_order();
    int si = 0, b = g;
    if (fstart) {
        if (g >= fstart[nsolves]) return;
        si = owner(fstart, nsolves, g);
        b = g - fstart[si];
    } else if (g >= nfaces1) {
        return;
    }
    const Solve2D& S = solves[si];
    const int nx = S.nx, ny = S.ny, nv = S.nv, RS = 2 * nv;
    const int nxf = (nx + 1) * ny;
    int dir, fi, side = -1, a = 0, c = 0, pa = 0, n, step;
    if (b < nxf) {
        dir = 0;
        fi = b;
        const int i = fi % (nx + 1) + 1, j = fi / (nx + 1) + 1;
        n = nx;
        step = 1;
        if (i == 1) side = 0, c = nx * (j - 1);
        else if (i == nx + 1) side = 1, c = (nx - 1) + nx * (j - 1);
        else a = (i - 2) + nx * (j - 1), pa = i - 1;
    } else {
        dir = 1;
        fi = b - nxf;
        const int i = fi % nx + 1, j = fi / nx + 1;
        n = ny;
        step = nx;
        if (j == 1) side = 2, c = i - 1;
        else if (j == ny + 1) side = 3, c = (i - 1) + nx * (ny - 1);
        else a = (i - 1) + nx * (j - 2), pa = j - 1;
    }
    if (side >= 0 && S.bc[side] != 1) return;  // :fix side: the stored face flux stays
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, NT = blockDim.x, NWV = NT >> 6;
    const long C = long(nx) * ny;
    const rsrc_t fin = make_rsrc(S.f, unsigned(C * RS * E));
    const long nfd = dir == 0 ? long(nx + 1) * ny : long(nx) * (ny + 1);
    const rsrc_t ffo = make_rsrc(dir == 0 ? S.xff : S.yff, unsigned(nfd * RS * E));
    double* fwout = (dir == 0 ? S.xfw : S.yfw) + (size_t)fi * 4;
    const int offf = fi * RS * E, ohb = nv * E;
    const T* __restrict__ U = (const T*)S.u;
    const T* __restrict__ V = (const T*)S.v;
    const T* __restrict__ WT = (const T*)S.wt;
    const double* geo = (dir == 0 ? S.xgeo : S.ygeo) + (size_t)fi * 3;
    const double n1 = geo[0], n2 = geo[1], len = geo[2];
    const T n1T = T(n1), n2T = T(n2);
    const double dt = S.dt, K = S.K;

    if (side < 0) {
        // ---- interior face between cells a (low side) and a + step
        const double* d = dir == 0 ? S.dx : S.dy;
        const int bc_ = a + step;
        // FP32 KCU / UGKS solves: the face's coefficients as k_face2d wrote them once per face
        const float* c32 = nullptr;
        if constexpr (sizeof(T) == 4 && FLUX != KFVS)
            if (S.xc32) c32 = (dir == 0 ? S.xc32 : S.yc32) + (size_t)fi * kFaceC32;
        T dxa, dxb;
        SlopeC<T> sa, sb;
        if (c32) {
            const int ka = !S.interp2 ? SK_NONE : (pa == 1 ? SK_RIGHT : (pa == n ? SK_LEFT : SK_LIM));
            const int kb2 = !S.interp2 ? SK_NONE : (pa + 1 == 1 ? SK_RIGHT : (pa + 1 == n ? SK_LEFT : SK_LIM));
            sa = SlopeC<T>{ka, T(c32[28]), T(c32[29])};
            sb = SlopeC<T>{kb2, T(c32[30]), T(c32[31])};
            dxa = T(c32[32]);
            dxb = T(c32[33]);
        } else {
            dxa = T(0.5 * d[a]);
            dxb = T(0.5 * d[bc_]);
            sa = slope_c<T>(S.interp2, d, a, pa, n, step);
            sb = slope_c<T>(S.interp2, d, bc_, pa + 1, n, step);
        }
        const int am = pa > 1 ? a - step : a, bp = pa + 1 < n ? bc_ + step : bc_;
        const int oam = am * RS * E, oa = a * RS * E, ob = bc_ * RS * E, obp = bp * RS * E;
        const int lim = S.limiter;
        // flux coefficients in the storage type, len folded in
        T A = 0, lam = 0, Ui = 0, Vi = 0, kb = 0, M0 = 0, M1 = 0, M2 = 0, M3 = 0, M4 = 0;
        T aL[4] = {0, 0, 0, 0}, aR[4] = {0, 0, 0, 0}, aT[4] = {0, 0, 0, 0};
        T c1L = 0, c2L = 0, c1R = 0, c2R = 0, c1T = 0, c2T = 0;
        const double* rec = (dir == 0 ? S.xcoef : S.ycoef) + (size_t)fi * kFaceCoef;
        if constexpr (FLUX == KFVS) {
            M0 = T(dt * len);
            M1 = T(0.5 * dt * dt * len);
        } else if (c32) {
            A = T(c32[0]);
            lam = T(c32[1]);
            Ui = T(c32[2]);
            Vi = T(c32[3]);
            kb = T(c32[4]);
            M0 = T(c32[5]);
            M1 = T(c32[6]);
            M2 = T(c32[7]);
            M3 = T(c32[8]);
            M4 = T(c32[9]);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                aL[m] = T(c32[10 + m]);
                aR[m] = T(c32[14 + m]);
                aT[m] = T(c32[18 + m]);
            }
            c1L = T(c32[22]);
            c2L = T(c32[23]);
            c1R = T(c32[24]);
            c2R = T(c32[25]);
            c1T = T(c32[26]);
            c2T = T(c32[27]);
        } else {
            if constexpr (FLUX == KCU) {
                const KcuFace<4>& F = *(const KcuFace<4>*)rec;
                A = T(F.prim[0] * (F.prim[3] / kPi));
                lam = T(F.prim[3]);
                Ui = T(F.prim[1]);
                Vi = T(F.prim[2]);
                kb = T(K / (2.0 * F.prim[3]));
                M1 = T(F.Mt1 * len);
                M2 = T(F.Mt2 * len);
            } else {
                const UgksFace2& F = *(const UgksFace2*)rec;
                A = T(F.prim[0] * (F.prim[3] / kPi));
                lam = T(F.prim[3]);
                Ui = T(F.prim[1]);
                Vi = T(F.prim[2]);
                const double kbd = K / (2.0 * F.prim[3]);
                kb = T(kbd);
                M0 = T(F.Mt[0] * len);
                M1 = T(F.Mt[1] * len);
                M2 = T(F.Mt[2] * len);
                M3 = T(F.Mt[3] * len);
                M4 = T(F.Mt[4] * len);
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    aL[m] = T(F.aL[m]);
                    aR[m] = T(F.aR[m]);
                    aT[m] = T(F.aT[m]);
                }
                c1L = T(0.5 * F.aL[3] * kbd);
                c2L = T(0.5 * F.aL[3] * F.Mxi2);
                c1R = T(0.5 * F.aR[3] * kbd);
                c2R = T(0.5 * F.aR[3] * F.Mxi2);
                c1T = T(0.5 * F.aT[3] * kbd);
                c2T = T(0.5 * F.aT[3] * F.Mxi2);
            }
        }
        T sm[8];
#pragma unroll
        for (int m = 0; m < 8; m++) sm[m] = T(0);
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int iv = tid + NT * k;
            if (iv < nv) {
                const T u = U[iv], v = V[iv], wq = WT[iv];
                const T x = u * n1T + v * n2T, y = v * n1T - u * n2T;
                const bool dd = x >= T(0);
                const int vo = iv * E;
                const int om = (dd ? oam : oa) + vo, oc = (dd ? oa : ob) + vo, op = (dd ? ob : obp) + vo;
                const T hm = bload((T*)0, fin, om, 0), hc = bload((T*)0, fin, oc, 0), hp = bload((T*)0, fin, op, 0);
                const T bm = bload((T*)0, fin, om, ohb), bc = bload((T*)0, fin, oc, ohb), bq = bload((T*)0, fin, op, ohb);
                const int kind = dd ? sa.kind : sb.kind;
                const T rL = dd ? sa.rL : sb.rL, rR = dd ? sa.rR : sb.rR;
                const T sh = slope_v<T>(kind, lim, hm, hc, hp, rL, rR);
                const T sg = slope_v<T>(kind, lim, bm, bc, bq, rL, rR);
                const T h = dd ? hc + sh * dxa : hc - sh * dxb;
                const T bb = dd ? bc + sg * dxa : bc - sg * dxb;
                const T q2 = x * x + y * y;
                T fh, fb;
                if constexpr (FLUX == KFVS) {
                    // (dt u f - dt^2 / 2 u^2 sf) len   flux_kfvs.jl:654-697
                    fh = x * (M0 * h - M1 * x * sh);
                    fb = x * (M0 * bb - M1 * x * sg);
                    const T p = wq * fh;
                    sm[0] += p;
                    sm[1] += x * p;
                    sm[2] += y * p;
                    sm[3] += q2 * p;
                    sm[4] += wq * fb;
                } else if constexpr (FLUX == KCU) {
                    // (Mt1 u g + Mt2 u f) len   flux_kcu.jl:898-967
                    const T cu = x - Ui, cv = y - Vi;
                    const T H = A * texp(-lam * (cu * cu + cv * cv));
                    fh = x * (M1 * H + M2 * h);
                    fb = x * (M1 * (H * kb) + M2 * bb);
                    const T p = wq * x * h;
                    sm[0] += p;
                    sm[1] += x * p;
                    sm[2] += y * p;
                    sm[3] += q2 * p;
                    sm[4] += wq * x * bb;
                } else {
                    // flux_ugks.jl:209-245 with the polynomial of the slopes factored out of H and B = H K / (2 lambda)
                    const T uu = x * x;
                    const T cu = x - Ui, cv = y - Vi;
                    const T H = A * texp(-lam * (cu * cu + cv * cv));
                    const T s0 = dd ? aL[0] : aR[0], s1 = dd ? aL[1] : aR[1], s2 = dd ? aL[2] : aR[2], s3 = dd ? aL[3] : aR[3];
                    const T c1 = dd ? c1L : c1R, c2 = dd ? c2L : c2R;
                    const T P = s0 + s1 * x + s2 * y + T(0.5) * s3 * q2;
                    const T PT = aT[0] + aT[1] * x + aT[2] * y + T(0.5) * aT[3] * q2;
                    const T GH = M0 * x + M1 * uu * (P + c1) + M2 * x * (PT + c1T);
                    const T GB = M0 * x * kb + M1 * uu * (kb * P + c2) + M2 * x * (kb * PT + c2T);
                    fh = H * GH + M3 * x * h - M4 * uu * sh;
                    fb = H * GB + M3 * x * bb - M4 * uu * sg;
                    const T p = wq * x * h, r = wq * uu * sh;
                    sm[0] += p;
                    sm[1] += r;
                    sm[2] += x * p;
                    sm[3] += x * r;
                    sm[4] += y * p;
                    sm[5] += y * r;
                    sm[6] += q2 * p + wq * x * bb;
                    sm[7] += q2 * r + wq * uu * sg;
                }
                bstore(fh, ffo, offf + vo, 0);
                bstore(fb, ffo, offf + vo, ohb);
            }
        }
        double r8[8];
#pragma unroll
        for (int m = 0; m < 8; m++) r8[m] = double(sm[m]);
        const double t = wave_reduce_scatter<8>(r8);
        if ((lane & 7) == 0) red[wave * 16 + (lane >> 3)] = t;
        __syncthreads();
        if (tid == 0) {
            // face totals in wavefront order, then fw in the face frame and back to the global one
            double r[8];
#pragma unroll
            for (int m = 0; m < 8; m++) {
                double q = red[m];
                for (int w = 1; w < NWV; w++) q += red[w * 16 + m];
                r[m] = q;
            }
            if (sizeof(T) == 4 && S.delta) {
                // the sums ran over departures: add the reference Maxwellian's sums in FP64, from its
                // moments in the global frame (x = n1 u + n2 v, y = n1 v - n2 u), and record its
                // coefficient in this face's flux for the cell update
                const double* mm = S.mmom;
                const double S1 = n1 * mm[0] + n2 * mm[1];
                const double Sx = n1 * n1 * mm[2] + 2.0 * n1 * n2 * mm[3] + n2 * n2 * mm[4];
                const double Sy = (n1 * n1 - n2 * n2) * mm[3] + n1 * n2 * (mm[4] - mm[2]);
                const double Sq = n1 * (mm[5] + mm[7]) + n2 * (mm[6] + mm[8]);
                const double Sb = n1 * mm[9] + n2 * mm[10];
                double cf;
                if constexpr (FLUX == KFVS) {
                    cf = dt * len;
                    r[0] += cf * S1;
                    r[1] += cf * Sx;
                    r[2] += cf * Sy;
                    r[3] += cf * Sq;
                    r[4] += cf * Sb;
                } else if constexpr (FLUX == KCU) {
                    cf = (*(const KcuFace<4>*)rec).Mt2 * len;
                    r[0] += S1;
                    r[1] += Sx;
                    r[2] += Sy;
                    r[3] += Sq;
                    r[4] += Sb;
                } else {
                    cf = (*(const UgksFace2*)rec).Mt[3] * len;
                    r[0] += S1;
                    r[2] += Sx;
                    r[4] += Sy;
                    r[6] += Sq + Sb;
                }
                (dir == 0 ? S.xcref : S.ycref)[fi] = cf;
            }
            double l0, l1, l2, l3;
            if constexpr (FLUX == KFVS) {
                l0 = r[0];
                l1 = r[1];
                l2 = r[2];
                l3 = 0.5 * (r[3] + r[4]);
            } else if constexpr (FLUX == KCU) {
                const KcuFace<4>& F = *(const KcuFace<4>*)rec;
                l0 = F.fw0[0] * len + F.Mt2 * r[0] * len;
                l1 = F.fw0[1] * len + F.Mt2 * r[1] * len;
                l2 = F.fw0[2] * len + F.Mt2 * r[2] * len;
                l3 = F.fw0[3] * len + F.Mt2 * 0.5 * (r[3] + r[4]) * len;
            } else {
                const UgksFace2& F = *(const UgksFace2*)rec;
                l0 = (F.fw0[0] + (F.Mt[3] * r[0] - F.Mt[4] * r[1])) * len;
                l1 = (F.fw0[1] + (F.Mt[3] * r[2] - F.Mt[4] * r[3])) * len;
                l2 = (F.fw0[2] + (F.Mt[3] * r[4] - F.Mt[4] * r[5])) * len;
                l3 = (F.fw0[3] + (F.Mt[3] * 0.5 * r[6] - F.Mt[4] * 0.5 * r[7])) * len;
            }
            // global_frame4 (geo_general.jl:6-26)
            fwout[0] = l0;
            fwout[1] = l1 * n1 - l2 * n2;
            fwout[2] = l1 * n2 + l2 * n1;
            fwout[3] = l3;
        }
        return;
    }

    // ---- Maxwell wall (flux_boundary.jl:190-220): outgoing velocities carry the cell's own h, b;
    // incoming ones the wall Maxwellian whose density rho_w gives zero mass flux
    const double* wp = (dir == 0 ? S.xwall : S.ywall) + (size_t)fi * 4;
    double bcl[4];
    local_frame4(wp, n1, n2, bcl);
    const T Uw = T(bcl[1]), Vw = T(bcl[2]), lw = T(bcl[3]);
    const T rotT = (side & 1) ? T(-1) : T(1);
    const int oc = c * RS * E;
    // FP32 departures: the cell's h, b are Mr + the stored values, and the stored wall flux leaves out
    // Mr's part x Mr len dt
    const bool del = sizeof(T) == 4 && S.delta;
    T hv[VPT], bv[VPT], ev[VPT], sm[10];
#pragma unroll
    for (int m = 0; m < 10; m++) sm[m] = T(0);
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        hv[k] = bv[k] = ev[k] = T(0);
        const int iv = tid + NT * k;
        if (iv < nv) {
            const T u = U[iv], v = V[iv], wq = WT[iv];
            const T x = u * n1T + v * n2T, y = v * n1T - u * n2T;
            const T q2 = x * x + y * y;
            T h = bload((T*)0, fin, oc + iv * E, 0), bb = bload((T*)0, fin, oc + iv * E, ohb);
            hv[k] = h;
            bv[k] = bb;
            if constexpr (sizeof(T) == 4) {
                if (del) {
                    h += S.mref[iv];
                    bb += S.mref[nv + iv];
                }
            }
            if (!(x * rotT >= T(0))) {
                sm[0] += wq * x * h;
                sm[1] += wq * (x * x) * h;
                sm[2] += wq * y * (h * x);
                sm[3] += wq * x * (h * q2);
                sm[4] += wq * x * bb;
            } else {
                const T cu = x - Uw, cv = y - Vw;
                const T e = texp(-lw * (cu * cu + cv * cv));
                ev[k] = e;
                sm[5] += wq * x * e;
                sm[6] += wq * (x * x) * e;
                sm[7] += wq * y * (e * x);
                sm[8] += wq * x * (e * q2);
            }
        }
    }
    double r16[16];
#pragma unroll
    for (int m = 0; m < 16; m++) r16[m] = m < 9 ? double(sm[m]) : 0.0;
    const double t = wave_reduce_scatter<16>(r16);
    if ((lane & 3) == 0 && (lane >> 2) < 9) red[wave * 16 + (lane >> 2)] = t;
    __syncthreads();
    double r[9];
#pragma unroll
    for (int m = 0; m < 9; m++) {
        double q = red[m];
        for (int w = 1; w < NWV; w++) q += red[w * 16 + m];
        r[m] = q;
    }
    const double rho = -r[0] / ((bcl[3] / kPi) * r[5]);
    const double Aw = rho * (bcl[3] / kPi), kbw = K / (2.0 * bcl[3]);
    const T AT = T(Aw), kbT = T(kbw), sc = T(len * dt);
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int iv = tid + NT * k;
        if (iv < nv) {
            const T u = U[iv], v = V[iv];
            const T x = u * n1T + v * n2T;
            const bool in = x * rotT >= T(0);
            const T H = AT * ev[k];
            T hW = in ? H : hv[k], bW = in ? H * kbT : bv[k];
            if constexpr (sizeof(T) == 4) {
                if (del) {
                    hW = in ? H - S.mref[iv] : hv[k];
                    bW = in ? H * kbT - S.mref[nv + iv] : bv[k];
                }
            }
            bstore(x * hW * sc, ffo, offf + iv * E, 0);
            bstore(x * bW * sc, ffo, offf + iv * E, ohb);
        }
    }
    if (tid == 0) {
        const double m0 = r[0] + Aw * r[5], m1 = r[1] + Aw * r[6], m2 = r[2] + Aw * r[7], m3 = r[3] + Aw * r[8];
        const double m4 = r[4] + Aw * kbw * r[5];
        const double l1 = m1 * len * dt, l2 = m2 * len * dt;
        fwout[0] = m0 * len * dt;
        fwout[1] = l1 * n1 - l2 * n2;
        fwout[2] = l1 * n2 + l2 * n1;
        fwout[3] = (0.5 * m3 + 0.5 * m4) * len * dt;
        if (del) (dir == 0 ? S.xcref : S.ycref)[fi] = len * dt;
    }
}

// update! (solver_update.jl:193-246, update_boundary.jl:172-306; step! step_2f.jl:179-249) of every
// stepped cell, in place, one workgroup per cell: Shakhov heat flux of the old distribution, then
// wavefront 0 runs the FP64 chain (w += sum fw / ds, conserve_prim, tau) and broadcasts it, then
// f = (f + (ff_L - ff_R + ff_D - ff_U) / ds + (dt / tau) M) / (1 + dt / tau).  Cells that are not
// stepped only clear their residual row.
enum { C_RT = 0, C_INV, C_RDS, C_AN, C_LN, C_UN, C_VN, C_KBN, C_PREU, C_PREV, C_AO, C_LO, C_UO, C_VO, C_KBO, C_CU, C_CV, C_N };

template <class T>
__global__ __launch_bounds__(step2d_max_threads<T>()) void k_cell2d(const Solve2D* __restrict__ solves, const int* __restrict__ wg0,
                                                             int nsolves, int ncells1, int rows) {
    constexpr int VPT = step2d_vpt<T>();
    __shared__ double red[kMaxWaves2 * 4];
    __shared__ double cst[C_N];
    const int g = xcd_order();
    int si = 0, c = g;
    if (wg0) {
        if (g >= wg0[nsolves]) return;
        si = owner(wg0, nsolves, g);
        c = g - wg0[si];
    } else if (g >= ncells1) {
        return;
    }
    const Solve2D& S = solves[si];
    const int nx = S.nx, ny = S.ny, nv = S.nv;
    const long RS = 2L * nv;
    const int i = c % nx + 1, j = c / nx + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, NT = blockDim.x, NWV = NT >> 6;
    double* rrow = (rows && S.rcell) ? S.rcell + (size_t)c * 16 : nullptr;
    const bool inner = i > 1 && i < nx && j > 1 && j < ny;
    const bool mid = i > 1 && i < nx;
    const bool stepped = inner || (i == 1 && S.bc[0] != 0) || (i == nx && S.bc[1] != 0) || (j == 1 && mid && S.bc[2] != 0) ||
                         (j == ny && mid && S.bc[3] != 0);
    if (!stepped) {
        if (rrow && tid < 16) rrow[tid] = 0.0;
        return;
    }
    T* __restrict__ h = (T*)S.f + (size_t)c * RS;
    T* __restrict__ bd = h + nv;
    const T* __restrict__ U = (const T*)S.u;
    const T* __restrict__ V = (const T*)S.v;
    const T* __restrict__ WT = (const T*)S.wt;
    const double K = S.K, dt = S.dt;
    const bool shak = S.coll == SHAKHOV;
    const bool del = sizeof(T) == 4 && S.delta;
    double po[4];
#pragma unroll
    for (int m = 0; m < 4; m++) po[m] = S.prim[(size_t)c * 4 + m];
    if (shak) {
        // heat flux of the old distribution (step_2f.jl:199-213)
        const T Uo = T(po[1]), Vo = T(po[2]);
        T a[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int iv = tid + NT * k;
            if (iv < nv) {
                const T cu = U[iv] - Uo, cv = V[iv] - Vo, c2 = cu * cu + cv * cv, wq = WT[iv];
                T hh = h[iv], bb = bd[iv];
                if constexpr (sizeof(T) == 4) {
                    if (del) {
                        hh += S.mref[iv];
                        bb += S.mref[nv + iv];
                    }
                }
                a[0] += wq * cu * c2 * hh;
                a[1] += wq * cu * bb;
                a[2] += wq * cv * c2 * hh;
                a[3] += wq * cv * bb;
            }
        }
        double r4[4] = {double(a[0]), double(a[1]), double(a[2]), double(a[3])};
        const double t = wave_reduce_scatter<4>(r4);
        if ((lane & 15) == 0) red[wave * 4 + (lane >> 4)] = t;
        __syncthreads();
    }
    const int L = (i - 1) + (nx + 1) * (j - 1), R = L + 1;
    const int D = (i - 1) + nx * (j - 1), Up = D + nx;
    if (wave == 0) {
        const double *fwL = S.xfw + (size_t)L * 4, *fwR = S.xfw + (size_t)R * 4;
        const double *fwD = S.yfw + (size_t)D * 4, *fwU = S.yfw + (size_t)Up * 4;
        const double ds = S.dx[c] * S.dy[c];
        double wo[4], wn[4], pn[4];
        for (int m = 0; m < 4; m++) {
            wo[m] = S.w[(size_t)c * 4 + m];
            wn[m] = wo[m] + (fwL[m] - fwR[m] + fwD[m] - fwU[m]) / ds;
        }
        conserve_prim<4>(wn, S.gam, pn);
        const double tau = vhs_tau(pn[0], pn[3], S.mu, S.om);
        const double r = dt / tau;
        if (lane == 0) {
            flag_health(S.health, prim_health<4>(pn));
            cst[C_RT] = r;
            cst[C_INV] = 1.0 / (1.0 + r);
            cst[C_RDS] = 1.0 / ds;
            cst[C_AN] = pn[0] * (pn[3] / kPi);
            cst[C_LN] = pn[3];
            cst[C_UN] = pn[1];
            cst[C_VN] = pn[2];
            cst[C_KBN] = K / (2.0 * pn[3]);
            if (del) {
                // the reference Maxwellian's share of the four face fluxes: sum of +-c (n1 u + n2 v) / ds
                const double *gL = S.xgeo + (size_t)L * 3, *gR = S.xgeo + (size_t)R * 3;
                const double *gD = S.ygeo + (size_t)D * 3, *gU = S.ygeo + (size_t)Up * 3;
                const double cL = S.xcref[L], cR = S.xcref[R], cD = S.ycref[D], cU = S.ycref[Up];
                cst[C_CU] = (cL * gL[0] - cR * gR[0] + cD * gD[0] - cU * gU[0]) / ds;
                cst[C_CV] = (cL * gL[1] - cR * gR[1] + cD * gD[1] - cU * gU[1]) / ds;
            }
            if (shak) {
                double q[4];
                for (int m = 0; m < 4; m++) {
                    double s = red[m];
                    for (int w = 1; w < NWV; w++) s += red[w * 4 + m];
                    q[m] = s;
                }
                const double q1 = 0.5 * (q[0] + q[1]), q2 = 0.5 * (q[2] + q[3]);
                const double pre = 0.8 * (1.0 - S.Pr) * (po[3] * po[3]) / po[0];
                cst[C_PREU] = pre * q1;
                cst[C_PREV] = pre * q2;
                cst[C_AO] = po[0] * (po[3] / kPi);
                cst[C_LO] = po[3];
                cst[C_UO] = po[1];
                cst[C_VO] = po[2];
                cst[C_KBO] = K / (2.0 * po[3]);
            }
        }
        if (lane < 4) {
            const int m = lane;
            S.w[(size_t)c * 4 + m] = wn[m];
            S.prim[(size_t)c * 4 + m] = pn[m];
            if (rrow) {
                rrow[(inner ? 0 : 2) * 4 + m] = (wn[m] - wo[m]) * (wn[m] - wo[m]);
                rrow[(inner ? 1 : 3) * 4 + m] = fabs(wn[m]);
                rrow[(inner ? 2 : 0) * 4 + m] = 0.0;
                rrow[(inner ? 3 : 1) * 4 + m] = 0.0;
            }
        }
    }
    __syncthreads();
    const T rT = T(cst[C_RT]), inv = T(cst[C_INV]), rds = T(cst[C_RDS]);
    const T Cu = del ? T(cst[C_CU]) : T(0), Cv = del ? T(cst[C_CV]) : T(0);
    const T An = T(cst[C_AN]), ln = T(cst[C_LN]), Un = T(cst[C_UN]), Vn = T(cst[C_VN]), kbn = T(cst[C_KBN]);
    T preu = 0, prev = 0, Ao = 0, lo = 0, Uo = 0, Vo = 0, kbo = 0;
    if (shak) {
        preu = T(cst[C_PREU]);
        prev = T(cst[C_PREV]);
        Ao = T(cst[C_AO]);
        lo = T(cst[C_LO]);
        Uo = T(cst[C_UO]);
        Vo = T(cst[C_VO]);
        kbo = T(cst[C_KBO]);
    }
    const T* fL = (const T*)S.xff + (size_t)L * RS;
    const T* fR = (const T*)S.xff + (size_t)R * RS;
    const T* fD = (const T*)S.yff + (size_t)D * RS;
    const T* fU = (const T*)S.yff + (size_t)Up * RS;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int iv = tid + NT * k;
        if (iv < nv) {
            const T u = U[iv], v = V[iv];
            const T cu = u - Un, cv = v - Vn;
            T MH = An * texp(-ln * (cu * cu + cv * cv));
            T MB = MH * kbn;
            if (shak) {
                const T ou = u - Uo, ov = v - Vo, c2 = ou * ou + ov * ov;
                const T Mo = Ao * texp(-lo * c2);
                const T gq = preu * ou + prev * ov;
                MH += gq * (T(2) * lo * c2 + T(K) - T(5)) * Mo;
                MB += gq * (T(2) * lo * c2 + T(K) - T(3)) * (Mo * kbo);
            }
            const int o = nv + iv;
            if constexpr (sizeof(T) == 4) {
                if (del) {
                    // departures: (d + flux of the departures + Mr's flux share + rT (M - Mr)) / (1 + rT)
                    const T mh = S.mref[iv], mb = S.mref[o];
                    const T jr = Cu * u + Cv * v;
                    h[iv] = (h[iv] + (fL[iv] - fR[iv] + fD[iv] - fU[iv]) * rds + jr * mh + rT * (MH - mh)) * inv;
                    bd[iv] = (bd[iv] + (fL[o] - fR[o] + fD[o] - fU[o]) * rds + jr * mb + rT * (MB - mb)) * inv;
                    continue;
                }
            }
            h[iv] = (h[iv] + (fL[iv] - fR[iv] + fD[iv] - fU[iv]) * rds + rT * MH) * inv;
            bd[iv] = (bd[iv] + (fL[o] - fR[o] + fD[o] - fU[o]) * rds + rT * MB) * inv;
        }
    }
}

// departure <-> full values of an FP32 fused solve, in place (grid-stride over cells and faces)
__global__ __launch_bounds__(256) void k_delta2d(const Solve2D* __restrict__ pS, int mode) {
    const Solve2D& S = *pS;
    const int nv = S.nv;
    const long RS = 2L * nv, C = long(S.nx) * S.ny, FX = long(S.nx + 1) * S.ny, FY = long(S.nx) * (S.ny + 1);
    const float* __restrict__ mr = S.mref;
    const float* __restrict__ U = (const float*)S.u;
    const float* __restrict__ V = (const float*)S.v;
    const long stride = long(gridDim.x) * blockDim.x, t0 = long(blockIdx.x) * blockDim.x + threadIdx.x;
    float* f = (float*)S.f;
    for (long e = t0; e < C * RS; e += stride) {
        const float m = mr[e % RS];
        f[e] = mode == 2 ? f[e] + m : f[e] - m;
    }
    for (int dir = 0; dir < 2; dir++) {
        const long nfc = dir == 0 ? FX : FY;
        float* ff = (float*)(dir == 0 ? S.xff : S.yff);
        double* cr = dir == 0 ? S.xcref : S.ycref;
        const double* geo = dir == 0 ? S.xgeo : S.ygeo;
        if (mode == 2) {
            for (long e = t0; e < nfc * RS; e += stride) {
                const long fc = e / RS;
                const double c = cr[fc];
                if (c != 0.0) {
                    const int k = int(e % RS), iv = k % nv;
                    ff[e] += float(c * (geo[fc * 3] * U[iv] + geo[fc * 3 + 1] * V[iv])) * mr[k];
                }
            }
        } else {
            // the face fluxes hold full values (as the phase kernels left them): no Mr part is missing
            for (long e = t0; e < nfc; e += stride) cr[e] = 0.0;
        }
    }
}

hipError_t launch_delta2d(const Solve2D& S, const Solve2D* d_S, int mode, hipStream_t st) {
    if (!S.delta || (mode != 2 && mode != 3)) return hipErrorInvalidValue;
    const long n = long(S.nx + 1) * (S.ny + 1) * 2 * S.nv;
    const int blocks = int(std::min<long>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_delta2d, dim3(blocks), dim3(256), 0, st, d_S, mode);
    return hipGetLastError();
}

hipError_t launch_flux2d(bool fp64, int flux, int threads, const Solve2D* d_solves, const int* d_fstart, int nsolves, int nfaces,
                         hipStream_t st) {
    if (nfaces <= 0) return hipSuccess;
    if (threads < 64 || threads > (fp64 ? step2d_max_threads<double>() : step2d_max_threads<float>()) || threads % 64)
        return hipErrorInvalidValue;
    const dim3 g(nfaces), b(flux2d_threads(fp64, threads));
    if (fp64) {
        if (flux == KFVS) hipLaunchKernelGGL((k_flux2d<double, KFVS>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
        else if (flux == KCU) hipLaunchKernelGGL((k_flux2d<double, KCU>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
        else hipLaunchKernelGGL((k_flux2d<double, UGKS>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
    } else {
        if (flux == KFVS) hipLaunchKernelGGL((k_flux2d<float, KFVS>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
        else if (flux == KCU) hipLaunchKernelGGL((k_flux2d<float, KCU>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
        else hipLaunchKernelGGL((k_flux2d<float, UGKS>), g, b, 0, st, d_solves, d_fstart, nsolves, nfaces);
    }
    return hipGetLastError();
}

hipError_t launch_cell2d(bool fp64, int threads, const Solve2D* d_solves, const int* d_wg0, int nsolves, int ncells, int rows,
                         hipStream_t st) {
    if (ncells <= 0) return hipSuccess;
    if (threads < 64 || threads > (fp64 ? step2d_max_threads<double>() : step2d_max_threads<float>()) || threads % 64)
        return hipErrorInvalidValue;
    const dim3 g(ncells), b(threads);
    if (fp64) hipLaunchKernelGGL(k_cell2d<double>, g, b, 0, st, d_solves, d_wg0, nsolves, ncells, rows);
    else hipLaunchKernelGGL(k_cell2d<float>, g, b, 0, st, d_solves, d_wg0, nsolves, ncells, rows);
    return hipGetLastError();
}

}  // namespace kbg

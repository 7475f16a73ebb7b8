// kbg_us.hip -- the unstructured-mesh DVM on triangles (2D1F2V and 2D2F2V, KFVS) for gfx950.
//
//   k_ustep   the fused time step: one wavefront per cell, lanes striding the velocity grid.  A
//             stepped cell (cellType 0 or 2) evaluates the fluxes of its three faces itself --
//             KitBase's KFVS flux between the face's two cells at interpOrder 1, where the cells'
//             slopes are zero and the face values are the cell values (flux_kfvs.jl:615-697,
//             with the face's normal and tangential velocities, solver_evolution.jl:872-1035) or the
//             diffuse Maxwell wall of a boundary face (flux_boundary.jl:160-220), the face moments
//             reduced in FP64 and turned to the global frame (geo_general.jl:6-26) -- and applies the
//             triangle step! (step_1f.jl:228-285, step_2f.jl:256-326): the FP64 chain (w -= sum
//             dirc fw / area, conserve_prim, tau, Shakhov heat flux), then every lane relaxes its
//             velocities.  f is read from one buffer and written to the other, so the face fluxes
//             never go to memory: per step f is read from HBM (own cell and the three neighbours,
//             the neighbours and the second pass mostly from L2) and written once.  Cells that are
//             not stepped copy f.
//   k_uface   evolve! alone: the same face fluxes, stored to fw / ff (one workgroup per face).
//   k_ucell   update! alone: the same triangle step from the stored fluxes, in place.
//   k_uavg    one workgroup, the cellType-3 cells in index order: w and f the mean of the two
//             neighbours (update_boundary.jl:520-590).
//
// k_ustep and the k_uface + k_ucell pair share the flux and cell code below; they sum the face
// moments in different orders, so a fused step and an evolve! + update! pair agree to rounding.
// k_ustep takes its cells in an XCD-aware order: the hardware deals workgroups to the 8 XCDs in
// turn, so workgroup b steps the four cells of group (b % 8) * chunk + b / 8 and each XCD walks a
// contiguous run of cells, whose neighbours' f it finds in its own L2.
//
// The wall is taken with rot = 1: the face normal of a boundary face points into the mesh, so the
// velocities with vn >= 0 come from the wall (KitBase's unstructured evolve! passes no `rot`, which
// its Maxwell-wall methods require).
#include <algorithm>

#include "kbg_device.h"
#include "kbg_types.h"

namespace kbg {

namespace {

constexpr int kUVpt = 4;  // velocities per thread of k_uface / k_ucell (at most 1024 threads)
constexpr int kUMaxWaves = 16;
constexpr int kXcd = 8;

// heaviside (math_general.jl:13): H(0) = 1
template <class T> KBG_D T heav(T x) { return x >= T(0) ? T(1) : T(0); }
// exp in the storage precision
KBG_D double uexp(double x) { return exp(x); }
KBG_D float uexp(float x) { return expf(x); }

template <class T> KBG_D gptr<T> fbuf(const SolveUS& S, int par) { return to_global((T*)(((S.cur + par) & 1) ? S.f1 : S.f)); }

// A face as the flux code sees it: normal, length, kind, its two cells (-1 outside) and, for a
// Maxwell wall, the wall state in the face's frame (geo_general.jl:130-150)
struct UFace {
    double n1, n2, len, bU, bV, bl;
    int kind, a, o;
};

KBG_D UFace uface_at(const SolveUS& S, int f) {
    UFace F;
    F.n1 = to_global(S.fgeo)[f * 3];
    F.n2 = to_global(S.fgeo)[f * 3 + 1];
    F.len = to_global(S.fgeo)[f * 3 + 2];
    F.kind = to_global(S.fkind)[f];
    F.a = to_global(S.fc)[2 * f];
    F.o = to_global(S.fc)[2 * f + 1];
    F.bU = F.bV = F.bl = 0.0;
    if (F.kind == 1) {
        const gptr<const double> wp = to_global(S.fwall) + (size_t)f * 4;
        F.bU = wp[1] * F.n1 + wp[2] * F.n2;
        F.bV = wp[2] * F.n1 - wp[1] * F.n2;
        F.bl = wp[3];
    }
    return F;
}

// One velocity of an interior face (flux_kfvs.jl:615-697 as the unstructured evolve! calls it: no
// slope argument, so (dt u f - dt^2 u^2 sf / 2) len = dt u f len; at interpOrder 1 the face values
// are the cell values): the upwind h (b) of the first cell (ha) or the second (ho), its flux and the
// five moment terms.
template <class T>
KBG_D void kfvs_vel(T u, T v, T wq, T ha, T ho, T ba, T bo, bool two, T dtT, T lenT, T& fh, T& fb, double (&s)[5]) {
    const T d = heav(u);
    const T h = ha * d + ho * (T(1) - d);
    fh = (dtT * u * h) * lenT;
    s[0] += double(wq * fh);
    s[1] += double(u * wq * fh);
    s[2] += double(v * wq * fh);
    s[3] += double((u * u + v * v) * wq * fh);
    if (two) {
        const T bb = ba * d + bo * (T(1) - d);
        fb = (dtT * u * bb) * lenT;
        s[4] += double(wq * fb);
    }
}

// Maxwell wall (flux_boundary.jl:160-220), one velocity: the two sums that fix the wall density
// (SF, SG before their constant factors) ...
template <class T> KBG_D void wall_pre_vel(T u, T v, T wq, T hc, T Uw, T Vw, T lw, double& sf, double& sg) {
    const T d = heav(u), cu = u - Uw, cv = v - Vw;
    sf += double(wq * u * hc * (T(1) - d));
    sg += double(wq * u * uexp(-lw * (cu * cu + cv * cv)) * d);
}

// ... and, with the wall Maxwellian's amplitude A, the flux of h (b) and the five moment terms
template <class T>
KBG_D void wall_vel(T u, T v, T wq, T hc, T bc, bool two, T Uw, T Vw, T lw, T A, T kbw, T sc, T& fh, T& fb, double (&s)[5]) {
    const T d = heav(u), cu = u - Uw, cv = v - Vw;
    const T H = A * uexp(-lw * (cu * cu + cv * cv));
    const T hW = H * d + hc * (T(1) - d);
    s[0] += double(wq * u * hW);
    s[1] += double(wq * (u * u) * hW);
    s[2] += double(wq * v * (hW * u));
    s[3] += double(wq * u * (hW * (u * u + v * v)));
    fh = u * hW * sc;
    if (two) {
        const T bW = (H * kbw) * d + bc * (T(1) - d);
        s[4] += double(wq * u * bW);
        fb = u * bW * sc;
    }
}

// the wall density factor from the reduced sums: A = rho_w lambda_w / pi with rho_w = -SF / SG
KBG_D double wall_amp(double sf, double sg, double bl) {
    const double SG = sg * (bl / kPi);
    const double rho = -sf / SG;
    return rho * (bl / kPi);
}

// the face's moments in the global frame (geo_general.jl:6-26) from its reduced sums
KBG_D void face_fw(const UFace& F, const double (&s)[5], bool two, double dt, double (&fw)[4]) {
    double fwl[4];
    if (F.kind == 0) {
        fwl[0] = s[0];
        fwl[1] = s[1];
        fwl[2] = s[2];
        fwl[3] = two ? 0.5 * (s[3] + s[4]) : 0.5 * s[3];
    } else {
        const double ld = F.len * dt;
        fwl[0] = s[0] * ld;
        fwl[1] = s[1] * ld;
        fwl[2] = s[2] * ld;
        fwl[3] = two ? (0.5 * s[3] + 0.5 * s[4]) * ld : 0.5 * s[3] * ld;
    }
    fw[0] = fwl[0];
    fw[1] = fwl[1] * F.n1 - fwl[2] * F.n2;
    fw[2] = fwl[1] * F.n2 + fwl[2] * F.n1;
    fw[3] = fwl[3];
}

// this thread's velocity k: index, node, weight, and the node in face F's frame
template <class T> struct UVel {
    int iv;
    T u, v, wq;
};
template <class T> KBG_D UVel<T> uvel(const SolveUS& S, int k) {
    const int iv = threadIdx.x + blockDim.x * k;
    UVel<T> q{iv, T(0), T(0), T(0)};
    if (iv < S.nv) {
        q.u = to_global((const T*)S.u)[iv];
        q.v = to_global((const T*)S.v)[iv];
        q.wq = to_global((const T*)S.wt)[iv];
    }
    return q;
}
template <class T> KBG_D void rotate(const UFace& F, const UVel<T>& q, T& un, T& vt) {
    const T n1T = T(F.n1), n2T = T(F.n2);
    un = q.u * n1T + q.v * n2T;
    vt = q.v * n1T - q.u * n2T;
}

// The cell's constants of the triangle step! (step_1f.jl:228-285, step_2f.jl:256-326): the FP64
// chain from the old prim po, the reduced heat-flux sums q and the dirc-weighted sum of the faces'
// moments dw = sum_j dirc_j fw_j -- w -= dw / area, conserve_prim, tau, the new Maxwellian and, for
// Shakhov, the old one with its heat-flux prefactors.  Lanes with writer set store w, prim and the
// residual row (rrow, or NULL) for moment `lane` (0..3); every caller's lanes have read w before.
struct UCell {
    double rT, inv, rds, A, l, U, V, kb, preu, prev, Ao, lo, Uo, Vo, kbo;
};

KBG_D UCell cell_chain(const SolveUS& S, int c, const double (&po)[4], const double (&q)[4], const double (&dw)[4], bool writer,
                       int lane, gptr<double> rrow) {
    const double K = S.K, ds = to_global(S.area)[c];
    double wo[4], wn[4], pn[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        wo[m] = to_global(S.w)[(size_t)c * 4 + m];
        wn[m] = wo[m] - dw[m] / ds;
    }
    conserve_prim<4>(wn, S.gam, pn);
    const double tau = vhs_tau(pn[0], pn[3], S.mu, S.om);
    const double r = S.dt / tau;
    UCell k;
    k.rT = r;
    k.inv = 1.0 / (1.0 + r);
    k.rds = 1.0 / ds;
    k.A = pn[0] * (pn[3] / kPi);
    k.l = pn[3];
    k.U = pn[1];
    k.V = pn[2];
    k.kb = K / (2.0 * pn[3]);
    k.preu = k.prev = k.Ao = k.lo = k.Uo = k.Vo = k.kbo = 0.0;
    if (S.nf == 2 && S.coll == SHAKHOV) {
        const double q1 = 0.5 * (q[0] + q[1]), q2 = 0.5 * (q[2] + q[3]);
        const double pre = 0.8 * (1.0 - S.Pr) * (po[3] * po[3]) / po[0];
        k.preu = pre * q1;
        k.prev = pre * q2;
        k.Ao = po[0] * (po[3] / kPi);
        k.lo = po[3];
        k.Uo = po[1];
        k.Vo = po[2];
        k.kbo = K / (2.0 * po[3]);
    }
    if (writer && lane == 0) flag_health(S.health, prim_health<4>(pn));
    if (writer && lane < 4) {
        const int m = lane;
        to_global(S.w)[(size_t)c * 4 + m] = wn[m];
        to_global(S.prim)[(size_t)c * 4 + m] = pn[m];
        if (rrow) {
            rrow[m] = (wn[m] - wo[m]) * (wn[m] - wo[m]);
            rrow[4 + m] = fabs(wn[m]);
        }
    }
    return k;
}

// the step's constants in the storage precision
template <class T> struct UCellT {
    T rT, inv, rds, A, l, U, V, kb, preu, prev, Ao, lo, Uo, Vo, kbo;
};
template <class T> KBG_D UCellT<T> to_t(const UCell& k) {
    return UCellT<T>{T(k.rT), T(k.inv), T(k.rds), T(k.A),  T(k.l),  T(k.U),  T(k.V),  T(k.kb),
                     T(k.preu), T(k.prev), T(k.Ao), T(k.lo), T(k.Uo), T(k.Vo), T(k.kbo)};
}

// one velocity of the relaxation: h (b) from its old value, the dirc-weighted face fluxes
// fsum = sum_j dirc_j F_j and the (Shakhov-corrected) new Maxwellian
template <class T>
KBG_D void relax_vel(const UCellT<T>& k, bool shak, T K, T u, T v, T h, T b, T fsh, T fsb, T& hn, T& bn) {
    const T cu = u - k.U, cv = v - k.V;
    T MH = k.A * uexp(-k.l * (cu * cu + cv * cv));
    T MB = MH * k.kb;
    if (shak) {
        const T ou = u - k.Uo, ov = v - k.Vo, c2 = ou * ou + ov * ov;
        const T Mo = k.Ao * uexp(-k.lo * c2);
        const T gq = k.preu * ou + k.prev * ov;
        MH += gq * (T(2) * k.lo * c2 + K - T(5)) * Mo;
        MB += gq * (T(2) * k.lo * c2 + K - T(3)) * (Mo * k.kbo);
    }
    hn = (h - fsh * k.rds + k.rT * MH) * k.inv;
    bn = (b - fsb * k.rds + k.rT * MB) * k.inv;
}

// Shakhov heat-flux terms of one velocity's old h, b (theory_moments.jl:567-578)
template <class T> KBG_D void heat_vel(T u, T v, T wq, T h, T b, T Uo, T Vo, double (&q)[4]) {
    const T cu = u - Uo, cv = v - Vo, c2 = cu * cu + cv * cv;
    q[0] += double(wq * cu * c2 * h);
    q[1] += double(wq * cu * b);
    q[2] += double(wq * cv * c2 * h);
    q[3] += double(wq * cv * b);
}

// this thread's h, b of one cell (0 past nv)
template <class T, int VPT> KBG_D void load_hb(const SolveUS& S, gptr<const T> p, T (&h)[VPT], T (&b)[VPT]) {
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int iv = threadIdx.x + blockDim.x * k;
        h[k] = iv < S.nv ? p[iv] : T(0);
        b[k] = (S.nf == 2 && iv < S.nv) ? p[S.nv + iv] : T(0);
    }
}

KBG_D bool stepped(int ct) { return ct == 0 || ct == 2; }

template <class T, int VPT> __global__ __launch_bounds__(1024) void k_uface(const SolveUS* __restrict__ pS) {
    __shared__ double red[kUMaxWaves * 8];
    const SolveUS& S = *pS;
    const int f = blockIdx.x;
    if (f >= S.nface || to_global(S.fkind)[f] == 2) return;
    const UFace F = uface_at(S, f);
    const int nv = S.nv, RS = S.nf * nv;
    const bool two = S.nf == 2;
    const gptr<const T> fcur = fbuf<T>(S, 0);
    const gptr<T> ff = to_global((T*)S.ff) + (size_t)f * RS;
    double s[5] = {0, 0, 0, 0, 0};
    if (F.kind == 0) {
        T ha[VPT], ba[VPT], ho[VPT], bo[VPT];
        load_hb(S, fcur + (size_t)F.a * RS, ha, ba);
        load_hb(S, fcur + (size_t)F.o * RS, ho, bo);
        const T dtT = T(S.dt), lenT = T(F.len);
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const UVel<T> x = uvel<T>(S, k);
            if (x.iv >= nv) continue;
            T un, vt, fh, fb;
            rotate(F, x, un, vt);
            kfvs_vel(un, vt, x.wq, ha[k], ho[k], ba[k], bo[k], two, dtT, lenT, fh, fb, s);
            ff[x.iv] = fh;
            if (two) ff[nv + x.iv] = fb;
        }
    } else {
        T hc[VPT], bc[VPT];
        load_hb(S, fcur + (size_t)(F.a >= 0 ? F.a : F.o) * RS, hc, bc);
        const T Uw = T(F.bU), Vw = T(F.bV), lw = T(F.bl);
        double s2[2] = {0, 0};
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const UVel<T> x = uvel<T>(S, k);
            if (x.iv >= nv) continue;
            T un, vt;
            rotate(F, x, un, vt);
            wall_pre_vel(un, vt, x.wq, hc[k], Uw, Vw, lw, s2[0], s2[1]);
        }
        block_allreduce<2>(s2, red);
        const T A = T(wall_amp(s2[0], s2[1], F.bl)), kbw = T(S.K / (2.0 * F.bl)), sc = T(F.len * S.dt);
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const UVel<T> x = uvel<T>(S, k);
            if (x.iv >= nv) continue;
            T un, vt, fh, fb;
            rotate(F, x, un, vt);
            wall_vel(un, vt, x.wq, hc[k], bc[k], two, Uw, Vw, lw, A, kbw, sc, fh, fb, s);
            ff[x.iv] = fh;
            if (two) ff[nv + x.iv] = fb;
        }
    }
    block_allreduce<5>(s, red);
    if (threadIdx.x == 0) {
        double fw[4];
        face_fw(F, s, two, S.dt, fw);
#pragma unroll
        for (int m = 0; m < 4; m++) to_global(S.fw)[(size_t)f * 4 + m] = fw[m];
    }
}

template <class T, int VPT> __global__ __launch_bounds__(1024) void k_ucell(const SolveUS* __restrict__ pS, int rows) {
    __shared__ double red[kUMaxWaves * 4];
    __shared__ double cst[sizeof(UCell) / sizeof(double)];
    const SolveUS& S = *pS;
    const int c = blockIdx.x;
    if (c >= S.ncell) return;
    const gptr<double> rrow = rows ? to_global(S.rcell) + (size_t)c * 8 : nullptr;
    if (!stepped(to_global(S.ctype)[c])) {
        if (rrow && threadIdx.x < 8) rrow[threadIdx.x] = 0.0;
        return;
    }
    const int nv = S.nv, RS = S.nf * nv, tid = threadIdx.x;
    const bool two = S.nf == 2, shak = two && S.coll == SHAKHOV;
    const gptr<T> hp = fbuf<T>(S, 0) + (size_t)c * RS;
    double po[4];
#pragma unroll
    for (int m = 0; m < 4; m++) po[m] = to_global(S.prim)[(size_t)c * 4 + m];
    T h[VPT], b[VPT], Fh[3][VPT], Fb[3][VPT];
    double fw[3][4];
    load_hb(S, hp, h, b);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int f = to_global(S.cf)[(size_t)c * 3 + j];
        load_hb(S, to_global((const T*)S.ff) + (size_t)f * RS, Fh[j], Fb[j]);
#pragma unroll
        for (int m = 0; m < 4; m++) fw[j][m] = to_global(S.fw)[(size_t)f * 4 + m];
    }
    const gptr<const double> dr = to_global(S.dirc) + (size_t)c * 3;
    double dw[4];
#pragma unroll
    for (int m = 0; m < 4; m++) dw[m] = fw[0][m] * dr[0] + fw[1][m] * dr[1] + fw[2][m] * dr[2];
    double q[4] = {0, 0, 0, 0};
    if (shak) {
        const T Uo = T(po[1]), Vo = T(po[2]);
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const UVel<T> x = uvel<T>(S, k);
            if (x.iv < nv) heat_vel(x.u, x.v, x.wq, h[k], b[k], Uo, Vo, q);
        }
        block_allreduce<4>(q, red);
    }
    // wavefront 0 runs the chain (every thread has read prim above; the reduction's barrier, or for
    // BGK the threads' not reading po again, orders it before the writes)
    if (tid < 64) {
        const UCell k = cell_chain(S, c, po, q, dw, true, tid, rrow);
        if (tid == 0) *reinterpret_cast<UCell*>(cst) = k;
    }
    __syncthreads();
    const UCellT<T> kt = to_t<T>(*reinterpret_cast<const UCell*>(cst));
    const T d0 = T(dr[0]), d1 = T(dr[1]), d2 = T(dr[2]);
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const UVel<T> x = uvel<T>(S, k);
        if (x.iv >= nv) continue;
        T hn, bn;
        relax_vel(kt, shak, T(S.K), x.u, x.v, h[k], b[k], Fh[0][k] * d0 + Fh[1][k] * d1 + Fh[2][k] * d2,
                  Fb[0][k] * d0 + Fb[1][k] * d1 + Fb[2][k] * d2, hn, bn);
        hp[x.iv] = hn;
        if (two) hp[nv + x.iv] = bn;
    }
}

// The fused step, one wavefront per cell (W cells per workgroup, no workgroup barrier: every LDS
// word is written and read back by the same lane).  A face's moments in the global frame are
// moments of its flux with the face-frame velocities turned back: with |n|^2 = n1^2 + n2^2 (KitBase's
// unit_normal leaves it slightly below 1), (un n1 - vt n2, un n2 + vt n1) = |n|^2 (u, v) and
// un^2 + vt^2 = |n|^2 (u^2 + v^2), so
//   fw = (sum w F, |n|^2 sum w u F, |n|^2 sum w v F, (|n|^2 sum w (u^2 + v^2) F + sum w Fb) / 2).
// Pass 1 takes one face at a time (after a first pass that fixes the wall densities when the cell
// has a Maxwell wall): it sums these four moments in FP64 (the Fb part of the last over 2 |n|^2, so
// that one factor |n|^2 turns the sum into fw) with a wavefront reduction per face, and adds the
// face's dirc-weighted flux to the relaxation's flux terms sum_j dirc_j F_j, sum_j dirc_j Fb_j in
// LDS (stored kind-2 fluxes included); the heat flux follows in a loop of its own.  The two cells of a face sum its
// moments over the same velocities in the same order and reduce them the same way, so they see the
// same fw and the face's mass, momentum and energy leave one cell exactly as they enter the other.
// Every lane then runs the FP64 chain (the same values in all lanes) and pass 2 relaxes from the
// cell's f and the LDS terms.  The moments are summed in another order than k_uface's local-frame
// sums: a fused step and an evolve! + update! pair agree to rounding.
template <class T> __global__ __launch_bounds__(256) void k_ustep(const SolveUS* __restrict__ pS, int par, int rows, int chunk, int W) {
    extern __shared__ double ldsd[];
    const SolveUS& S = *pS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the wavefront's cell, a wavefront-uniform value: the cell's and its faces' data go to scalar registers
    const int c = __builtin_amdgcn_readfirstlane((int(blockIdx.x % kXcd) * chunk + int(blockIdx.x / kXcd)) * W + wave);
    if (c >= S.ncell) return;  // per wavefront: no workgroup barrier follows
    const int nv = S.nv, RS = S.nf * nv;
    const bool two = S.nf == 2, shak = two && S.coll == SHAKHOV;
    const gptr<const T> fin = fbuf<T>(S, par);
    const gptr<T> fout = fbuf<T>(S, par + 1);
    const gptr<const T> fc = fin + (size_t)c * RS;
    const gptr<T> fo = fout + (size_t)c * RS;
    const gptr<double> rrow = rows ? to_global(S.rcell) + (size_t)c * 8 : nullptr;
    if (!stepped(to_global(S.ctype)[c])) {
        // carried over (a cellType-3 cell is then averaged by k_uavg)
        for (int i = lane; i < RS; i += 64) fo[i] = fc[i];
        if (rrow && lane < 8) rrow[lane] = 0.0;
        return;
    }
    T* g = reinterpret_cast<T*>(ldsd) + (size_t)wave * RS;  // [G0 | B0] of this wavefront's velocities
    double po[4];
#pragma unroll
    for (int m = 0; m < 4; m++) po[m] = to_global(S.prim)[(size_t)c * 4 + m];
    const gptr<const double> dr = to_global(S.dirc) + (size_t)c * 3;
    const gptr<const T> U = to_global((const T*)S.u);
    const gptr<const T> V = to_global((const T*)S.v);
    const gptr<const T> WT = to_global((const T*)S.wt);
    UFace F[3];
    gptr<const T> nb[3];  // the neighbour's f (interior face) or the stored flux (kind 2)
    bool first[3];        // this cell is the face's first cell
    bool wall = false;
    T d[3];       // dirc_j
    double hb[3];  // 1 / (2 |n_j|^2): the b part of the energy sum, taken out of |n|^2 again below
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int f = to_global(S.cf)[(size_t)c * 3 + j];
        F[j] = uface_at(S, f);
        first[j] = F[j].a == c;
        nb[j] = F[j].kind == 0 ? fin + (size_t)(first[j] ? F[j].o : F[j].a) * RS
                               : (F[j].kind == 2 ? to_global((const T*)S.ff) + (size_t)f * RS : nullptr);
        wall = wall || F[j].kind == 1;
        d[j] = T(dr[j]);
        hb[j] = 0.5 / (F[j].n1 * F[j].n1 + F[j].n2 * F[j].n2);
    }
    T A[3] = {T(0), T(0), T(0)};
    if (wall) {
        double s2[6] = {0, 0, 0, 0, 0, 0};
        for (int iv = lane; iv < nv; iv += 64) {
            const UVel<T> x{iv, U[iv], V[iv], WT[iv]};
            const T h = fc[iv];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if (F[j].kind != 1) continue;
                T un, vt;
                rotate(F[j], x, un, vt);
                wall_pre_vel(un, vt, x.wq, h, T(F[j].bU), T(F[j].bV), T(F[j].bl), s2[2 * j], s2[2 * j + 1]);
            }
        }
        wave_allreduce<6>(s2);
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (F[j].kind == 1) A[j] = T(wall_amp(s2[2 * j], s2[2 * j + 1], F[j].bl));
    }
    const T dtT = T(S.dt), KT = T(S.K);
    // pass 1, one face at a time: its moments [sum w F, sum w u F, sum w v F, sum w (u^2 + v^2) F / 2 +
    // sum w Fb / (2 |n|^2)] in FP64 and the flux terms sum_j dirc_j F_j, sum_j dirc_j Fb_j in LDS
    double fw[3][4];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (F[j].kind == 2) {
            for (int iv = lane; iv < nv; iv += 64) {
                const T gh = nb[j][iv] * d[j];
                g[iv] = j == 0 ? gh : g[iv] + gh;
                if (two) {
                    const T gb = nb[j][nv + iv] * d[j];
                    g[nv + iv] = j == 0 ? gb : g[nv + iv] + gb;
                }
            }
#pragma unroll
            for (int m = 0; m < 4; m++) fw[j][m] = to_global(S.fw)[(size_t)to_global(S.cf)[(size_t)c * 3 + j] * 4 + m];
            continue;
        }
        double a4[4] = {0, 0, 0, 0};
        const T lenT = T(F[j].len), Uw = T(F[j].bU), Vw = T(F[j].bV), lw = T(F[j].bl), kbw = T(S.K / (2.0 * F[j].bl)),
                sc = T(F[j].len * S.dt);
        for (int iv = lane; iv < nv; iv += 64) {
            const UVel<T> x{iv, U[iv], V[iv], WT[iv]};
            const T h = fc[iv], b = two ? fc[nv + iv] : T(0);
            double t[5] = {0, 0, 0, 0, 0};  // local-frame sums: unused here
            T un, vt, fh = T(0), fb = T(0);
            rotate(F[j], x, un, vt);
            if (F[j].kind == 0) {
                const T hn = nb[j][iv], bn = two ? nb[j][nv + iv] : T(0);
                kfvs_vel(un, vt, x.wq, first[j] ? h : hn, first[j] ? hn : h, first[j] ? b : bn, first[j] ? bn : b, two, dtT, lenT,
                         fh, fb, t);
            } else {
                wall_vel(un, vt, x.wq, h, b, two, Uw, Vw, lw, A[j], kbw, sc, fh, fb, t);
            }
            const double e = double(x.wq * fh);
            a4[0] += e;
            a4[1] += double(x.u) * e;
            a4[2] += double(x.v) * e;
            a4[3] += 0.5 * double(x.u * x.u + x.v * x.v) * e;
            if (two) a4[3] += hb[j] * double(x.wq * fb);
            const T gh = fh * d[j];
            g[iv] = j == 0 ? gh : g[iv] + gh;
            if (two) {
                const T gb = fb * d[j];
                g[nv + iv] = j == 0 ? gb : g[nv + iv] + gb;
            }
        }
        wave_allreduce<4>(a4);
        const double nn = F[j].n1 * F[j].n1 + F[j].n2 * F[j].n2;
        fw[j][0] = a4[0];
        fw[j][1] = nn * a4[1];
        fw[j][2] = nn * a4[2];
        fw[j][3] = nn * a4[3];
    }
    double q[4] = {0, 0, 0, 0};
    if (shak) {
        const T Uo = T(po[1]), Vo = T(po[2]);
        for (int iv = lane; iv < nv; iv += 64) heat_vel(U[iv], V[iv], WT[iv], fc[iv], fc[nv + iv], Uo, Vo, q);
        wave_allreduce<4>(q);
    }
    double dw[4];
#pragma unroll
    for (int m = 0; m < 4; m++) dw[m] = fw[0][m] * dr[0] + fw[1][m] * dr[1] + fw[2][m] * dr[2];
    const UCellT<T> kt = to_t<T>(cell_chain(S, c, po, q, dw, true, lane, rrow));
    // pass 2: the relaxation
    for (int iv = lane; iv < nv; iv += 64) {
        const T h = fc[iv], b = two ? fc[nv + iv] : T(0);
        T hn, bn;
        relax_vel(kt, shak, KT, U[iv], V[iv], h, b, g[iv], two ? g[nv + iv] : T(0), hn, bn);
        fo[iv] = hn;
        if (two) fo[nv + iv] = bn;
    }
}

template <class T> __global__ __launch_bounds__(1024) void k_uavg(const SolveUS* __restrict__ pS, int par) {
    const SolveUS& S = *pS;
    const int RS = S.nf * S.nv, tid = threadIdx.x, NT = blockDim.x;
    const gptr<T> f = fbuf<T>(S, par);
    for (int a = 0; a < S.navg; a++) {
        const int c = to_global(S.avg)[a];
        int id[2] = {-1, -1}, k = 0;
        for (int j = 0; j < 3; j++) {
            const int n = to_global(S.nb)[(size_t)c * 3 + j];
            if (n != -1 && k < 2) id[k++] = n;
        }
        if (k == 2) {
            if (tid < 4) to_global(S.w)[(size_t)c * 4 + tid] = 0.5 * (to_global(S.w)[(size_t)id[0] * 4 + tid] + to_global(S.w)[(size_t)id[1] * 4 + tid]);
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

bool threads_ok(int threads, const SolveUS& S) {
    return threads >= 64 && threads <= 1024 && threads % 64 == 0 && threads * kUVpt >= S.nv;
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

int ustep_threads(int nv) { return 64 * ((nv + 64 * kUVpt - 1) / (64 * kUVpt)); }

hipError_t launch_uface(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, hipStream_t st) {
    if (S.nface <= 0) return hipSuccess;
    if (!threads_ok(threads, S)) return hipErrorInvalidValue;
    if (fp64) hipLaunchKernelGGL((k_uface<double, kUVpt>), dim3(S.nface), dim3(threads), 0, st, d_S);
    else hipLaunchKernelGGL((k_uface<float, kUVpt>), dim3(S.nface), dim3(threads), 0, st, d_S);
    return hipGetLastError();
}

hipError_t launch_ucell(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, int rows, hipStream_t st) {
    if (S.ncell <= 0) return hipSuccess;
    if (!threads_ok(threads, S)) return hipErrorInvalidValue;
    if (fp64) hipLaunchKernelGGL((k_ucell<double, kUVpt>), dim3(S.ncell), dim3(threads), 0, st, d_S, rows);
    else hipLaunchKernelGGL((k_ucell<float, kUVpt>), dim3(S.ncell), dim3(threads), 0, st, d_S, rows);
    return hipGetLastError();
}

hipError_t launch_ustep(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, int par, int rows, hipStream_t st) {
    (void)threads;  // one wavefront per cell
    if (S.ncell <= 0) return hipSuccess;
    // W cells per workgroup, W the most of 4, 2, 1 whose LDS flux terms (W nf nv values) fit 64 KiB;
    // workgroup b -> cell group (b % 8) chunk + b / 8 covers [0, 8 chunk) once (each XCD a contiguous
    // run); cells past ncell exit
    const size_t per = size_t(S.nf) * S.nv * (fp64 ? 8 : 4);
    const int W = 4 * per <= 65536 ? 4 : (2 * per <= 65536 ? 2 : 1);
    if (per > 65536) return hipErrorInvalidValue;
    const int groups = (S.ncell + W - 1) / W, chunk = (groups + kXcd - 1) / kXcd;
    const dim3 grid(unsigned(chunk) * kXcd), block(64 * W);
    if (fp64) hipLaunchKernelGGL(k_ustep<double>, grid, block, W * per, st, d_S, par, rows, chunk, W);
    else hipLaunchKernelGGL(k_ustep<float>, grid, block, W * per, st, d_S, par, rows, chunk, W);
    return hipGetLastError();
}

hipError_t launch_uavg(bool fp64, const SolveUS& S, const SolveUS* d_S, int par, hipStream_t st) {
    if (S.navg <= 0) return hipSuccess;
    if (fp64) hipLaunchKernelGGL(k_uavg<double>, dim3(1), dim3(256), 0, st, d_S, par);
    else hipLaunchKernelGGL(k_uavg<float>, dim3(1), dim3(256), 0, st, d_S, par);
    return hipGetLastError();
}

}  // namespace kbg

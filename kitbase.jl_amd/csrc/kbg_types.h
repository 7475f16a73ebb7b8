// kbg_types.h -- device-visible solve descriptors and launch entry points of the kbgpu kernels.
//
// A solve's state lives in HBM for its whole life.  Distributions are stored in the solve's
// precision (FP32 or FP64) with explicit strides: element (cell k, distribution q, velocity v) of f
// is at k fcs + q fqs + v fvs.  Solves stepped by the fused sweep (1D, one velocity dimension) are
// cell-fastest, f[q][v][np] with each row padded to np cells; the others are velocity-contiguous
// per cell, f[cell][nf][nv] (h then b).  Slopes sf and face fluxes ff are always [cell|face][nf][nv].
// Conserved and primitive variables are FP64 [cell][nw].  1D cells are indexed by storage
// k = i - 1 + ng for the physical 1-based index i in [1-ng, nx+ng]; 1D faces by k - 1 for the
// physical face k in [1, nx+1] (between cells k-1 and k).  2D cells (i, j) by (i-1) + nx (j-1);
// x-faces (i, j), i in [1, nx+1], by (i-1) + (nx+1)(j-1); y-faces (i, j), j in [1, ny+1], by
// (i-1) + nx (j-1).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace kbg {

struct Solve1D {
    int nx, ng, nv, nf, vdim, nw;
    int flux, coll, limiter, interp2;
    int upd_lo, upd_hi;       // physical cells stepped by update!
    int slope_lo, slope_hi;   // physical cells whose slopes reconstruct! computes
    int face_lo, face_hi;     // physical faces evolve! computes from the two neighbours
    int faceL_buf, faceR_buf; // physical face 1 / nx+1 comes from the boundary-face buffers
    int wallL, wallR;         // Maxwell wall at face 1 / nx+1 (2F only)
    int cur;                  // index of the current f / w buffer
    int swept;                // f is cell-fastest and stepped by k_sweep1d
    int delta;                // FP32 swept: f buffers hold f - M~(prim) while the fused sweep owns them
    long np;                  // padded cells per velocity row (swept layout)
    long fcs, fqs, fvs;       // strides of f: cell, distribution, velocity
    double K, gam, mu, om, Pr, dt;
    const double* dx;                              // [nx + 2 ng]
    const void* u;                                 // [nv] velocity nodes (storage type)
    const void* v;                                 // [nv] (3V only)
    const void* wv;                                // [nv] (3V only)
    const void* wt;                                // [nv] quadrature weights
    double* w[2];                                  // [N][nw] ping-pong conserved variables
    double* prim;                                  // [N][nw]
    void* f[2];                                    // [N][nf][nv] ping-pong distributions
    double* sw;                                    // [N][nw]          (phase-exact path)
    void* sf;                                      // [N][nf][nv]      (phase-exact path)
    double* fw;                                    // [nx+1][nw]       (phase-exact path)
    void* ff;                                      // [nx+1][nf][nv]   (phase-exact path)
    double* bfw;                                   // [2][nw]  faces 1 and nx+1 (fused path)
    void* bff;                                     // [2][nf][nv]
    const double* bcprim;                          // [2][nw]  wall primitive vectors
    const void* vtab;                              // [nv + 8] per-velocity constants (swept path)
    int* health;                                   // failure flags (kbg_device.h prim_health)
    const double* fsmq;                            // collision :fsm: Q of the cells fsm_c0 .. [ncell][nv]
    int fsm_c0;                                    // first physical cell of fsmq
};

struct Solve2D {
    int nx, ny, nv;
    int flux, coll, limiter, interp2;
    int bc[4];                                     // left, right, down, up: 0 fix, 1 Maxwell
    double K, gam, mu, om, Pr, dt;
    const double* dx;                              // [nx*ny]
    const double* dy;
    const double* xgeo;                            // [(nx+1)*ny][3]  n1, n2, len
    const double* ygeo;                            // [nx*(ny+1)][3]
    const double* xwall;                           // [(nx+1)*ny][4]  wall prim (global frame)
    const double* ywall;                           // [nx*(ny+1)][4]
    const void* u;                                 // [nv]
    const void* v;
    const void* wt;
    double* w;                                     // [nx*ny][4]
    double* prim;
    void* f;                                       // [nx*ny][2][nv]
    double* sw;                                    // [nx*ny][2 dirs][4]          (phase-exact path)
    void* sf;                                      // [nx*ny][2 dirs][2][nv]      (phase-exact path)
    double* xfw;                                   // [(nx+1)*ny][4]
    void* xff;                                     // [(nx+1)*ny][2][nv]
    double* yfw;
    void* yff;
    double* xcoef;                                 // [(nx+1)*ny][kFaceCoef] KCU / UGKS face chain records
    double* ycoef;                                 // [nx*(ny+1)][kFaceCoef]
    float* xc32;                                   // FP32 KCU / UGKS solves: [(nx+1)*ny][kFaceC32] the face
    float* yc32;                                   //   flux's FP32 coefficients (k_face2d writes, k_flux2d reads)
    double* rcell;                                 // [nx*ny][16] residual rows
    int* health;                                   // failure flags (kbg_device.h prim_health)
    // FP32 solves of the fused step store f - Mr, Mr a per-solve reference Maxwellian (h and b), and
    // the face fluxes of the fused step without Mr's part c x Mr, where x is the normal velocity and c
    // the face's coefficient of the distribution in its flux (0 for a stored :fix face)
    int delta;
    const float* mref;                             // [2][nv] Mr (h, b) as stored in FP32
    double mmom[11];                               // sum w u^a v^b Mr_h: (a,b) = 10 01 20 11 02 30 21 12 03; Mr_b: 10 01
    double* xcref;                                 // [(nx+1)*ny] c of each x-face
    double* ycref;                                 // [nx*(ny+1)]
};

// An unstructured triangle solve (2D1F2V / 2D2F2V, KFVS): faces with their two cells, normal (from
// the first cell to the second, into the mesh on a boundary face; KitBase's unit_normal) and
// length; cells with their three faces, dirc signs and area (solver_evolution.jl:872-1035,
// solver_update.jl:359-471).
struct SolveUS {
    int ncell, nface, nv, nf, coll, navg;
    double K, gam, mu, om, Pr, dt;
    const int* fc;        // [nface][2] cells (-1 outside)
    const int* fkind;     // [nface] 0 KFVS, 1 Maxwell wall, 2 stored flux
    const double* fgeo;   // [nface][3] n1, n2, len
    const double* fwall;  // [nface][4] wall prim, global frame
    const int* cf;        // [ncell][3] faces
    const double* dirc;   // [ncell][3]
    const double* area;   // [ncell]
    const int* ctype;     // [ncell] KitBase cellType (0, 2 stepped)
    const int* nb;        // [ncell][3] neighbours (-1 outside)
    const int* avg;       // [navg] cellType-3 cells in index order
    const void* u;        // [nv]
    const void* v;
    const void* wt;
    double* w;            // [ncell][4]
    double* prim;
    void* f;              // [ncell][nf][nv], buffer 0
    void* f1;             // buffer 1: the fused step writes the buffer it does not read
    int cur;              // buffer holding the state at the start of a call (kernels add their parity)
    double* fw;           // [nface][4] global frame
    void* ff;             // [nface][nf][nv]
    double* rcell;        // [ncell][8] residual rows (sum of squared change, sum of |w|)
    int* health;
};

// doubles per face-coefficient record (a UgksFace2 or a KcuFace<4>, padded)
constexpr int kFaceCoef = 32;
// FP32 coefficient record of an interior face: A, lambda, U, V, K / (2 lambda), Mt[0..4] len,
// aL[4], aR[4], aT[4], the six aX[3] products of the UGKS flux, the reciprocal centre spacings of
// the two cells (rL, rR each) and their half widths
constexpr int kFaceC32 = 40;
// Threads per workgroup of the fused 2D step (one face or cell per workgroup): 64 ceil(nv / 256), 4
// velocities per thread, at most 1024 in FP32 and 512 in FP64 (the register budget of an FP64
// thread); 0 when the velocity grid is larger (the phase kernels then step the solve)
template <class T> constexpr int step2d_vpt() { return 4; }
template <class T> constexpr int step2d_max_threads() { return sizeof(T) == 4 ? 1024 : 512; }
constexpr int step2d_threads(bool fp64, int nv) {
    return (nv + 255) / 256 * 64 <= (fp64 ? 512 : 1024) ? (nv + 255) / 256 * 64 : 0;
}

// One unit of work of the fused 1D sweep: cells [c0, c1] (physical, at most 60) of solve `solve`,
// updated by nw wavefronts (1, 2, 4 or 8) that split the velocity grid evenly in multiples of 8.
struct WorkItem {
    int solve, c0, c1, nw;
};
// A workgroup's share of a sweep launch: items item0 .. item0 + count - 1, all with the same nw,
// packed side by side (count nw <= the launch's wavefronts per workgroup).
struct Pack {
    int item0, count, nw, pad_;
};

// ---- launchers (kbg_sweep.h, kbg_1d.hip, kbg_2d.hip) --------------------------------------------------------
// All return hipSuccess or the launch error; `fp64` selects the storage type.
// nstep: time steps per launch, 1, or 2 for KFVS solves without boundary-face buffers (items of at most
// kSweepOut2 = 56 cells); par counts launches, not steps
hipError_t launch_sweep1d(bool fp64, int nf, int flux, int wg_waves, const Solve1D* d_solves, const WorkItem* d_items,
                          const Pack* d_packs, int npacks, int par, double* d_rpart, int nstep, hipStream_t st);
// velocities per wavefront of the sweep for a storage type / distribution count (a multiple of 8),
// and the largest velocity grid it takes (8 wavefronts)
#ifndef KBG_SWEEP_VW32
#define KBG_SWEEP_VW32 64
#endif
constexpr int kSweepMaxWaves = 8;
constexpr int sweep_vw_raw(bool fp64, int nf) { return (fp64 ? KBG_SWEEP_VW32 / 2 : KBG_SWEEP_VW32) / nf; }
constexpr int sweep_vw(bool fp64, int nf) { return sweep_vw_raw(fp64, nf) < 8 ? 8 : sweep_vw_raw(fp64, nf); }
constexpr int sweep_max_nv(bool fp64, int nf) { return kSweepMaxWaves * sweep_vw(fp64, nf); }
// wavefronts of a sweep item: the smallest power of two whose wavefronts hold nv velocities
constexpr int sweep_nw(bool fp64, int nf, int nv) {
    return (nv + sweep_vw(fp64, nf) - 1) / sweep_vw(fp64, nf) <= 1 ? 1
         : (nv + sweep_vw(fp64, nf) - 1) / sweep_vw(fp64, nf) <= 2 ? 2
         : (nv + sweep_vw(fp64, nf) - 1) / sweep_vw(fp64, nf) <= 4 ? 4 : 8;
}
// velocities per wavefront of an item (a multiple of 8, at most sweep_vw)
constexpr int sweep_item_vw(int nv, int nw) { return ((nv + nw - 1) / nw + 7) / 8 * 8; }
// Conversions of an FP32 swept solve's f buffer `buf` between its stored departure f - M~(prim) and f:
// mode 0: FP64 staging (the swept layout) -> departure; 1: departure -> FP64 staging (exact, f = M~ + d);
// 2: departure -> f in place (FP32); 3: f -> departure in place
hipError_t launch_delta1d(const Solve1D& S, const Solve1D* d_S, double* stage, int mode, int buf, hipStream_t st);
hipError_t launch_recon1d(bool fp64, const Solve1D& S, const Solve1D* d_S, hipStream_t st);
hipError_t launch_evolve1d(bool fp64, const Solve1D& S, const Solve1D* d_S, hipStream_t st);
hipError_t launch_wall1d(bool fp64, const Solve1D& S, const Solve1D* d_S, int fused, int par, hipStream_t st);
hipError_t launch_update1d(bool fp64, const Solve1D& S, const Solve1D* d_S, double* d_rcell, hipStream_t st);

// The fast spectral collision of a 1D1F3V solve (kbg_fsm.hip): kernel modes and work buffers on the
// device, hipFFT plans for the cells c0 .. c0 + ncell - 1
struct FsmPlan {
    int nu = 0, nv = 0, nw = 0, M = 0, ncell = 0, c0 = 0;
    double kn = 0.0;
    bool f32 = false;           // transforms, spectra and modes in FP32 (FP32 storage), else FP64
    void* spec = nullptr;       // complex [ncell][n3]
    void* work = nullptr;       // complex [ncell][2 M(M-1) + 2][n3]
    void* phi = nullptr;        // real [M(M-1)][n3]
    void* psi = nullptr;        // real [M(M-1)][n3]
    void* phipsi = nullptr;     // real [n3]
    double* Q = nullptr;        // [ncell][n3]
    // cubic grids of side cube_n (fsm_cube_size, else 0) take the in-house forward transforms, with
    // twiddles tw[cube_n] = exp(-2 pi i j / cube_n) in the plan's precision: FP32 in LDS per direction
    // pair (products part [ncell][M(M-1) + 1][n3], no work), FP64 through work in three passes
    int cube_n = 0;
    float* part = nullptr;
    void* tw = nullptr;
    struct hipfftHandle_t* plan_b = nullptr;
    struct hipfftHandle_t* plan_f = nullptr;
    bool planned = false;
};
// Q of the plan's cells from the f buffer f (cell stride fcs, storage index of cell c0 = k0)
hipError_t fsm_collision(const FsmPlan& P, bool fp64, const void* f, long fcs, int k0, hipStream_t st);
hipError_t fsm_plan_create(FsmPlan& P);
// side of the cubic velocity grids whose forward transforms are the library's own (8, 12, 16, 18, 24), else 0
int fsm_cube_size(int nu, int nv, int nw);
void fsm_plan_destroy(FsmPlan& P);
hipError_t launch_resid1d(const Solve1D& S, const double* d_rcell, int nrows, int boundary_split, double* d_res,
                          hipStream_t st);

// Conversions of an FP32 fused 2D solve between departures (f - Mr, face fluxes without c x Mr) and
// the full values, in place: mode 2 departure -> full, mode 3 full -> departure (the face fluxes are
// then taken as full, c = 0)
hipError_t launch_delta2d(const Solve2D& S, const Solve2D* d_S, int mode, hipStream_t st);
// unstructured triangles: face fluxes (evolve!), cell update (update!) and the cellType-3 means
// unstructured kernels; par is the step's buffer parity relative to S.cur, threads that of
// ustep_threads(nv) (the kernels' velocities per thread follow from nv)
int ustep_threads(int nv);
hipError_t launch_uface(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, hipStream_t st);
hipError_t launch_ucell(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, int rows, hipStream_t st);
hipError_t launch_ustep(bool fp64, int threads, const SolveUS& S, const SolveUS* d_S, int par, int rows, hipStream_t st);
hipError_t launch_uavg(bool fp64, const SolveUS& S, const SolveUS* d_S, int par, hipStream_t st);
hipError_t launch_uresid(const double* d_rows, int nrows, double* d_out, hipStream_t st);
hipError_t launch_ucfl(const double* d_prim, const double* d_dmin, int n, double u1, double v1, double gam, double* d_out,
                       hipStream_t st);
hipError_t launch_recon2d(bool fp64, const Solve2D& S, const Solve2D* d_S, hipStream_t st);
hipError_t launch_evolve2d(bool fp64, const Solve2D& S, const Solve2D* d_S, hipStream_t st);
hipError_t launch_update2d(bool fp64, const Solve2D& S, const Solve2D* d_S, double* d_rcell, hipStream_t st);
hipError_t launch_resid2d(const Solve2D& S, const double* d_rcell, double* d_res, hipStream_t st);
// max over cells k0 .. k0+n-1 of the CFL rate (kbg_2d.hip k_cfl); d_dy NULL in 1D
hipError_t launch_cfl(const double* d_prim, int nw, int k0, int n, const double* d_dx, const double* d_dy, double u1, double v1,
                      double gam, double* d_out, hipStream_t st);
// Face-chain records of the interior faces (KCU / UGKS; nothing for KFVS).  d_fstart: prefix sums of
// interior-face counts over nsolves solves, or NULL for the single solve d_solves[0] with nfaces faces.
// inline_sw rebuilds the conserved slopes from w instead of reading the stored sw.
hipError_t launch_face2d(bool inline_sw, int flux, const Solve2D* d_solves, const int* d_fstart, int nsolves, int nfaces,
                         hipStream_t st);
// Fused 2D step, workgroups of `threads` (step2d_threads of every solve of the launch): face fluxes of
// all faces (d_fstart: prefix sums of (nx+1) ny + nx (ny+1) per solve), then the cell update of all
// cells (d_wg0: prefix sums of nx ny).  rows: write each solve's residual rows.
hipError_t launch_flux2d(bool fp64, int flux, int threads, const Solve2D* d_solves, const int* d_fstart, int nsolves, int nfaces,
                         hipStream_t st);
hipError_t launch_cell2d(bool fp64, int threads, const Solve2D* d_solves, const int* d_wg0, int nsolves, int ncells, int rows,
                         hipStream_t st);

}  // namespace kbg

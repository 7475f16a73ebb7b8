// kbg_types.h -- device-visible solve descriptors and launch entry points of the kbgpu kernels.
//
// A solve's state lives in HBM for its whole life.  Distributions are velocity-contiguous per cell:
// f[cell][nf][nv] (h then b for two distributions), stored in the solve's precision (FP32 or FP64).
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
    double* sw;                                    // [nx*ny][2 dirs][4]
    void* sf;                                      // [nx*ny][2 dirs][2][nv]
    double* xfw;                                   // [(nx+1)*ny][4]
    void* xff;                                     // [(nx+1)*ny][2][nv]
    double* yfw;
    void* yff;
};

// One unit of work of the fused 1D kernel: cells [c0, c1] (physical) of solve `solve`, marched by
// one wavefront whose lanes hold vpt velocities each.
struct WorkItem {
    int solve, c0, c1, vpt;
};

// ---- launchers (kbg_march.hip, kbg_1d.hip, kbg_2d.hip) --------------------------------------------------------
// All return hipSuccess or the launch error; `fp64` selects the storage type.
hipError_t launch_march1d(bool fp64, int nf, int flux, int vpt, const Solve1D* d_solves, const WorkItem* d_items,
                          int nitems, int grid, int par, double* d_rpart, int seg, hipStream_t st);
hipError_t launch_recon1d(bool fp64, const Solve1D& S, const Solve1D* d_S, hipStream_t st);
hipError_t launch_evolve1d(bool fp64, const Solve1D& S, const Solve1D* d_S, hipStream_t st);
hipError_t launch_wall1d(bool fp64, const Solve1D& S, const Solve1D* d_S, int fused, int par, hipStream_t st);
hipError_t launch_update1d(bool fp64, const Solve1D& S, const Solve1D* d_S, double* d_rcell, hipStream_t st);
hipError_t launch_resid1d(const Solve1D& S, const double* d_rcell, int nrows, int boundary_split, double* d_res,
                          hipStream_t st);

hipError_t launch_recon2d(bool fp64, const Solve2D& S, const Solve2D* d_S, hipStream_t st);
hipError_t launch_evolve2d(bool fp64, const Solve2D& S, const Solve2D* d_S, hipStream_t st);
hipError_t launch_update2d(bool fp64, const Solve2D& S, const Solve2D* d_S, double* d_rcell, hipStream_t st);
hipError_t launch_resid2d(const Solve2D& S, const double* d_rcell, double* d_res, hipStream_t st);

}  // namespace kbg

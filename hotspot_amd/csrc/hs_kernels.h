// Device data layout and kernels of the grid thermal model (gfx950).
//
// Every kernel works on a work-list of (problem, tile) items over an array of
// problem descriptors, so one launch serves one model or a batch of models of
// different sizes.  A tile is a TX x (TY * rpt) patch of one layer (TX*TY =
// 256 threads, TX along the contiguous column index so each wavefront reads
// 64 consecutive values; each thread owns rpt consecutive rows of its
// column); the last item of each problem is a "package" workgroup that
// handles the 12 (or 28) package nodes, whose rows couple to whole boundary
// rows of the spreader / sink (and secondary) layers.
//
// Per-problem scalars (CG coefficients, RK4 step sizes, flags) live in device
// memory; reductions are two-level and deterministic: each workgroup writes
// its partial into a fixed slot with plain stores, and a finish kernel (one
// workgroup per problem, launched after the sweep) combines the slots in
// slot order.
#pragma once

#include <hip/hip_runtime.h>

#include "hs_internal.h"

namespace hs {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kRed = 4;   // values reduced per slot
constexpr int kGmM = 32;  // GMRES restart length (non-symmetric steady solves)
constexpr unsigned short kCellMixed = 0xffff;   // cell_u16: not inside one block

// Per-problem Taylor record of a group (doubles): [0] the sub-step tau and
// [1] its term count (0: nothing to do), written by the finish of the first
// term; then, per Horner step j = 2..kTayMaxK, the factors of the step
//   h <- E_1 + s_j M h + t_j h,   M = -C^-1 G,
// at [2 j] (s_j) and [2 j + 1] (t_j), written by the host with the plan,
// and the same factors rounded to FP32 at float index 2 j and 2 j + 1 from
// [kTrecF] on (uniform FP32 operands for the FP32 sweeps).
constexpr int kTayMaxK = 60;
constexpr int kTrecF = 2 * (kTayMaxK + 1);
constexpr int kTrec = kTrecF + (kTayMaxK + 1);

// FP64 buffers of a problem
enum Buf64 {
    B_STEADY = 0,  // steady solution (last_steady)
    B_Y0, B_Y1,    // transient state, ping-pong by parity (last_trans)
    B_RHS,         // P + ambient / inlet sources
    B_K1, B_KA, B_KB, B_ACC, B_T1,   // RK4 work
    B_R, B_P0, B_P1, B_Q, B_Z,       // CG work
    B_TMP,
    NB64
};
// FP32 buffers of a problem
enum Buf32 {
    F_R = 0, F_P0, F_P1, F_Q, F_X, F_Z,   // inner CG of the mixed steady solve
    F_D0, F_D1, F_ACC,               // Taylor terms (mixed transient)
    F_PW,                            // cell power without sources (FP32 Taylor groups)
    NB32
};

// per-problem scalar slots (doubles)
enum Scal {
    S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_TOL2, S_ITERS, S_DONE, S_PPAR,
    S_RR0,                       // inner solve: initial residual
    S_IDONE, S_ITOL2,            // inner solve flags
    S_H, S_NEWH, S_T, S_DT, S_STATUS, S_YPAR, S_CALLS, S_ERR,
    S_TAU, S_TMAX, S_TDONE, S_TK,   // Taylor: step, term max, done flag, substeps
    S_NREFINE,
    S_STEP,                      // interval index of a device-resident sequence
    S_TITERS,                    // mixed steady: inner iterations summed over refinements
    S_RHO,                       // Taylor: bound on ||C^-1 G||_inf
    S_TCALLS,                    // RK4 calls summed over a sequence
    S_KNEED,                     // Taylor: terms of the current sub-step
    S_NSTEPS,                    // intervals of a device-resident sequence
    S_CORR,                      // mixed steady: a correction e is pending
    S_TTOL,                      // Taylor: truncation tolerance per sub-step
    S_KPLAN,                     // Taylor: most terms a sub-step may take
    S_TTERMS,                    // Taylor: term sweeps (E_k -> E_k+1) summed over a sequence
    S_MGINIT,                    // multigrid-preconditioned inner solve: first V-cycle pending
    S_TBYTES,                    // Taylor: algorithmic bytes per node of the dominant sweep kernel, summed
    S_TAU1,                      // Taylor: factor of the slope in E_1 (tau phi_1(-sigma); tau unshifted)
    S_TSIG,                      // Taylor: shift sigma = rho tau / 2 of the series (0: unshifted)
    S_THOLD,                     // Taylor: interval held (the group launched fewer steps than it needs)
    S_NHOLD,                     // Taylor: held intervals summed over a sequence
    NSCAL = 48
};
static_assert(S_NHOLD < NSCAL, "scalar slots");
enum RkStatus { RK_NEED_K1 = 0, RK_ATTEMPT = 1, RK_DONE = 2 };

struct ProbDesc {
    // geometry and tiling
    int nl, nr, nc, nx;
    int tx_shift, nti, ntj, ntiles;   // tiles: nl * nti * ntj, then the package item
    int hetero, has_adv, nblocks;
    int rpt;                          // grid rows per thread in a tile
    int nti2;                         // column tiles along the rows (multigrid: all layers per thread)
    // tiling of the float4 Taylor sweep (chosen by timing; the arithmetic of
    // a cell does not depend on it): tile width 1 << t_shift, t_rpt rows per
    // thread, t_nti tiles along the rows, t_ntiles tiles, then the package item
    int t_shift, t_rpt, t_nti, t_ntiles;
    int t_z, pad0;                    // 1: the sweep z-marches (column tiles through every layer)
    double ambient;
    double cell_area;
    // homogeneous layer coefficients
    double gx[kMaxLayers];     // row neighbour (i +- 1): 1/rx
    double gy[kMaxLayers];     // column neighbour (j +- 1): 1/ry
    double gz[kMaxLayers];     // layer n <-> n+1: 1/rz[n]
    double cinv[kMaxLayers];   // 1/c
    double gamb[kMaxLayers];   // per-cell conductance to ambient
    double ring_gns[kMaxLayers], ring_gew[kMaxLayers];
    int ring_node[kMaxLayers][4];  // package node of the N,S,E,W edge, -1 none
    // package nodes
    double x_gamb[kMaxExtra], x_cinv[kMaxExtra];
    int x_ring_layer[kMaxExtra], x_ring_edge[kMaxExtra];
    double x_ring_g[kMaxExtra];
    int x_ncpl[kMaxExtra];
    int x_cpl_node[kMaxExtra][3];
    int x_cpl_steady_only[kMaxExtra][3];
    double x_cpl_g[kMaxExtra][3];
    // device buffers
    double *b64[NB64];
    float *b32[NB32];
    // heterogeneous (detailed 3D / microchannel) per-cell coefficients
    const double *gE, *gS, *gD, *cinvc, *dinvc;
    const double *aN, *aS, *aE, *aW, *aSelf;
    const float *fE, *fS, *fD, *fcinv, *fdinv;
    const float *faN, *faS, *faE, *faW, *faSelf;
    // backward-Euler descriptors of heterogeneous models: sigma C per cell,
    // added to the ambient conductance of the stencil (nullptr otherwise)
    const double *bshift;
    // block <-> grid maps (per layer)
    const int *map_ptr[kMaxLayers];     // cell -> entry range
    const int *map_unit[kMaxLayers];
    const double *map_occ[kMaxLayers];
    const double *unit_area[kMaxLayers];
    // per cell of a layer: its blocks (up to four) and their weights occ *
    // cell area / block area, as [2 c] {u0, u1, u2, u3} and [2 c + 1] the
    // weights bit-cast (unused slots: u0 with weight 0; u0 = -1 for a cell
    // of five or more blocks, which takes its map entries)
    const int4 *cell_uw[kMaxLayers];
    // per cell of a layer: the one block that covers it whole, or kCellMixed
    // (the cell takes cell_uw), and per block of the layer its weight
    // (float)(cell area / block area); null where a layer's floorplan has
    // kCellMixed blocks or more
    const unsigned short *cell_u16[kMaxLayers];
    const float *unit_w[kMaxLayers];
    // with cell_u16: the cells marked kCellMixed listed by tile of the
    // standard tiling, in tile order (mix_off: per tile of a layer, the first
    // entry, then the count), each with its plane index and cell_uw record
    const int *mix_off[kMaxLayers];
    const int *mix_cell[kMaxLayers];
    const int4 *mix_uw[kMaxLayers];
    int layer_base[kMaxLayers];         // block-vector offset per layer
    const int4 *blk_range;              // per block of the vector: i1,i2,j1,j2
    const int *blk_layer;
    // per block of the vector: unit_w of its layer, by block index; set when
    // every layer has cell_u16 and every kCellMixed cell has a cell_uw record
    // of one to four blocks (null otherwise)
    const float *blk_w;
    int map_mode, pad1;
    double *blk_in;                     // block vector (device staging)
    double *blk_out;
    double *seq_in;                     // [steps][nblocks + nx] power sequence
    double *seq_out;                    // [steps][nblocks + nx] temperatures
    double *bconst;                     // per-cell constant source (inlet), hetero only
    int *lzero;                         // per layer: F_PW holds no power (all blocks 0), not written
    // GMRES (non-symmetric steady solves; allocated on first use)
    double *gm;                         // per-problem state (GmState)
    double *gpart;                      // [ntiles+1][kGmM+1] partials
    float *gmv32;                       // [kGmM+1][vlen] basis, FP32 inner solves
    double *gmv64;                      // [kGmM+1][vlen] basis, FP64 inner solves
    unsigned vlen;
    int seq_one;                        // 1: every interval of the sequence takes seq_in's one row
    // reduction state
    double *partial;                    // [ntiles+1][kRed]
    unsigned *counter;
    double *scal;                       // [NSCAL]
};

struct WorkItem {
    int prob;
    int tile;
};

}  // namespace hs

// ifl_adjoint.hpp - device buffers and launchers of the discrete adjoint (adjoint.hip).
//
// One adjoint iteration lambda <- G_x^T lambda + J_x^T transposes one SIMPLE outer iteration at
// its fixed point (u* = smoothed velocity, p' = correction, reverse stage order):
//   k_adj_pcorr      p'-bar from the corrections (and b'-bar := 0)
//   k_adj_sweep_t    x 2 n_p   transposed red-black half sweeps of the p' system (2D: k_adj_tile2d,
//                              up to four sweeps per launch)
//   k_adj_rc         u*-bar, the Rhie-Chow pressure terms, aC-bar, inlet theta-bar (divergence)
//   k_adj_sweep_t    x 2 n_d   transposed half sweeps of each momentum component
//   k_adj_coef       coefficient adjoints before the BC rows, new p-bar, mu-bar partials
//   k_adj_face       new u-bar and face adjoints, inlet theta-bar (coefficients)
//   k_adj_reduce     increments, scaled increment, mu-bar (one workgroup)
// Every output element is a gather over the inputs that reach it, in a fixed order, so the
// FP32 result equals the oracle's (oracle/adjoint_oracle.hpp) bit for bit.
#pragma once

#include "ifl_device.hpp"
#include "ifl_kernels.hpp"

namespace ifl {

struct AdjState {
    double inc[7];        // squared increments of u, v, w, p, uf, vf, wf in the last iteration
    double incMax[7];     // their maxima over the first 5 iterations
    double scaled;        // max_q sqrt(inc[q] / incMax[q])
    double mub;           // dJ/dmu of the last iteration
    long long iters;
};

struct AdjBufs {
    float *lu[3], *lp, *lf[3];       // adjoint state (face arrays laid out as uf, vf, wf)
    float *pcb, *pcbAlt;             // p'-bar and its ping-pong partner
    float *bpc;                      // b'-bar
    float *us[3], *usAlt[3];         // u*-bar and partners
    float *acb;                      // aC-bar
    float *rb[3];                    // adjoint of the momentum right-hand sides
    float *abar;                     // coefficient adjoints (1 + 2 dim planes)
    float *pacc;                     // p-bar partial (Rhie-Chow terms)
    const float *gu[3], *gp;         // objective weights (nullptr = 0)
    double *gprof[6];                // theta-bar per face cell (inlet faces)
    double *part;                    // per-workgroup partials: [mu-bar, inc p] then [inc u,v,w,uf,vf,wf]
    int nblk;                        // workgroups of the cell grid
    AdjState *st;
};

void launch_adj_pcorr(const Lc &L, const AdjBufs &A);
void launch_adj_sweep_t(const Lc &L, const AdjBufs &A, const float *in, float *out, float *bb, const float *coef,
                        float rlx, int col);
// 2D: ns (<= adj_tile_sweeps(shape)) transposed sweeps in one launch; shapes 0..5 are the tiles
// 64x16, 32x16, 64x8, 32x8, 64x16, 64x8 with 4, 4, 2, 4, 2, 4 sweeps per launch
constexpr int kAdjShapes = 6;
int adj_tile_sweeps(int shape);
void launch_adj_tile2d(const Lc &L, int shape, const float *in, float *out, float *bb, const float *coef, float rlx,
                       int ns);
void launch_adj_rc(const Lc &L, const AdjBufs &A);
void launch_adj_coef(const Lc &L, const AdjBufs &A);
void launch_adj_face(const Lc &L, const AdjBufs &A);
void launch_adj_reduce(const Lc &L, const AdjBufs &A);
// Workgroups of the adjoint's cell kernels (their partial-sum records).
int adj_blocks(const CaseParams &P);

}  // namespace ifl

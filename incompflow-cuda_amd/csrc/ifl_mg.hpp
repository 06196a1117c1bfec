// ifl_mg.hpp - the multigrid p' solve (IFL_PSOLVE_MG) on 2D grids: level layout and launchers.
//
// Level 0 is the case's p' system: x = p' (P.pc), b = b' (P.bp), the +x/+y couplings and the
// centre of the assembled p' matrix (kernels.cuh:502-540, planes of P.pcoef) plus, in cells next
// to an outlet face, the coupling rho A^2 / aC to the ghost p' = 0 of the velocity correction
// (kernels.cuh:713-727), computed from P.rinv on the fly.  Level l+1 joins 2x2 blocks of level l
// (the last block of an odd side holds one column or row) while both sides of level l are at
// least 4.  A coarse coupling is the sum of the fine couplings across the coarse face, the coarse
// outlet term the sum of its children's and the centre -(E + W + N + S) + outlet term (the
// couplings towards -x and -y of a cell are the +x and +y ones of its neighbours, bit for bit).
//
// One V-cycle: down the levels (unrelaxed red-black sweeps from x = 0, FP64 residual, the
// children's residuals summed in FP64 and rounded once into the coarse b), the coarsest level
// smoothed `ncoarse` times, and up (x += over * x_parent, sweeps).  Levels of at most
// kMgBottomCells cells run in one single-workgroup kernel with everything in LDS; it also builds
// the matrices of the levels below its first one, so the coarsening kernel stops at that level.
#pragma once

#include "ifl_device.hpp"
#include "ifl_kernels.hpp"

namespace ifl {

constexpr int kMgMaxLevels = 16;
constexpr int kMgBottomCells = 4096;   // levels at most this large run in the bottom kernel
constexpr int kMgCoarsenDepth = 6;     // levels one coarsening pass makes (64x64 source tiles)

// A level l >= 1: nx x ny cells; the correction after the down pass (x) and after the up pass
// (xu: the up pass reads the down pass's values around its tile while other tiles write theirs,
// so it writes another array), the restricted residual, centre (with the outlet term), +x and +y
// couplings, outlet term.  For level 0 only nx and ny are used: its down pass writes the p'
// partner buffer (P.alt[3], idle under the multigrid) and its up pass p' itself (P.pc).
struct MgLevel {
    int nx, ny;
    float *x, *xu, *b, *C, *E, *N, *dg;
};

// The coarsening pass from level `src` to the next m (<= kMgCoarsenDepth) levels.
struct MgCoarsen {
    MgLevel s;                            // the source level (level 0: from the case's p' system)
    MgLevel d[kMgCoarsenDepth];           // levels src+1 .. src+m
    int m;
};

// A p' solve as the host launches it.
struct MgLaunch {
    int nl = 0;                   // levels, level 0 included
    int lb = 0;                   // first level of the bottom kernel
    int cycles = 1, nu = 2, ncoarse = 4;
    float over = 1.0f;
    MgLevel lv[kMgMaxLevels] = {};
    const MgLevel *dev = nullptr; // device copy of lv (the bottom kernel reads its levels there)
};

// A level l >= 1 of a 3D grid: the correction x (and its partner for ping-pong sweeps), the restricted residual b, the
// centre (with the outlet term), the +x/+y/+z couplings and the outlet term.  Level 0 is the case's
// p' system (x = P.pc, b = P.bp, the planes of P.pcoef); only its sizes are used.
struct MgLevel3 {
    int nx, ny, nz;
    float *x, *xa, *b, *C, *E, *N, *T, *dg;   // xa: the z-march sweeps' partner of x
};

// A 3D p' solve as the host launches it (level-wide kernels, mg3d.hip).
constexpr int kMg3BottomFloats = 30 * 1024;   // LDS of the 3D bottom kernel (x, b and four matrix planes)

struct MgLaunch3 {
    int nl = 0;
    int lb = 0;                   // first level of the single-workgroup bottom kernel (nl: none)
    int cycles = 1, nu = 2, ncoarse = 4;
    float over = 1.0f;
    MgLevel3 lv[kMgMaxLevels] = {};
};

// Level sizes of an nx x ny x nz grid (dims[l][0..2]); returns the number of levels.
int mg3_levels(int nx, int ny, int nz, int (*dims)[3]);
// First level (>= 1) from which all coarser levels fit the bottom kernel's LDS (nl: none).
int mg3_bottom_level(int nl, const int (*dims)[3]);
// Build the coarse matrices, then run G.cycles V-cycles (one case, not batched).
void launch_mg3_solve(const Lc &L, const MgLaunch3 &G);

// Level sizes of an nx x ny grid (dims[l][0..1]); returns the number of levels.
int mg_levels(int nx, int ny, int (*dims)[2]);
// First level that the bottom kernel takes (-1: the coarsest level does not fit in its LDS).
int mg_bottom_level(int nl, const int (*dims)[2]);
// Coarsen the matrix of every level, then run G.cycles V-cycles (one case, not batched).
void launch_mg_solve(const Lc &L, const MgLaunch &G);

}  // namespace ifl

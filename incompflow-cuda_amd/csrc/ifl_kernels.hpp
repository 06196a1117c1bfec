// ifl_kernels.hpp - host-side launchers of the gfx950 kernels (kernels.hip, sweep2d.hip,
// sweep3d.hip).  Every launcher runs its kernel on one case or on a batch of cases (Lc).
#pragma once

#include "ifl_device.hpp"

namespace ifl {

// One equation of a smoother launch.  The field, its ping-pong partner, source and matrix follow
// from the slot (u, v, w, p', refinement correction); launch parity selects the direction.
struct SweepSel {
    int slot;
    int q;       // equation index for relax / tol (u, v, w, p')
    int k;       // sweeps of this equation in this launch (0..K)
    int kPrev;   // sweeps of this equation in the previous launch of the call
    int flip = 0;    // 1: launch L reads the partner buffer when L is even (odd launch counts end
                     //    in the caller's field without a copy; the partner holds the field)
    int accum = 0;   // 1: also add the result into p' (last launch of a refinement correction)
};

enum { COEF_STORED = 0, COEF_MOMENTUM = 1, COEF_PCORR = 2, COEF_TEMP = 3 };
// 2D sweep tiles; 128x32 runs 1024-thread workgroups, the others 256
enum { TILE_AUTO = 0, TILE_64x16 = 1, TILE_64x8 = 2, TILE_32x8 = 3, TILE_32x4 = 4, TILE_128x32 = 5 };
enum { RED_MOM = 1, RED_MASS = 2, RED_UPD = 4, RED_FINISH = 8 };

// Launch context: one case (P, grids from P) or a batch (device map, total workgroups per kind).
struct Lc {
    const CaseParams *P = nullptr;   // the case (batch: any case, only its dim/tile are read)
    BatchMap M{};
    bool batch = false;
    int total[MAP_KINDS] = {};
    int tile = TILE_64x16;           // 2D sweep tile
    hipStream_t s = nullptr;
};

int cell_blocks(const CaseParams &P);
int copy_blocks(const CaseParams &P);
int tile_count(int nx, int ny, int tile);
int tile_dimx(int nx, int tile);
int default_tile(int nx, int ny);
int zmarch_blocks(const CaseParams &P, bool pc);   // pc: the p' sweep's column shape
int zmarch_chunk(const CaseParams &P, bool pc);    // z planes per column chunk (sets P.zcM / P.zcP)
int zmarch_dimx(const CaseParams &P, bool pc);
int cell_dimx(const CaseParams &P);

void launch_init_faces(const Lc &L);
void launch_face_bc(const Lc &L);
void launch_mom_assemble(const Lc &L);
// fin: smoother calls closed by workgroup 0 of this launch (instead of k_smooth_finish)
void launch_rhie_chow(const Lc &L, const Finish &fin = Finish{});
// writeAlt: also clear the p' smoother's partner buffer (a p' call with an odd launch count starts there)
void launch_pcorr_assemble(const Lc &L, bool writeAlt = true);
// writeAlt: also store the corrected u, v, w into the momentum smoothers' partner buffers
void launch_update(const Lc &L, const Finish &fin = Finish{}, bool writeAlt = true);
void launch_pcorr_defect(const Lc &L, const Finish &fin = Finish{});
void launch_pcorr_accumulate(const Lc &L);
void launch_reduce_cells(const Lc &L, int mask);
void launch_sweep_jacobi(const Lc &L, const SweepSel &e, int launch);
void launch_smooth_finish(const Lc &L, int slot, int q, int nL, int kLast, int sweepKind, int nprev);
void launch_copy_back(const Lc &L, int slot, int dir);   // dir 0: partner -> field, 1: field -> partner
// 2D fused red-black sweeps (K = block sweeps per launch), one or two equations sharing a matrix
void launch_rb_sweep2d(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch, int block);
// 3D z-marching red-black sweep
// 3D red-black: neq (1..3) momentum equations together, or the p' system (kind COEF_PCORR, neq 1)
void launch_rb_sweep3d(const Lc &L, int kind, const SweepSel *e, int neq, int launch);

}  // namespace ifl

// ifl_kernels.hpp - host-side launchers of the gfx950 kernels (kernels.hip, sweep2d.hip).
#pragma once

#include "ifl_device.hpp"

namespace ifl {

// One equation of a fused sweep: read `in`, write `out` (ping-pong), source `src`.
struct SweepEq {
    const float *in;
    float *out;
    const float *src;
    float rlx;
    double tol;
    int slot;
    int eq;
    int k;       // sweeps of this equation in this launch (0..K)
    int kPrev;   // sweeps of this equation in the previous launch of the call
};

enum { COEF_STORED = 0, COEF_MOMENTUM = 1, COEF_PCORR = 2 };
enum { TILE_AUTO = 0, TILE_64x16 = 1, TILE_64x8 = 2, TILE_32x8 = 3, TILE_32x4 = 4 };
enum { RED_MOM = 1, RED_MASS = 2, RED_UPD = 4, RED_FINISH = 8 };

int cell_blocks(const CaseParams &P);
int rb_blocks(const CaseParams &P);
int tile_count(int nx, int ny, int tile);
int default_tile(int nx, int ny);

void launch_init_faces(const CaseParams &P, hipStream_t s);
void launch_face_bc(const CaseParams &P, hipStream_t s);
void launch_mom_assemble(const CaseParams &P, hipStream_t s);
void launch_rhie_chow(const CaseParams &P, hipStream_t s);
void launch_pcorr_assemble(const CaseParams &P, hipStream_t s);
void launch_update(const CaseParams &P, hipStream_t s);
void launch_reduce_cells(const CaseParams &P, int mask, hipStream_t s);
void launch_pcorr_defect(const CaseParams &P, hipStream_t s);
void launch_pcorr_accumulate(const CaseParams &P, hipStream_t s);
void launch_sweep_rb_half(const CaseParams &P, float *fld, const float *coef, const float *src, float rlx, double tol,
                          int slot, int sweep, hipStream_t s);
void launch_sweep_jacobi(const CaseParams &P, const float *in, float *out, const float *coef, const float *src,
                         float rlx, double tol, int slot, int sweep, hipStream_t s);
void launch_smooth_finish(const CaseParams &P, int slot, int eq, int nL, int kLast, int nprev, double tol,
                          hipStream_t s);
void launch_rb_sweep2d(const CaseParams &P, int kind, const float *coef, const SweepEq &e0, const SweepEq &e1, int neq,
                       int launch, int block, int tile, hipStream_t s);
void launch_copy(const CaseParams &P, float *dst, const float *src, hipStream_t s);
int zmarch_blocks(const CaseParams &P);
void launch_rb_sweep3d(const CaseParams &P, const float *coef, const SweepEq &e, int launch, hipStream_t s);

}  // namespace ifl

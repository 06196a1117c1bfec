// ifl_kernels.hpp - host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include "ifl_device.hpp"

namespace ifl {

void launch_init_faces(const CaseParams &P, hipStream_t s);
void launch_face_bc(const CaseParams &P, hipStream_t s);
void launch_mom_assemble(const CaseParams &P, hipStream_t s);
void launch_rhie_chow(const CaseParams &P, hipStream_t s);
void launch_pcorr_assemble(const CaseParams &P, hipStream_t s);
void launch_update(const CaseParams &P, int finish, hipStream_t s);
void launch_pcorr_defect(const CaseParams &P, hipStream_t s);
void launch_pcorr_accumulate(const CaseParams &P, hipStream_t s);
void launch_sweep(const CaseParams &P, int smoother, float *fld, const float *coef, const float *src, float rlx,
                  double tol, int slot, int eq, int sweep, hipStream_t s);

}  // namespace ifl

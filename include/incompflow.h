/*
 * incompflow.h - C-ABI of the MI355X steady SIMPLE solver (collocated structured grids).
 *
 * This is the drop-in boundary for the reference's `Solver` class and its compile-time
 * `config.h` contract.  Every entry point takes plain pointers and sizes, returns an int
 * status (IFL_OK == 0) and leaves a thread-local message for ifl_last_error().  No C++
 * or torch type crosses it.  A handle owns all of its device memory and one HIP stream;
 * buffers passed in are copied, never retained.  Handles are independent: one handle per
 * host thread, no device-wide synchronisation.
 *
 * Reference surface replaced (IncompFlow-CUDA):
 *   include/config.h:6-81     compile-time case     -> ifl_config (runtime, same fields)
 *   include/constants.h:7-36  face/slot/BC codes    -> IFL_EAST.., IFL_SLOT_*, IFL_WALL..
 *   include/Solver.h:46       Solver(u,v,w,p,temp)  -> ifl_create (ifl_config.init[])
 *   include/Solver.h:48       ~Solver()             -> ifl_destroy
 *   include/Solver.h:50       solve()               -> ifl_solve
 *   include/Solver.h:52       writeVTK(path)        -> ifl_write_vtk
 *   include/Solver.h:54-71    private stage methods -> ifl_run_stage (tests only)
 */
#ifndef INCOMPFLOW_H
#define INCOMPFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFL_ABI_VERSION 3

/* ---- status codes ---- */
enum {
    IFL_OK = 0,
    IFL_ERR_ARG = 1,        /* invalid argument / config */
    IFL_ERR_DEVICE = 2,     /* HIP runtime error */
    IFL_ERR_DIVERGED = 3,   /* non-finite residual detected on the device */
    IFL_ERR_NOT_CONVERGED = 4, /* stop rule not met within the iterations allowed (only with
                                  ifl_config.fail_if_not_converged = 1) */
    IFL_ERR_IO = 5,
    IFL_ERR_NOMEM = 6
};

/* ---- boundary faces, include/constants.h:7-12 ---- */
enum { IFL_EAST = 0, IFL_WEST = 1, IFL_NORTH = 2, IFL_SOUTH = 3, IFL_TOP = 4, IFL_BOTTOM = 5 };

/* ---- coefficient slots, include/constants.h:14-20 (plane a of a coefficient array) ---- */
enum { IFL_SLOT_C = 0, IFL_SLOT_E = 1, IFL_SLOT_W = 2, IFL_SLOT_N = 3, IFL_SLOT_S = 4,
       IFL_SLOT_T = 5, IFL_SLOT_B = 6 };

/* ---- velocity BC codes, include/constants.h:34-36 ---- */
enum { IFL_WALL = 0, IFL_INLET = 1, IFL_OUTLET = 2 };

/* ---- temperature BC codes, include/config.h:46-53 (TempBCs::type) ---- */
enum { IFL_TEMP_DIRICHLET = 0, IFL_TEMP_NEUMANN = 1 };

/* ---- smoother (linear solver) ---- */
enum {
    IFL_SMOOTHER_RBGS = 0,   /* red-black Gauss-Seidel (deterministic form of GaussSeidelIterate) */
    IFL_SMOOTHER_JACOBI = 1, /* point Jacobi, pointJacobiIterate with the result copied back */
    IFL_SMOOTHER_LEX = 2     /* lexicographic Gauss-Seidel: CPU oracle only */
};

/* ---- p' solver ---- */
enum {
    IFL_PSOLVE_SWEEPS = 0,   /* the reference's relaxed smoother call (Solver.cu:196), optionally refined */
    IFL_PSOLVE_MG = 1        /* aggregation-multigrid V-cycles of controlled strength (mg_* fields) */
};

/* ---- outer stop rule ---- */
enum {
    IFL_STOP_NONE = 0,      /* run exactly num_outer_iter outer iterations */
    IFL_STOP_FAITHFUL = 1,  /* Solver.cu:237-268 as written (all maxima tracked from pNorm) */
    IFL_STOP_FIXED = 2,     /* same test, each variable normalised by its own running max */
    IFL_STOP_SCALED = 3     /* scaled FP64 residual < tol_residual (see ifl_iter_record) */
};

/* ---- field identifiers for ifl_get_field / ifl_set_field ---- */
enum {
    IFL_FIELD_U = 0, IFL_FIELD_V = 1, IFL_FIELD_W = 2, IFL_FIELD_P = 3, IFL_FIELD_TEMP = 4,
    IFL_FIELD_PCORR = 5,
    IFL_FIELD_UF = 6,   /* (nx+1)*ny*nz, index i + (nx+1)(j + ny k) */
    IFL_FIELD_VF = 7,   /* nx*(ny+1)*nz, index i + nx(j + (ny+1) k) */
    IFL_FIELD_WF = 8,   /* nx*ny*(nz+1), index i + nx(j + ny k), k in [0,nz] */
    IFL_FIELD_MCOEF = 9,  /* momentum matrix, (1+2dim) planes of N (one set for u,v,w) */
    IFL_FIELD_PCOEF = 10, /* pressure-correction matrix, (1+2dim) planes of N */
    IFL_FIELD_BU = 11, IFL_FIELD_BV = 12, IFL_FIELD_BW = 13,
    IFL_FIELD_BP = 14,    /* mass imbalance b' */
    IFL_FIELD_TCOEF = 15, /* temperature matrix, (1+2dim) planes of N (energy on) */
    IFL_FIELD_BT = 16,    /* temperature source (energy on) */
    IFL_FIELD_COUNT = 17
};

/* ---- stages for ifl_run_stage: one SIMPLE outer iteration is 1..9 in order ---- */
enum {
    IFL_STAGE_INIT_FACES = 0,   /* Solver.cu:369-408  initFaceVel (first iteration only) */
    IFL_STAGE_FACE_BC = 1,      /* Solver.cu:410-465  applyBCsToFaceVel */
    IFL_STAGE_MOM_ASSEMBLE = 2, /* Solver.cu:167-175  coefficients + sources + BC rows (and the
                                   temperature matrix and source, energy on) */
    IFL_STAGE_MOM_RESIDUAL = 3, /* FP64 ||b - A phi|| per momentum component and T (no reference) */
    IFL_STAGE_SMOOTH_U = 4,     /* Solver.cu:178 */
    IFL_STAGE_SMOOTH_V = 5,     /* Solver.cu:180 */
    IFL_STAGE_SMOOTH_W = 6,     /* Solver.cu:183 */
    IFL_STAGE_RHIE_CHOW = 7,    /* Solver.cu:186-187  RhieChowInterpolate + applyBCsToFaceVel */
    IFL_STAGE_PCORR_ASSEMBLE = 8, /* Solver.cu:189-193 coefficients, b', massNorm, p'=0 */
    IFL_STAGE_SMOOTH_P = 9,     /* Solver.cu:196 (+ FP64-defect refinement cycles) */
    IFL_STAGE_UPDATE = 10,      /* Solver.cu:198      updateField */
    IFL_STAGE_OUTER = 11,       /* one whole outer iteration incl. the stop test */
    IFL_STAGE_SMOOTH_T = 12,    /* temperature sweeps (energy on; after SMOOTH_U..W in an iteration) */
    IFL_STAGE_COUNT = 13
};

/* Runtime case description; field-for-field the reference's config.h (line cited). */
typedef struct ifl_config {
    int32_t dim;                 /* config.h:6   2 or 3 */
    int32_t nx, ny, nz;          /* config.h:8-10 (nz = 1 in 2D) */
    double xmin, xmax;           /* config.h:12-13 */
    double ymin, ymax;           /* config.h:14-15 */
    double zmin, zmax;           /* config.h:16-17 */
    double density;              /* config.h:21 */
    double viscosity;            /* config.h:22 dynamicViscosity */
    int32_t bc_type[6];          /* config.h:26-33 velBCs::type[face] */
    double bc_val[6][3];         /* config.h:35-42 velBCs::val[face][u,v,w] */
    int32_t num_outer_iter;      /* config.h:65 */
    int32_t n_iter[4];           /* config.h:67-70 nIter_u, nIter_v, nIter_w, nIter_p */
    double relax[4];             /* config.h:72-75 relax_u, relax_v, relax_w, relax_p */
    double tol[4];               /* config.h:77-80 tol_u, tol_v, tol_w, tol_p */
    double tol_mass;             /* config.h:81 */
    double init[5];              /* Solver.h:46 uniform initial u, v, w, p, temp */
    /* additions with no reference field */
    int32_t smoother;            /* IFL_SMOOTHER_* (reference: Gauss-Seidel, Solver.cu:178) */
    int32_t stop_rule;           /* IFL_STOP_* (reference: IFL_STOP_FAITHFUL) */
    double tol_residual;         /* threshold for IFL_STOP_SCALED */
    int32_t p_refine;            /* FP64-defect refinement cycles of the p' solve (0 = off) */
    int32_t log_every;           /* print the Solver.cu:270-285 block every k iterations (0 = never) */
    int32_t check_every;         /* outer iterations between host reads of the stop flag (0 = auto) */
    int32_t tile;                /* 2D sweep tile: 0 auto, 1 64x16, 2 64x8, 3 32x8, 4 32x4,
                                    5 128x32 (1024-thread workgroups) */
    int32_t sweep_block;         /* red-black sweeps per launch of the p' system: 1, 2, 4 or 5 in 2D,
                                    1 or 2 in 3D (larger values run 2); a stop test met inside a
                                    block takes effect at the block's end (1 = the reference's
                                    per-sweep exit) */
    /* Energy (temperature) transport.  The reference declares the thermal properties and the
     * temperature BC table (config.h:19-23, 45-63) but never solves for temperature; here the
     * steady equation  div(rho u T) = div((k / c_p) grad T)  is discretised like the momentum
     * equations (kernels.cuh:116-131 with mu -> k / c_p; Dirichlet faces take the wall row of
     * kernels.cuh:304-320, Neumann faces the zero-gradient row of :321-331) on the face
     * velocities of the outer iteration, and swept after the momentum equations.  Temperature
     * is passive: nothing of the flow depends on it. */
    int32_t energy;              /* 1: solve for temperature in every outer iteration (0: carry it) */
    int32_t n_iter_t;            /* temperature sweeps per outer iteration */
    double conductivity;         /* config.h:19 thermalConductivity */
    double specific_heat;        /* config.h:20 specificHeatCapacity */
    double relax_t;              /* under-relaxation of the temperature sweeps */
    double tol_t;                /* sweep stop tolerance of the temperature sweeps */
    int32_t temp_bc_type[6];     /* config.h:46-53 TempBCs::type[face]: IFL_TEMP_DIRICHLET / NEUMANN */
    double temp_bc_val[6];       /* config.h:55-62 TempBCs::val[face] (Dirichlet values) */
    int32_t sweep_block_mom;     /* the same for the momentum and temperature equations: 0 = the
                                    smaller of sweep_block and 2; 1, 2, 4 or 5.  Four or five pay on
                                    large grids with 128x32 tiles; on small tiles the halo costs more */
    /* The p' solve.  The reference relaxes p' with nIter_p Gauss-Seidel sweeps (Solver.cu:196) and
     * notes that large cases converge slowly (README.md:416).  IFL_PSOLVE_MG replaces that call by
     * mg_cycles V-cycles of an aggregation multigrid on the same matrix (kernels.cuh:502-540):
     * coarse cells join 2x2(x2) fine cells, coarse couplings are the sums of the fine couplings
     * across the coarse face, every level is smoothed by unrelaxed red-black Gauss-Seidel, the
     * residual and its restriction are FP64, and the coarse correction is scaled by mg_over.  The
     * operator adds, on outlet faces, the coupling to the ghost value p' = 0 that the correction
     * step applies there (kernels.cuh:713-727).  A converged state is the same for every p' solver
     * (b' = 0 gives p' = 0); the solver only changes how fast it is reached.  A near-exact p' solve
     * makes this SIMPLE variant diverge, so the strength (mg_cycles, mg_over) is a tuning choice
     * (ifl_autotune, IFL_TUNE_PSOLVE).  2D and 3D; not with p_refine or in a batch. */
    int32_t p_solver;            /* IFL_PSOLVE_SWEEPS (default) or IFL_PSOLVE_MG */
    int32_t mg_cycles;           /* V-cycles per p' solve, 1..8 (0 = 1) */
    int32_t mg_sweeps;           /* red-black sweeps before and after each coarse correction, 1..3 (0 = 2) */
    int32_t mg_coarse_sweeps;    /* red-black sweeps on the coarsest level, 1..200 (0 = 4:
                                    it has at most a few cells across) */
    double mg_over;              /* coarse-correction factor, (0, 2] (0 = 1.0) */
    int32_t fail_if_not_converged; /* 1: ifl_solve and ifl_batch_solve return IFL_ERR_NOT_CONVERGED when
                                      the stop rule was not met within the iterations allowed */
    int32_t reserved[3];
} ifl_config;

/* One record per outer iteration (device-written, FP64). */
typedef struct ifl_iter_record {
    double p_norm, u_norm, v_norm, w_norm;  /* RMS of applied increments, Solver.cu:226-235 */
    double uf_norm, vf_norm, wf_norm;
    double mass_norm;                       /* RMS of b' before correction */
    double res_cont, res_u, res_v, res_w;   /* FP64 L2: ||b'||, ||b_u - A u|| ... before sweeps */
    double scaled;                          /* max_x res_x / max(res_x over iterations 1..5), x over
                                               continuity, momentum (and temperature, energy on) */
    double rel_p, rel_u, rel_v, rel_w, rel_uf, rel_vf, rel_wf; /* stop-rule relative norms */
    int32_t sweeps[4];                      /* smoother sweeps executed: u, v, w, p' (incl. refine) */
    int32_t converged;                      /* stop rule met at this iteration */
    int32_t sweeps_t;                       /* temperature sweeps executed (energy on) */
    double res_t;                           /* FP64 L2 ||b_T - A_T T|| before the sweeps (energy on;
                                               part of the scaled residual then) */
} ifl_iter_record;

typedef struct ifl_report {
    int32_t iterations;          /* outer iterations executed by this call */
    int32_t converged;           /* stop rule met */
    int32_t status;              /* IFL_OK or IFL_ERR_DIVERGED */
    int32_t pad;
    double elapsed_ms;           /* device time of the call (HIP events) */
    double final_scaled;         /* last scaled residual */
    double sweeps_total;         /* smoother cell-sweeps summed over equations and iterations */
} ifl_report;

typedef struct ifl_solver ifl_solver;

/* Thread-local description of the last failure on this thread. */
const char *ifl_last_error(void);
int ifl_abi_version(void);

/* Fill cfg with the reference's shipped defaults (include/config.h). */
void ifl_config_default(ifl_config *cfg);

/* Solver.h:46 / Solver.cu:13-104. Allocates all device state, uniform initial fields. */
int ifl_create(const ifl_config *cfg, ifl_solver **out);
/* Solver.h:48 / Solver.cu:106-151. */
int ifl_destroy(ifl_solver *h);

/* Per-face profile of the face-normal velocity (replaces velBCs::val[face][normal] for
 * every boundary cell of that face; n = cells on the face: W/E ny*nz, N/S nx*nz, T/B nx*ny).
 * vals == NULL restores the uniform value.  No reference counterpart (config.h:35-42 is uniform). */
int ifl_set_face_profile(ifl_solver *h, int face, const double *vals, size_t n);

/* Solver.h:50 / Solver.cu:153-299: run outer iterations until the stop rule holds or
 * max_iter (<=0: cfg.num_outer_iter) iterations ran.  history (optional) receives one record
 * per executed iteration, at most hist_cap. */
int ifl_solve(ifl_solver *h, int max_iter, ifl_report *report,
              ifl_iter_record *history, size_t hist_cap);

/* No counterpart in the reference: for callers that cannot allocate device memory (serving
 * loops, code that a harness forbids to call hipMalloc).  One solve of cfg from the caller's device
 * fields u, v, p (w in 3D, temp with energy = 1; each nx*ny*nz floats, read as the initial state
 * and overwritten with the result) on `stream` (a hipStream_t of the current device; NULL: the
 * null stream), with max_iter, report and history as in ifl_solve.  face_profiles (NULL, or 6
 * device pointers, NULL per face for none) holds the face-normal velocity of each face's boundary
 * cells as ifl_set_face_profile would, in float.  Every other buffer of the case (faces,
 * matrices, partials, monitor, multigrid levels) lives in a static device array of
 * IFL_NOALLOC_SCRATCH_BYTES held by each GPU the library runs on until the process ends; a case
 * that needs more fails with IFL_ERR_ARG, before any device call (about 121 bytes per cell: up to
 * some 550k cells in 2D, e.g. 740x740, and 530k in 3D, e.g. 80^3).  No device
 * allocation, no graph: the kernels of ifl_solve are launched directly, and the fields equal those
 * of ifl_create + ifl_set_field + ifl_solve from the same fields, bit for bit.  Calls on one GPU
 * share the array: within a process they are serialised by a lock (each call drains its stream
 * before returning); overlapping calls from other processes have their own copies.  The call
 * starts at iteration 0, as the first ifl_solve of a new handle does (the face velocities are
 * built from the cell fields), so it does not continue an earlier call. */
#define IFL_NOALLOC_SCRATCH_BYTES (64u << 20)
int ifl_solve_noalloc(const ifl_config *cfg, void *stream, float *u, float *v, float *w, float *p, float *temp,
                      const float *const *face_profiles, int max_iter, ifl_report *report,
                      ifl_iter_record *history, size_t hist_cap);

/* Host copies of device state.  n must equal the field's element count. */
int ifl_get_field(ifl_solver *h, int field, float *dst, size_t n);
int ifl_set_field(ifl_solver *h, int field, const float *src, size_t n);
size_t ifl_field_size(const ifl_solver *h, int field);

/* Solver.h:52 / Solver.cu:301-367: legacy-VTK ASCII RECTILINEAR_GRID at path (as given;
 * the reference prefixes "output/", which ifl::Solver in Solver.h reproduces). */
int ifl_write_vtk(ifl_solver *h, const char *path);

/* ---- Binary state export / import (no reference counterpart) ----
 * The reference's constructor takes only uniform initial values (Solver.h:46) and its only output,
 * writeVTK (Solver.cu:301-367), is ASCII and write-only, so a run cannot be continued or warm
 * started.  A state file holds what the next outer iteration reads: the fields u, v, w, p, temp,
 * uf, vf, wf (FP32) and the outer-iteration monitor (iteration count, the running maxima of the
 * stop rules, the scaled-residual normalisers).  Loading it into a handle with the same dim and
 * grid size, then solving on, gives the same iterates bit for bit as the solve that wrote it would
 * have continued with; loading it into a handle whose other settings differ (viscosity, boundary
 * values, inlet profiles, relaxation, sweeps) is a warm start of that case from the saved state.
 *
 * File layout (little-endian): one ifl_state_header, then the eight fields in the order u, v, w,
 * p, temp, uf, vf, wf, each as ifl_get_field returns it (float32, its full element count). */
#define IFL_STATE_MAGIC "IFLSTATE"
#define IFL_STATE_VERSION 1

typedef struct ifl_state_header {
    char magic[8];               /* IFL_STATE_MAGIC, no terminating 0 */
    uint32_t version;            /* IFL_STATE_VERSION */
    int32_t dim, nx, ny, nz;
    int32_t last_converged;      /* the stop rule held at the last iteration */
    int64_t iter;                /* outer iterations executed */
    double norm_max[7];          /* running maxima of the p, u, v, w, uf, vf, wf increment norms */
    double res_max[5];           /* scaled-residual normalisers: continuity, u, v, w momentum, T */
    double last_scaled;          /* scaled residual of the last iteration */
    uint64_t payload_bytes;      /* bytes of field data after the header */
    uint64_t checksum;           /* FNV-1a (64-bit) of the field data */
} ifl_state_header;

/* Write the handle's state to path (created or replaced). */
int ifl_save_state(ifl_solver *h, const char *path);
/* Read a state file into the handle.  Fails with IFL_ERR_ARG (and leaves the handle unchanged)
 * when dim/nx/ny/nz differ from the handle's, and with IFL_ERR_IO when the file is missing,
 * truncated, of another version or fails its checksum. */
int ifl_load_state(ifl_solver *h, const char *path);

/* Run one stage on the handle's device state (parity tests).  out (optional, >= 8 doubles)
 * receives stage-specific FP64 values:
 *   SMOOTH_U/V/W/P/T : [0] sweeps executed, [1] sum of squared increments of the last sweep
 *   MOM_RESIDUAL   : [0..3] ||b_u - A u||, ||b_v - A v||, ||b_w - A w||, ||b_T - A_T T|| (L2)
 *   PCORR_ASSEMBLE : [0] sum of b'^2
 *   UPDATE         : [0..6] sums of squared increments of p, u, v, w, uf, vf, wf
 *   OUTER          : [0] converged flag, [1] scaled residual */
int ifl_run_stage(ifl_solver *h, int stage, double *out, size_t n_out);

/* Kernels timed by ifl_time_kernel. */
enum {
    IFL_KERNEL_SWEEP = 0,         /* one launch of the p' smoother: sweep_block full red-black
                                     sweeps (2D), one red-black sweep (3D) or one Jacobi sweep */
    IFL_KERNEL_MOM_ASSEMBLE = 1,
    IFL_KERNEL_RHIE_CHOW = 2,
    IFL_KERNEL_PCORR_ASSEMBLE = 3,
    IFL_KERNEL_UPDATE = 4,
    IFL_KERNEL_MOM_SWEEP = 5,     /* one launch of the momentum smoother (u, v and, in 3D, w swept
                                     together on their shared matrix) */
    IFL_KERNEL_TEMP_SWEEP = 6,    /* one launch of the temperature smoother (energy on) */
    IFL_KERNEL_PSOLVE = 7,        /* the whole p' solve of an outer iteration (its sweeps and
                                     refinement cycles, or its multigrid cycles), timed as one unit */
    IFL_KERNEL_COUNT = 8
};

/* Measurement hook (no reference counterpart): launch one kernel `reps` times on the handle's
 * stream, each between its own pair of HIP events, and return the mean device time per launch
 * in microseconds (*launches_per_rep receives 1).  The handle's whole device state is restored
 * afterwards (one copy of its arena is made for that), so a solve continues as if nothing had
 * been timed. */
int ifl_time_kernel(ifl_solver *h, int kernel, int reps, double *avg_us, int *launches_per_rep);

/* ---- Autotuner (no reference counterpart) ----
 * Picks, by measurement on the handle's own case, the settings that the reference fixes at
 * compile time (config.h:67-70) or that only exist here:
 *   IFL_TUNE_LAUNCH : 2D sweep tile, sweep_block and sweep_block_mom (3D: sweep_block 1 or 2),
 *                     by the device time of one outer iteration;
 *   IFL_TUNE_SWEEPS : sweeps per outer iteration (n_iter of u, v, w and of p'), by the drop of
 *                     the scaled residual per second over trial_iters outer iterations.
 * Trials start from the current state (a case that has run fewer than 20 outer iterations is
 * first advanced to 20 with its current settings; those iterations stay) and the state is
 * restored after every trial, so the solve continues from where it was, with the chosen
 * settings written into the handle's configuration (p_refine is set to 0 when the multigrid is
 * chosen).  On an error return the configuration is the one the call started with.
 * trial_iters <= 0 selects 300. */
enum { IFL_TUNE_LAUNCH = 1, IFL_TUNE_SWEEPS = 2,
       IFL_TUNE_PSOLVE = 4,     /* the p' solver and its strength (the configured sweeps, or one
                                   multigrid V-cycle with mg_over 1.0, 1.4 or 1.8), by the drop of the
                                   scaled residual per second over trial_iters outer iterations */
       IFL_TUNE_PSOLVE_TARGET = 8 };  /* the same candidates, by the device time a copy of the case
                                   coarsened 4x per side (2x, or none, on small grids; same domain,
                                   physics, boundary conditions and inlet profiles, averaged) takes from
                                   rest to the handle's scaled-residual target (1e-4 unless the stop rule
                                   is IFL_STOP_SCALED); the handle's own state is not touched.  The
                                   SIMPLE residual oscillates too much for short trials to rank the
                                   coarse-correction factors; the time to the target does.  Wins over
                                   IFL_TUNE_PSOLVE when both are set */

typedef struct ifl_tune_result {
    int32_t tile;              /* chosen tile (0 in 3D) */
    int32_t sweep_block;       /* chosen p' sweeps per launch */
    int32_t n_iter[4];         /* chosen sweeps per outer iteration (u, v, w, p') */
    double us_per_iter;        /* device time of one outer iteration with the choice */
    double decades_per_s;      /* scaled-residual decades per second with the choice (0: not measured) */
    int32_t trials;            /* configurations measured */
    int32_t sweep_block_mom;   /* chosen momentum sweeps per launch (0: the default rule) */
    int32_t p_solver;          /* chosen p' solver (IFL_PSOLVE_*) */
    int32_t mg_cycles;         /* its V-cycles (multigrid) */
    double mg_over;            /* its coarse-correction factor (multigrid) */
} ifl_tune_result;

int ifl_autotune(ifl_solver *h, int what, int trial_iters, ifl_tune_result *out);

/* ---- Discrete adjoint (no reference counterpart) ----
 * Gradient of an objective J of the converged state with respect to the normal-velocity
 * profile of an inlet face and to the viscosity mu, by the fixed-point adjoint of the SIMPLE
 * outer iteration: lambda <- G_x^T lambda + J_x^T (transposed stencils of the momentum, Rhie-
 * Chow, p' and correction stages, the upwind switch and the 8/3 boundary rows, solved with
 * transposed red-black sweeps), then dJ/dtheta = lambda^T G_theta.  Call it after a solve has
 * converged; only the converged state is used.  The adjoint iterate converges at the forward
 * iteration's own rate. */
enum { IFL_OBJ_PRESSURE_DROP = 0,  /* mean p of the first cell layer along x minus the last */
       IFL_OBJ_LINEAR = 1 };       /* sum of weights_u*u + weights_v*v + weights_w*w + weights_p*p */

typedef struct ifl_adjoint_opts {
    int32_t objective;             /* IFL_OBJ_* */
    int32_t max_iter;              /* adjoint iterations at most (0: 200000) */
    double tol;                    /* stop when the scaled adjoint increment is below tol (0: 1e-7) */
    double grad_tol;               /* or when dJ/dmu and (if dJ_dprofile is asked for) the profile
                                      gradient changed by less than grad_tol (relative; the profile
                                      gradient in the L2 norm) over the last check_every iterations
                                      (0: off).  In FP32 the increment stalls near 1e-4 while the
                                      gradient has long settled; this is the practical stop rule. */
    const float *weights[4];       /* IFL_OBJ_LINEAR: weights of u, v, w, p (cells each; NULL = 0) */
    int32_t warm_start;            /* 1: continue from the previous call's lambda */
    int32_t check_every;           /* iterations between host reads of the monitor (0: 64) */
    int32_t tile;                  /* 2D transposed sweeps, tile and sweeps per launch: 0 = by grid
                                      size; 1..6 = 64x16 K4, 32x16 K4, 64x8 K2, 32x8 K4, 64x16 K2,
                                      64x8 K4 (every choice gives the same result, bit for bit) */
    int32_t reserved;              /* 0 */
} ifl_adjoint_opts;

typedef struct ifl_adjoint_report {
    int32_t iterations;            /* adjoint iterations run by this call */
    int32_t converged;
    double J;                      /* objective at the current state */
    double final_scaled;           /* last scaled increment */
    double elapsed_ms;
    double grad_change;            /* grad_tol rule: the larger relative change of dJ/dmu and of the
                                      profile gradient over the last check (-1: not watched) */
} ifl_adjoint_report;

void ifl_adjoint_opts_default(ifl_adjoint_opts *o);
/* dJ_dprofile (optional, n_profile = cells of `face`, which must be an inlet) and dJ_dmu
 * (optional) receive the gradient. */
int ifl_adjoint(ifl_solver *h, const ifl_adjoint_opts *o, int face, double *dJ_dprofile, size_t n_profile,
                double *dJ_dmu, ifl_adjoint_report *report);
/* Host copy of an adjoint component (field = IFL_FIELD_U, V, W, P, UF, VF, WF). */
int ifl_adjoint_lambda(ifl_solver *h, int field, float *dst, size_t n);

/* ---- Batched solves (no reference counterpart) ----
 * A batch runs many independent cases, each with its own grid size, domain, boundary
 * conditions, fluid properties, relaxation, tolerances and stop rule, through one outer
 * iteration graph: every kernel launch covers the workgroups of all cases, so small grids
 * fill the GPU together.  What fixes the launch sequence must agree across the cases: dim,
 * smoother, n_iter[4], p_refine, sweep_block and tile.  A case stops on its own stop rule;
 * its workgroups then return at once.  Every case's result (fields, iteration count,
 * history) equals that of ifl_solve on the same case with the same tile, bit for bit. */
typedef struct ifl_batch ifl_batch;

/* Create n cases from cfgs[0..n-1]. */
int ifl_batch_create(const ifl_config *cfgs, int n, ifl_batch **out);
int ifl_batch_destroy(ifl_batch *b);
int ifl_batch_size(const ifl_batch *b);
/* Handle of case i for field access, face profiles and VTK output (owned by the batch: do not
 * destroy; do not call ifl_solve on it while the batch solves). */
int ifl_batch_case(ifl_batch *b, int i, ifl_solver **out);
/* Run outer iterations until every case has stopped or max_iter (<=0: the largest
 * num_outer_iter) iterations ran.  reports (optional, n entries) receive each case's report;
 * elapsed_ms is the batch's wall time.  Returns IFL_ERR_DIVERGED if any case diverged. */
int ifl_batch_solve(ifl_batch *b, int max_iter, ifl_report *reports);
/* Autotuner of a batch (no reference counterpart): the tile, sweep_block and sweep_block_mom that
 * every case of the batch shares, by the device time of trial_iters (<= 0: 30) outer iterations of
 * the whole batch per candidate (2D red-black; tiles 128x32, 64x16, 64x8, 32x8, 32x4 x sweep_block
 * 2, 4, 5 x sweep_block_mom 0, 4 or 5, at most sweep_block).  Every case is restored after each
 * trial, so the batch solves on from
 * where it was, with the chosen settings; each case still equals its own ifl_solve with the same
 * settings, bit for bit.  On an error return the batch keeps its previous settings. */
int ifl_batch_autotune(ifl_batch *b, int trial_iters, ifl_tune_result *out);

#ifdef __cplusplus
}
#endif

#endif /* INCOMPFLOW_H */

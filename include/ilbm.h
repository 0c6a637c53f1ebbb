/*
 * ilbm.h -- C ABI of the MI355X IB-LBM level (libilbm_hip.so).
 *
 * Plain pointers and sizes only.  Device work is ordered on one HIP stream
 * per level; every entry point returns an ilbm_status (0 = ok) and never
 * aborts.  One host thread per handle.
 *
 * Correspondence with the reference's C++ call surface (iLBMReX, which has
 * no FFI of its own; these are the members a maintainer would route here):
 *
 *   ilbm_level_create        AmrCoreLBM::AmrCoreLBM + ReadParameters +
 *                            MakeNewLevelFromScratch + InitEquilibrium
 *                            (Source/Core/AmrCoreLBM.cpp:190-356, 931-984,
 *                            1188-1507, 2117-2164) and the IBMarkerDF ctor /
 *                            build_markers (Source/IBM/IBMarkerDF.cpp:186-497)
 *   ilbm_level_advance       Evolve's loop body for max_level = 0:
 *                            timeStepWithSubcycling -> AdvancePhiAtLevel
 *                            (Source/Core/AdvanceLBMAtLevel.cpp:20-417), the
 *                            Sum(Rho) reduction (AmrCoreLBM.cpp:378) and
 *                            ComputeIBForce (AmrCoreLBM.cpp:407-410, 456-701)
 *   ilbm_level_get_f /       f_new[lev], macro_new[lev] comps 0..3 and
 *   ilbm_level_get_macro /   forcing[lev] on the valid box (AmrCoreLBM.H:175-180)
 *   ilbm_level_get_forcing
 *   ilbm_level_mass_history  "Sum(Rho)" printed per coarse step (AmrCoreLBM.cpp:412-414)
 *   ilbm_level_force_history force.dat rows (AmrCoreLBM.cpp:636-700)
 *   ilbm_level_num_markers   IBMarkerDF::num_markers (IBMarkerDF.H:222)
 *   ilbm_level_marker_force_sum  IBMarkerDF::last_marker_force_sum (IBMarkerDF.H:223)
 *   ilbm_level_marker_state  marker positions / U / F (IBMarkerDF.H:182-195)
 */
#ifndef ILBM_H
#define ILBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ILBM_ABI_VERSION 3

typedef enum {
  ILBM_OK = 0,
  ILBM_ERR_INVALID = 1,     /* bad argument or case description */
  ILBM_ERR_UNSUPPORTED = 2, /* lattice table / BC type / option not built */
  ILBM_ERR_HIP = 3,         /* HIP runtime error */
  ILBM_ERR_ALLOC = 4,       /* device allocation failed */
  ILBM_ERR_STATE = 5,       /* call not valid in the current state */
  ILBM_ERR_FULL = 6         /* batch admission refused: HBM budget or case slots */
} ilbm_status;

typedef enum { ILBM_STORE_F64 = 0, ILBM_STORE_F32 = 1, ILBM_STORE_F16S = 2 } ilbm_storage;

/* One level of an iLBMReX run; field names follow the reference's inputs
 * keys (geometry.*, amrbc.*, lbm2d/lbm3d.*, lbmPhysicalParameters.*, ibm.*)
 * and MarkerIBParams (Source/IBM/MarkerIBParams.H:8-132). */
typedef struct {
  int dim;                 /* 2 or 3 */
  int n[3];                /* valid cells (n[2] = 1 in 2D) */
  double plo[3];           /* geometry.prob_lo */
  double dx[3];            /* cell size */
  int periodic[3];         /* geometry.is_periodic */
  int bc_lo[3], bc_hi[3];  /* AMReX BCType: 0 int_dir, 2 foextrap, 3 ext_dir, 1001 user_1, 1002 user_2 */
  double bc_lo_val[3][4];  /* rho, ux, uy, uz */
  double bc_hi_val[3][4];
  int q;                   /* lbm*.ndir */
  double cx[27], cy[27], cz[27], w[27]; /* lbm*.dir{x,y,z}, wis/wim */
  double nu, U0;
  int nmac;
  int use_prescribed_force;
  double prescribed_force[3];
  int use_ib;
  double x0, y0, z0, R;
  int coupling_method;     /* 0 explicit, 1 ivc (wu_shu), 2 explicit_diag */
  int delta_type;          /* 0/4: 4-point, 1/3: 3-point */
  int loop_ns;
  double mdf_relax, explicit_diag_eps;
  int geometry_type;       /* 0 cylinder/sphere, 1 box (2D), 2 user pitching section */
  int n_marker;
  double rd;
  double ubx, uby, ubz, omx, omy, omz;
  double ivc_diag_reg;
  int ivc_rebuild_matrix;
  int renormalize_delta;
  int force_eval_method;   /* 0 eulerian, 1 marker, 2 momentum_exchange (falls back to eulerian) */
  double airfoil_chord, airfoil_thickness;
  double pitch_mean, pitch_amplitude, pitch_frequency, pitch_phase, pivot_x;
  double box_lo[3], box_hi[3], box_ds;
  int box_lid_only;
  double box_lid_u[3];
  double box_wall_tol;
  int storage;             /* ilbm_storage */
} ilbm_case_desc;

typedef struct ilbm_level ilbm_level;

/* init_macro: host array [4][n2][n1][n0] (rho, ux, uy, uz; x fastest),
 * the initdata values of the valid region. */
int ilbm_level_create(const ilbm_case_desc* desc, const double* init_macro,
                      ilbm_level** out);
int ilbm_level_destroy(ilbm_level* L);

/* Advances nsteps coarse steps.  With want_state != 0 the last step also
 * writes f_new and macro_new for ilbm_level_get_f / get_macro. */
int ilbm_level_advance(ilbm_level* L, int nsteps, int want_state);
int ilbm_level_synchronize(ilbm_level* L);

int ilbm_level_get_f(ilbm_level* L, double* host_out);       /* [q][cells] */
int ilbm_level_get_macro(ilbm_level* L, double* host_out);   /* [4][cells] */
int ilbm_level_get_forcing(ilbm_level* L, double* host_out); /* [3][cells] */
/* UpdateDerivedMacroFields (AmrCoreLBM.cpp:150-168, visPara.H:10-54) of the
 * exported state: [0] vorticity, [1] pressure (rho - 1)/3, [2][cells]. */
int ilbm_level_get_derived(ilbm_level* L, double* host_out);

/* Sum(Rho) after each step since creation; *n_out = steps recorded. */
int ilbm_level_mass_history(ilbm_level* L, double* host_out, int max_rows, int* n_out);
/* force.dat rows (step, time, 21 columns in 2D / 31 in 3D). */
int ilbm_level_force_history(ilbm_level* L, double* host_out, int max_rows,
                             int* n_rows, int* n_cols);

int ilbm_level_num_markers(ilbm_level* L, int* out);
int ilbm_level_marker_dv(ilbm_level* L, double* out);
int ilbm_level_marker_state(ilbm_level* L, double* xyz, double* uvw, double* fxyz);
int ilbm_level_marker_force_sum(ilbm_level* L, double out[3]);
int ilbm_level_marker_torque(ilbm_level* L, double out[3]);

/* Checkpoint / restart of a level's dynamic state (populations, ghost and
 * boundary layers, IB band macros and force, time-centred force state,
 * clock) into a host buffer, restored into a level created from the same
 * case description.  The reference's checkpoint writer is stubbed out
 * (AmrCoreLBM.cpp:1932-2115); this is a compact per-case dump for
 * preemption.  Histories restart empty after a restore. */
int ilbm_level_checkpoint_size(ilbm_level* L, int64_t* bytes);
int ilbm_level_checkpoint(ilbm_level* L, void* host_buf, int64_t bytes);
int ilbm_level_restore(ilbm_level* L, const void* host_buf, int64_t bytes);

/* Level facts: [0] tau, [1] dt, [2] time, [3] steps, [4] band cells,
 * [5] storage bytes per population, [6] layout: 0 wrapped (all periodic),
 * 1 padded with ghost slabs and 64-bit offsets, 2 padded with 32-bit offsets
 * (populations and slabs under 4 GiB), 3 as 2 with non-periodic faces in x
 * only, 4 as 3 with the x-face ghost slabs written by the stream-collide
 * kernel itself (ext_dir / user_2 / foextrap x faces), [7] stream-collide
 * blocks. */
int ilbm_level_info(ilbm_level* L, double out[8]);

/* Timing hooks for measurement: average device time (ms) of the fused
 * stream-collide kernel over the last advance() call, measured with HIP
 * events on the level's stream, and the number of launches it covers. */
int ilbm_level_kernel_time(ilbm_level* L, double* avg_ms, int* launches);
int ilbm_level_set_timing(ilbm_level* L, int enable);
/* IB level of at least 2^20 cells: stream-collide the cells farther than
 * one cell from the IB band on a second stream while the IB forcing runs
 * (default on; off while timing is on).  The results are the same either
 * way; Sum(Rho) is summed in another order.  No counterpart in the
 * reference. */
int ilbm_level_set_overlap(ilbm_level* L, int enable);

/* Measurement helper (no counterpart in the reference): average time of a
 * plain device copy of `bytes` bytes over `iters` launches, used to state
 * achievable HBM bandwidth and to calibrate profiler byte counters. */
int ilbm_copy_probe(int64_t bytes, int iters, double* avg_ms);

/* ------------------------------------------------------------------------
 * Batched cases (no counterpart in the reference, which runs one case per
 * process).  Many independent cases stay resident on the device and advance
 * together: each kernel of a step is launched once for all cases of a
 * lattice / storage class, and the step sequence is replayed from a
 * hipGraph of `graph_steps` steps.  Cases are admitted while others run
 * (continuous batching) as long as their device footprint fits the budget;
 * each stops at its own step count and keeps its histories until removed.
 * Results of a case do not depend on which other cases share the batch.
 * Moving bodies (geometry_type 2) are not batched.
 * ------------------------------------------------------------------------ */
typedef struct ilbm_batch ilbm_batch;

int ilbm_batch_create(int64_t hbm_budget_bytes, int max_cases, int graph_steps,
                      ilbm_batch** out);
int ilbm_batch_destroy(ilbm_batch* B);
/* Device bytes a case of nsteps steps is charged at admission. */
int ilbm_batch_footprint(const ilbm_case_desc* desc, int nsteps, int64_t* bytes);
/* ILBM_ERR_FULL when the case does not fit; nothing is then allocated. */
int ilbm_batch_add(ilbm_batch* B, const ilbm_case_desc* desc, const double* init_macro,
                   int nsteps, int* case_id);
/* Advances every unfinished case by min(max_steps, smallest remaining
 * count) steps (*steps_done); histories of those steps are copied back. */
int ilbm_batch_advance(ilbm_batch* B, int max_steps, int use_graph, int* steps_done);
int ilbm_batch_finished(ilbm_batch* B, int* ids, int max_ids, int* n);
int ilbm_batch_progress(ilbm_batch* B, int case_id, int* done, int* target);
int ilbm_batch_force_history(ilbm_batch* B, int case_id, double* host_out, int max_rows,
                             int* n_rows, int* n_cols);
int ilbm_batch_mass_history(ilbm_batch* B, int case_id, double* host_out, int max_rows,
                            int* n_out);
int ilbm_batch_remove(ilbm_batch* B, int case_id);
/* Preemption: dump a resident case, and admit a case from such a dump
 * (same case description) for nsteps more steps. */
int ilbm_batch_checkpoint_size(ilbm_batch* B, int case_id, int64_t* bytes);
int ilbm_batch_checkpoint(ilbm_batch* B, int case_id, void* host_buf, int64_t bytes);
int ilbm_batch_add_from_checkpoint(ilbm_batch* B, const ilbm_case_desc* desc, const void* host_buf,
                                   int64_t bytes, int nsteps, int* case_id);
/* Tile width (cells along x per workgroup) the autotuner chose for a case;
 * 0: linear tiles of 256 consecutive cells in x-fastest order (no padding
 * of rows that are not a multiple of the width). */
int ilbm_batch_case_tile(ilbm_batch* B, int case_id, int* tile_bx);
/* [0] resident cases, [1] unfinished cases, [2] bytes charged, [3] budget,
 * [4] graph launches, [5] graph captures, [6] steps launched without a
 * graph, [7] lattice updates done. */
int ilbm_batch_info(ilbm_batch* B, double out[8]);
/* Free and total device memory (hipMemGetInfo), for admission budgets. */
int ilbm_device_memory(int64_t* free_bytes, int64_t* total_bytes);

/* ------------------------------------------------------------------------
 * AMR hierarchy: level 0 is the case's domain, levels 1..nlev-1 (nlev <= 3)
 * one box each at refinement ratio 2, advanced with subcycling.  Replaces
 * timeStepWithSubcycling / AdvancePhiAtLevel for lev > 0 and the
 * coarse-fine transfer of the reference (Source/Core/AmrCoreLBM.cpp:
 * FillPatchMesoscopic lev > 0 :1545-1627, FillPatchMacro :2166-2198,
 * GetDataMacro :2239-2269, AverageDownTo :1520-1526,
 * MakeNewLevelFromCoarse :751-830, UpdateDerivedMacroFields :150-168).
 * The boxes stay fixed unless ilbm_amr_set_regrid switches regridding on
 * (below).  box_lo / box_hi hold, for
 * level l >= 1, its box in level-l cell indices at [3*(l-1) + d]; a box
 * covers whole coarse cells and keeps 3 coarse cells from its coarse box.
 * The immersed body lives on the finest level; the initial state
 * (rho, ux, uy, uz) is uniform on every level.
 * ------------------------------------------------------------------------ */
typedef struct ilbm_amr ilbm_amr;

int ilbm_amr_create(const ilbm_case_desc* desc, int nlev, const int* box_lo, const int* box_hi,
                    int regrid_int, const double init_state[4], ilbm_amr** out);
int ilbm_amr_destroy(ilbm_amr* A);
/* nsteps coarse steps; want_state also exports level 0 on the last one
 * (refined levels always export their last substep). */
int ilbm_amr_advance(ilbm_amr* A, int nsteps, int want_state);
int ilbm_amr_level_cells(ilbm_amr* A, int lev, int64_t* cells);
int ilbm_amr_get_f(ilbm_amr* A, int lev, double* host_out);      /* [q][box cells] */
int ilbm_amr_get_macro(ilbm_amr* A, int lev, double* host_out);  /* [4][box cells] */
/* UpdateDerivedMacroFields of an exported level (AmrCoreLBM.cpp:150-168):
 * vorticity and pressure from FillPatchMacro with one ghost, [2][box cells] */
int ilbm_amr_get_derived(ilbm_amr* A, int lev, double* host_out);
/* Sum(Rho) of level 0 and the finest level's force.dat rows per coarse step */
int ilbm_amr_mass_history(ilbm_amr* A, double* host_out, int max_rows, int* n_out);
int ilbm_amr_force_history(ilbm_amr* A, double* host_out, int max_rows, int* n_rows, int* n_cols);
int ilbm_amr_num_markers(ilbm_amr* A, int* out);
/* [0] levels, [1] coarse steps, [2] time, [3..5] cells of levels 0..2,
 * [6] tau of the finest level, [7] tau of level 0 */
int ilbm_amr_info(ilbm_amr* A, double out[8]);

/* Regridding (timeStepWithSubcycling's regrid, AmrCoreLBM.cpp:1692-1720):
 * every regrid_int steps of level l < finest, ErrorEst tags level l
 * (AmrCoreLBM.cpp:986-1180: |vorticity| > thresholdRatio[l] * max
 * vorticity, the level-set band of a cylinder / sphere, the body
 * refinement region :56-104, cleared in the no-refine x bands), the tags
 * are grown by n_error_buf and each finer level gets ONE box: the bounding
 * box of its parent's tags (and of its own child's footprint grown by 3),
 * aligned to blocking_factor and kept 3 cells inside its parent; AMReX's
 * Berger-Rigoutsos clustering into several boxes is not reproduced.  A
 * changed level is remade as RemakeLevel does (:835-914: macros from the
 * old level where it was valid and coarse interpolation elsewhere,
 * f = feq(macros), new marker backend on the finest level). */
typedef struct {
  double threshold_ratio[4];   /* amrvorth.thresholdRatio per level */
  double no_refine_xlo, no_refine_xhi;                      /* amrvorth.no_refine_xlo / xhi */
  double refine_upstream, refine_downstream, refine_cross;  /* ibm.refine_* (cylinder / sphere) */
  int blocking_factor;         /* amr.blocking_factor, fine cells, even */
  int n_error_buf;             /* amr.n_error_buf */
} ilbm_regrid_params;
/* before the first ilbm_amr_advance; ILBM_ERR_STATE afterwards */
int ilbm_amr_set_regrid(ilbm_amr* A, const ilbm_regrid_params* params);
/* current boxes of levels 1..nlev-1 (level-l indices, [3*(l-1) + d]) and the
 * number of regrids that changed a box */
int ilbm_amr_boxes(ilbm_amr* A, int* box_lo, int* box_hi, int* n_regrids);

/* ------------------------------------------------------------------------
 * Per-box entry points on caller-owned device memory in the reference's
 * layout: an AMReX Array4 of FP64 (one contiguous array per component over
 * the box grown by its ghost cells, x fastest; AmrCoreLBM.H:276).  One call
 * per MFIter box of AdvancePhiAtLevel, in the reference's order; AMReX's
 * ParallelCopy / FillBoundary between boxes stay with the caller
 * (ilbm_box_fill_periodic covers a level of one box).  The domain is
 * [0, n-1] of desc (the level's cell counts); tau = 3 nu / dx + 1/2.
 * `stream` is a hipStream_t (NULL: the default stream).
 * ------------------------------------------------------------------------ */
typedef enum { ILBM_FAB_F64 = 0, ILBM_FAB_F32 = 1 } ilbm_fab_dtype;

/* amrex::Array4: element (i, j, k, n) is at
 * p[(i - lo0) + (j - lo1) jstride + (k - lo2) kstride + n nstride].
 * Strides of 0 mean the dense Array4 strides of the box (x fastest, one
 * contiguous array per component); a component alias (MultiFab make_alias)
 * is p advanced by that component's offset with ncomp = 1.  The per-box
 * path computes in FP64 (amrex::Real of the reference's PRECISION=DOUBLE
 * build) and takes ILBM_FAB_F64 fabs; other dtypes are refused with
 * ILBM_ERR_UNSUPPORTED. */
typedef struct {
  void* p;            /* device pointer to component 0 at cell lo */
  int lo[3], hi[3];   /* fab box (valid box grown by the ghost cells) */
  int ncomp;
  int dtype;          /* ilbm_fab_dtype */
  int64_t jstride, kstride, nstride;   /* elements; 0: dense */
} ilbm_fab;

/* InitEquilibrium (AmrCoreLBM.cpp:2117-2164): f = feq(rho, u) on [lo, hi] */
int ilbm_box_equilibrium(const ilbm_case_desc* desc, const ilbm_fab* macro, ilbm_fab* f,
                         const int lo[3], const int hi[3], void* stream);
/* FillBoundary with the domain's periodicity for a level of one box */
int ilbm_box_fill_periodic(const ilbm_case_desc* desc, ilbm_fab* fab, const int valid_lo[3],
                           const int valid_hi[3], void* stream);
/* PhysBCFunct<GpuBndryFuncFab<mesoscopicBcFill / macroBcFill>> on every fab
 * cell outside the domain: FilccCell foextrap, then the user functor
 * (Source/BC/bc_fill.H:263-347), faces before edges before corners */
int ilbm_box_fill_physical(const ilbm_case_desc* desc, ilbm_fab* fab, int mesoscopic, void* stream);
/* collide / collide_forced on the valid box, in place on f, with feq of
 * macro (rho, ux, uy, uz) and the force fab (NULL: no force)
 * (Source/LBM/collide_forced.H:10-51, AdvanceLBMAtLevel.cpp:196-240) */
int ilbm_box_collide(const ilbm_case_desc* desc, ilbm_fab* f, const ilbm_fab* macro,
                     const ilbm_fab* force, const int valid_lo[3], const int valid_hi[3], void* stream);
/* halfway bounce-back into the user_1 ghost cells (AdvanceLBMAtLevel.cpp:265-373) */
int ilbm_box_bounce_back(const ilbm_case_desc* desc, ilbm_fab* f, const int valid_lo[3],
                         const int valid_hi[3], void* stream);
/* stream + calculateMacroForcing on the valid box (stream.H:9-22,
 * calculateMacroForcing.H:9-32) */
int ilbm_box_stream_macro(const ilbm_case_desc* desc, const ilbm_fab* f_in, ilbm_fab* f_out,
                          ilbm_fab* macro_out, const ilbm_fab* force, const int valid_lo[3],
                          const int valid_hi[3], void* stream);
/* visPara (Source/LBM/visPara.H:10-54): vorticity into comp 4 and pressure
 * into comp 5 of macro_out from a macro patch with one ghost cell */
int ilbm_box_derived(const ilbm_case_desc* desc, const ilbm_fab* macro_patch, ilbm_fab* macro_out,
                     const int valid_lo[3], const int valid_hi[3], void* stream);
/* ParallelCopy of [lo, hi] between two fabs of the same index space, and
 * setVal of a whole fab (the copy / fill parts of FillPatchSingleLevel and
 * the force scratch's setVal, AdvanceLBMAtLevel.cpp:43-56) */
int ilbm_box_copy(const ilbm_fab* src, ilbm_fab* dst, const int lo[3], const int hi[3], int ncomp,
                  void* stream);
int ilbm_box_setval(ilbm_fab* fab, double value, void* stream);
/* Device memory for callers without an allocator of their own (hipMalloc /
 * hipMemcpy / hipDeviceSynchronize); measurement and test plumbing. */
int ilbm_device_alloc(int64_t bytes, void** out);
int ilbm_device_free(void* p);
int ilbm_copy_to_device(void* dst, const void* src, int64_t bytes);
int ilbm_copy_to_host(void* dst, const void* src, int64_t bytes);
int ilbm_device_synchronize(void);
/* sum of one component over a box (MultiFab::sum of one box; Sum(Rho),
 * the force integral of ComputeIBForce), fixed order; synchronizes */
int ilbm_box_sum(const ilbm_fab* fab, int comp, const int lo[3], const int hi[3], double* host_out,
                 void* stream);
/* ------------------------------------------------------------------------
 * Coarse-fine transfer between two levels at refinement ratio 2, on
 * caller-owned fabs (level indices; the fine box keeps its ghost cells
 * inside the fine domain, so every fine ghost is a coarse-fine ghost).
 * ------------------------------------------------------------------------ */
/* FillPatchMesoscopic / FillPatchMacro for lev > 0 (AmrCoreLBM.cpp:
 * 1545-1627, 2180-2197): every cell of the fine fab outside
 * [valid_lo, valid_hi] gets the coarse macros at `time` -- GetDataMacro's
 * choice of crse_old (at t_old), crse_new (at t_new) or their linear mix
 * (:2239-2269) -- interpolated by cell_cons_interp over `ncomp` components
 * with one limiter per direction; macro_fine (may be NULL) receives them,
 * f_fine (may be NULL) receives their equilibrium (:1605-1620).  desc is the
 * fine level (its n, lattice). */
int ilbm_box_cf_fill(const ilbm_case_desc* desc, const ilbm_fab* crse_old, const ilbm_fab* crse_new, double t_old,
                     double t_new, double time, ilbm_fab* f_fine, ilbm_fab* macro_fine, int ncomp,
                     const int valid_lo[3], const int valid_hi[3], void* stream);
/* InterpFromCoarseLevel of MakeNewLevelFromCoarse / FillCoarsePatchMacro
 * (AmrCoreLBM.cpp:751-830, 2214-2237): cell_cons_interp of crse over the
 * fine cells [lo, hi] */
int ilbm_box_interp_from_coarse(const ilbm_case_desc* desc, const ilbm_fab* crse, ilbm_fab* macro_fine, int ncomp,
                                const int lo[3], const int hi[3], void* stream);
/* AverageDownTo (AmrCoreLBM.cpp:1520-1526): the coarse cells under the fine
 * box [fine_lo, fine_hi] (whole coarse cells) get the mean of their 2^dim
 * children, f (q comps) and macros (nmac comps) */
int ilbm_box_average_down(const ilbm_case_desc* desc, const ilbm_fab* f_fine, const ilbm_fab* macro_fine,
                          const int fine_lo[3], const int fine_hi[3], ilbm_fab* f_crse, ilbm_fab* macro_crse,
                          int nmac, void* stream);

/* ------------------------------------------------------------------------
 * IBMarkerDF on caller-owned fabs (Source/IBM/IBMarkerDF.H:68-91).
 * ------------------------------------------------------------------------ */
typedef struct ilbm_ib ilbm_ib;
/* IBMarkerDF(geom, dm, ba, x0, y0, z0, R, par) + build_markers
 * (IBMarkerDF.cpp:186-497): desc is the level (n = its domain, dx, BCs, the
 * ibm.* parameters); [valid_lo, valid_hi] its box (NULL: the domain).  A box
 * inside the domain is a refined level's: its faces are coarse-fine and the
 * marker stencils must stay inside it (ILBM_ERR_UNSUPPORTED otherwise). */
int ilbm_ib_create(const ilbm_case_desc* desc, const int valid_lo[3], const int valid_hi[3], ilbm_ib** out);
int ilbm_ib_destroy(ilbm_ib* ib);
/* update_forcing(rhocc, ucc, vcc, wcc, Fx, Fy, Fz, dt, time)
 * (IBMarkerDF.cpp:1522-1917): reads rho, u, v, w (one component each; w
 * may be NULL in 2D) on the band around the marker stencils, as
 * FillPatchMacro leaves them, and writes F = dt * rho * EulerForceTotal on
 * every cell of the valid box of Fx, Fy, Fz (Fz may be NULL in 2D), zero
 * outside the band.  Periodic images are taken from the valid cells and
 * physical ghost values from the case's BCs (macroBcFill), which is what
 * FillPatchMacro puts there.  The reference's prev_F* arguments are unused
 * by it and have no counterpart. */
int ilbm_ib_update_forcing(ilbm_ib* ib, const ilbm_fab* rho, const ilbm_fab* u, const ilbm_fab* v,
                           const ilbm_fab* w, ilbm_fab* fx, ilbm_fab* fy, ilbm_fab* fz, double dt, double time,
                           void* stream);
int ilbm_ib_num_markers(ilbm_ib* ib, int64_t* out);          /* num_markers */
int ilbm_ib_force_sum(ilbm_ib* ib, double out[3]);           /* last_marker_force_sum */
/* sum over markers of the moments r x (F dv) about the body centre that the
 * last spreading stored (IBMarkerDF.cpp:1500-1503; stored, never reduced,
 * by the reference) */
int ilbm_ib_torque(ilbm_ib* ib, double out[3]);
/* positions, U and F per marker, [nm][3] each (any may be NULL) */
int ilbm_ib_marker_state(ilbm_ib* ib, double* xyz, double* uvw, double* fxyz);

/* ilbm_box_sum without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * bits as ilbm_box_sum; the partials live in a fixed 1 MiB device array per
 * GPU, so boxes above ilbm_box_sum_noalloc_max_cells cells (33,554,176) are
 * rejected with ILBM_ERR_INVALID, and calls that overlap on one GPU race. */
int ilbm_box_sum_noalloc(const ilbm_fab* fab, int comp, const int lo[3], const int hi[3], double* host_out,
                         void* stream);
int ilbm_box_sum_noalloc_max_cells(int64_t* cells);

const char* ilbm_status_string(int status);
/* Text of the last error raised on the calling thread. */
const char* ilbm_last_error(void);
int ilbm_abi_version(void);
/* sizeof(ilbm_case_desc) as compiled, for binding layout checks. */
int ilbm_case_desc_size(void);

#ifdef __cplusplus
}
#endif
#endif

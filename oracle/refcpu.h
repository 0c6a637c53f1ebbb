/*
 * refcpu -- CPU restatement of iLBMReX's per-timestep IB-LBM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: it is used
 * by tests/, by __graft_entry__.smoke() (as the checker) and by the
 * cpu_baseline leg of bench.py.  The product (ilbmrex_amd) never links,
 * loads or calls it.
 *
 * It restates, in plain C99 and FP64, the reference's single-level step
 * (Source/Core/AdvanceLBMAtLevel.cpp:20-417), its per-cell kernels
 * (the Source/LBM headers), the FillPatch ghost semantics of the BC functors
 * (Source/BC/bc_fill.H:12-347) together with the AMReX behaviour they sit
 * in (ParallelCopy, FilccCell foextrap, face->edge->corner fill order,
 * FillBoundary, SumBoundary), the marker direct-forcing backend
 * (Source/IBM/IBMarkerDF.cpp:70-1917) and the force evaluation
 * (Source/Core/AmrCoreLBM.cpp:456-701).  Expression forms and operation
 * order follow the reference so that FP64 results are reproducible bit for
 * bit by an implementation that evaluates the same operations.
 *
 * One box covers the whole level (AMReX max_grid_size >= domain); the
 * reference's multi-box layouts give the same values up to the summation
 * order of reductions and of SumBoundary.
 */
#ifndef ILBM_REFCPU_H
#define ILBM_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AMReX BCType codes as used in the reference inputs (amrbc.bc_lo/hi). */
enum { RC_INT_DIR = 0, RC_FOEXTRAP = 2, RC_EXT_DIR = 3, RC_USER_1 = 1001, RC_USER_2 = 1002 };

/* MarkerIBParams (Source/IBM/MarkerIBParams.H:8-132). */
enum { RC_GEOM_CYLINDER = 0, RC_GEOM_BOX = 1, RC_GEOM_USER = 2 };
enum { RC_COUPLING_EXPLICIT = 0, RC_COUPLING_IVC = 1, RC_COUPLING_EXPLICIT_DIAG = 2 };
enum { RC_FORCE_EVAL_EULERIAN = 0, RC_FORCE_EVAL_MARKER = 1, RC_FORCE_EVAL_ME = 2 };

typedef struct {
  /* geometry */
  int dim;             /* 2 or 3 (AMREX_SPACEDIM) */
  int n[3];            /* valid cells per dim (n[2] = 1 in 2D) */
  double plo[3];       /* geometry.prob_lo */
  double dx[3];        /* cell size per dim */
  int periodic[3];
  int bc_lo[3], bc_hi[3];
  double bc_lo_val[3][4]; /* rho, ux, uy, uz per dim (BCVals) */
  double bc_hi_val[3][4];
  /* lattice: directions stored as Real, like the reference tables */
  int q;
  double cx[27], cy[27], cz[27], w[27];
  /* physics */
  double nu;
  double U0;           /* lbmPhysicalParameters.U0 (Cd normalisation) */
  int nmac;            /* 6 */
  int use_prescribed_force;
  double prescribed_force[3];
  /* immersed boundary */
  int use_ib;          /* ibm.use_cylinder && ibm.method == marker */
  double x0, y0, z0, R;
  int coupling_method;
  int delta_type;
  int loop_ns;
  double mdf_relax;
  double explicit_diag_eps;
  int geometry_type;
  int n_marker;
  double rd;
  double ubx, uby, ubz, omx, omy, omz;
  double ivc_diag_reg;
  int ivc_rebuild_matrix;
  int renormalize_delta;
  int force_eval_method;
  /* user-defined marker geometry (pitching NACA section); used when
   * geometry_type == RC_GEOM_USER.  Kinematics follow
   * Examples/Pitching_airfoil/IBMUserDefinedGeometry.H:144-175 with the
   * compile-time constants made run-time parameters. */
  double airfoil_chord, airfoil_thickness;
  double pitch_mean, pitch_amplitude, pitch_frequency, pitch_phase;
  double pivot_x;      /* pivot offset along the chord from the leading edge */
  /* box geometry (2D cavity walls) */
  double box_lo[3], box_hi[3], box_ds;
  int box_lid_only;
  double box_lid_u[3];
  double box_wall_tol;
} rc_case;

typedef struct rc_level rc_level;

/* Creates a level and initialises it as MakeNewLevelFromScratch +
 * InitEquilibrium do (Source/Core/AmrCoreLBM.cpp:931-984, 2117-2164):
 * macro_new = macro_old = given (rho,u,v,w) on the valid region, f = feq.
 * init_macro is [4][n2][n1][n0] (x fastest). */
rc_level *rc_create(const rc_case *c, const double *init_macro);
void rc_destroy(rc_level *L);

/* One coarse step: timeStepWithSubcycling(0) at max_level = 0, i.e.
 * AdvancePhiAtLevel, then Evolve's mass sum and (if use_ib) ComputeIBForce.
 * Source/Core/AmrCoreLBM.cpp:366-414, 1690-1754. */
void rc_step(rc_level *L);
void rc_advance(rc_level *L, int nsteps);

/* Outputs after the last step. */
double rc_mass_sum(const rc_level *L);           /* Sum(Rho) of macro_new */
int rc_get_f(const rc_level *L, double *out);     /* f_new valid [q][k][j][i] */
int rc_get_macro(const rc_level *L, double *out); /* macro_new comps 0..3 valid */
int rc_get_forcing(const rc_level *L, double *out); /* forcing[lev] (3 comps) valid */
/* ComputeIBForce record: step, time, then the 21 2D columns / 31 3D columns
 * of force.dat in file order (Source/Core/AmrCoreLBM.cpp:651-685). */
int rc_force_record(const rc_level *L, double *out, int maxn);
int rc_num_markers(const rc_level *L);
void rc_marker_state(const rc_level *L, double *xyz, double *uvw, double *fxyz);
void rc_last_marker_force_sum(const rc_level *L, double out[3]);
double rc_marker_dv(const rc_level *L);
double rc_tau(const rc_level *L);

/* ---- per-function entry points (unit checks) ---- */
double rc_feq(double w, double cx, double cy, double cz, double rho, double u,
              double v, double wv, int dim);
/* collide_forced on one cell's populations (in place). */
void rc_collide_forced(double *f, const double *feq, int q, double tau,
                       const double *w, const double *cx, const double *cy,
                       const double *cz, double rho, double ux, double uy,
                       double uz, double fx, double fy, double fz, int dim);
double rc_delta(double xf, double xp, double h, int type);
int rc_delta_type_to_kernel(int delta_type);
double rc_overlap_1d(double xa, double xb, double h, int kernel_type,
                     double plo, int ilo, int ihi);
int rc_lu_factorize(double *a, int *pivot, int n, double tiny);
void rc_lu_solve(const double *a, const int *pivot, const double *rhs,
                 double *x, int n);
/* Marker set as build_markers + update_lagrangian_marker(time) produce it. */
/* AMR hierarchy: nlev levels, one box per level at ratio 2 (level l box
 * given in level-l cell indices for l >= 1: box_lo/box_hi[3*(l-1)+d]),
 * subcycling, equilibrium coarse-fine ghosts, average-down, derived
 * fields every regrid_int steps.  init_macro: level-0 [4][cells]. */
typedef struct rc_amr rc_amr;
rc_amr *rc_amr_create(const rc_case *c, int nlev, const int *box_lo, const int *box_hi,
                      int regrid_int, const double *init_macro);
void rc_amr_destroy(rc_amr *A);
void rc_amr_step(rc_amr *A);
void rc_amr_advance(rc_amr *A, int nsteps);
/* Regrid (ErrorEst / RemakeLevel with one box per level): per-level
 * amrvorth.thresholdRatio, amrvorth.no_refine_xlo/hi, ibm.refine_*,
 * amr.blocking_factor and amr.n_error_buf. */
typedef struct {
  double threshold_ratio[4];
  double no_refine_xlo, no_refine_xhi;
  double refine_upstream, refine_downstream, refine_cross;
  int blocking_factor;
  int n_error_buf;
} rc_regrid;
void rc_amr_set_regrid(rc_amr *A, const rc_regrid *p);
int rc_amr_boxes(const rc_amr *A, int *lo, int *hi);
int rc_amr_regrids(const rc_amr *A);
void rc_amr_regrid_now(rc_amr *A, int lbase);
rc_level *rc_amr_level(rc_amr *A, int l);
int rc_get_macro_all(const rc_level *L, double *out);
void rc_cons_interp_test(const double *crse, const int *clo, const int *chi, int dim, int ncomp,
                         const int *fiv, double *out);

int rc_build_markers(const rc_case *c, double time, double *xyz, int maxn,
                     double *dv_out);

/* Stage probes of IBMarkerDF::update_forcing (markers placed at `time`):
 * interpolation of a caller-given EulerVel ([3][valid box grown by 2]) into
 * marker U [nm][3], and spreading of caller-given marker forces [nm][3] into
 * EulerForce [3][valid cells] with the moments r x (F dv) [nm][3]. */
int rc_ib_interp_probe(rc_level *L, double time, const double *ev, double *U_out);
int rc_ib_spread_probe(rc_level *L, double time, const double *F, double *efi_out, double *mom_out);

/* Planted mistakes, for tests that show a pin fails on a plausible error:
 * 1 ext_dir non-equilibrium sign, 2 spreading without dv, 3 export without
 * dt, 4 interpolation without the cell volume, 5 Cd without the 1/2,
 * 6 explicit forcing sign, 7 moment sign, 8 forced macros without F/2.
 * 0 restores the restatement.  Process-global. */
void rc_set_mutation(int id);

#ifdef __cplusplus
}
#endif
#endif

/*
 * adh_sw2_hip.h -- C ABI of the MI355X (gfx950) SW2 Newton-step library.
 *
 * The library replaces the reference's legacy serial SW2 Newton path:
 * the five callbacks fe_newton receives (src/fe/fe_newton.c:220-224, bound in
 * src/fe/sw2/fe_sw2_solve.c:45-49) and the linear-solve pair fe_newton calls
 * (src/fe/fe_newton.c:547-557).  The reference passes SSUPER_MODEL*; a C shim
 * in the reference forwards the arrays those structs
 * hold -- plain pointers and sizes -- to the entry points below.  All state
 * and the linear system stay resident in GPU memory between calls; host
 * arrays are written only where an output pointer is non-NULL.
 *
 * Layouts are the reference's: per-node velocity as interleaved (x,y)
 * (SVECT2D vel[N], src/structs/ssw_2d.h), 3 DOF per node ordered (u, v, h)
 * with equations (x-momentum, y-momentum, continuity)
 * (src/fe/sw2/fe_sw2_inc.c:47-51), 3x3 diagonal blocks diagonal[9*N]
 * and off-diagonal blocks in SPARSE_VECT row order (src/structs/ssolve.h:16-21,
 * block value order src/fe/fe_assemble_matrix.c:200-208).
 *
 * Return codes: ADH_SW2_OK (0) or a negative error; the solve and Newton
 * calls additionally report the reference's YES (1) / NO (-3) verdicts in
 * their out-parameters.  adh_sw2_hip_last_error() returns a message for the
 * last failure on the calling thread.  Handles are not thread-safe; one
 * handle per model, any number of handles per process.
 */
#ifndef ADH_SW2_HIP_H
#define ADH_SW2_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ADH_SW2_OK 0
#define ADH_SW2_EINVAL -1
#define ADH_SW2_EHIP -2
#define ADH_SW2_ENOMEM -4
#define ADH_SW2_EUNSUPPORTED -5

#define ADH_SW2_YES 1   /* src/include/define.h:169 */
#define ADH_SW2_NO (-3) /* src/include/define.h:170 */

/* Mesh, strings and model constants: the fields of SGRID, SELEM_2D/1D,
 * STR_VALUE, SMAT_SW and SMODEL that the SW2 path reads (see the
 * reference's src/structs).  Element geometry (djac, grad_shp, 1D djac and
 * outward normal) may be passed as computed by the reference's sgrid set-up
 * or left NULL, in which case the library computes it with the same formulas
 * (src/elem/element_djac_nrml_gradPhi.c:24-36, 193-274;
 * src/structs/selem_1d.c:118-160). */
typedef struct {
    int nnodes, nelems2d, nelems1d, nmat;
    const double *x, *y, *z;          /* [N] node coordinates (SNODE.x/y/z) */
    const int *node_bc;               /* [N] bc_flag of the node's node string, 0 = none */
    const double *node_bc_val;        /* [3N] Dirichlet u, v, h values (NULL if no Dirichlet nodes) */
    const int *e2d;                   /* [3E] triangle node ids (SELEM_2D.nodes) */
    const int *e2d_mat;               /* [E] material (SELEM_2D.mat) */
    const double *e2d_ks;             /* [E] equivalent roughness height of the element's string */
    const int *e2d_src_flag;          /* [E] 1 if the element string is BCT_WATER_SOURCE */
    const double *e2d_src;            /* [E] water source series value */
    const double *e2d_djac;           /* [E] optional: SELEM_2D.djac */
    const double *e2d_grad;           /* [6E] optional: SELEM_2D.grad_shp (x0,y0,x1,y1,x2,y2) */
    const int *e1d;                   /* [2B] boundary edge node ids (SELEM_1D.nodes) */
    const int *e1d_adj;               /* [B] triangle adjacent to the edge (outward-normal test) */
    const int *e1d_bc;                /* [B] 10 NB VEL, 9 NB OTW, 4 NB OUT */
    const double *e1d_val;            /* [B] series value: VEL flux series / OTW tail elevation */
    const double *e1d_djac;           /* [B] optional: SELEM_1D.djac */
    const double *e1d_nrml;           /* [2B] optional: SELEM_1D.nrml */
    const int *mat_turb;              /* [M] 0 none, 1 EVS, 2 EEV (SMAT_SW.EVSF/EEVF) */
    const int *mat_eev_mode;          /* [M] */
    const double *mat_eev_coef;       /* [M] */
    const double *mat_ev;             /* [3M] user eddy viscosity xx, yy, xy */
    const double *mat_cor_speed;      /* [M] get_coriolis_angular_speed(coriolis) */
    double gravity, dt, tau_temporal, tau_pg, drying_lower_limit, drying_upper_limit, viscosity;
    /* Winds (MP WND STR / ATT, SERIES WIND): surface stress on fully wet
     * triangles, src/fe/sw2/fe_sw2_body_resid.c:674-719 with
     * swind_elem2d_get_local_stress (src/structs/smeteor.c:241-308).
     * wind = 0 switches them off (the node values then need not be set). */
    int wind;
    const int *mat_wind_flag;         /* [M] stress law (SMAT_SW.wind_flag): only 0, the node value is the
                                         stress, is supported; NULL = 0 */
    const double *mat_windatt;        /* [M] attenuation (SMAT_SW.windatt); NULL = 1.0 */
    double density;                   /* MP RHO (SMODEL.density) */
} adh_sw2_mesh;

/* Solver settings.  Defaults (adh_sw2_hip_default_opts) reproduce the
 * reference's single-model settings (src/solver/solv_initialize.c:54-66,
 * src/solver/solv_bcgstab.c:171-188). */
typedef struct {
    double tol_nonlin;   /* IP NTL (absolute max-norm) */
    double inc_nonlin;   /* IP ITL */
    int max_nonlin_it;   /* IP NIT */
    int max_lin_it;      /* 500 */
    int max_cuts;        /* 5 line-search cuts */
    double lin_rtol;     /* BiCGSTAB stop: ||r||inf <= lin_atol + lin_rtol*||M^-1 b||inf */
    double lin_atol;
    int precision;       /* 0: FP64; 1: FP32 element blocks, matrix and preconditioner, FP64 Krylov vectors;
                            2: FP32 matrix and Krylov; 1 and 2 form the FD Jacobian in FP64 and keep the
                            residual, update and Newton tests in FP64 */
    int preconditioner;  /* 2: 3x3 block Jacobi; 1: profile LU of node blocks (OP PRE 1 with OP BLK
                            nblock, src/solver/solv_blk_*.c), FP64 only; 3: block ILU(0) in multicolour
                            order (no counterpart in the reference), FP64 or FP32 factors (precision
                            0 or 1) */
    int check_every;     /* Krylov iterations before the first host convergence check of a solve that has
                            no previous count (later solves start from the previous count, then check every
                            2; with graph replay both are rounded up to multiples of 4) */
    int nblock;          /* OP BLK: node blocks of the profile-LU preconditioner (default 1) */
    int graphs;          /* 1: replay captured HIP graphs of Krylov iteration chunks; 0 (default): launch eagerly */
    int krylov;          /* 0 (default): BiCGSTAB (solv_bcgstab.c); 1: restarted GMRES(30), left-preconditioned,
                            stopping on the Arnoldi 2-norm of the preconditioned residual (no counterpart in the
                            reference; single models, precision 0 or 1) */
    int lu_tiled;        /* preconditioner 1: 0 (default) factors node blocks of up to 2,730 nodes with one
                            workgroup each and larger ones (up to OP BLK 1, the whole matrix) with the tiled
                            band LU; 1 uses the tiled band LU for every block (an ensemble handle
                            always does) */
} adh_sw2_opts;

typedef struct {
    int converged;       /* fe_newton's YES/NO as 1/0 */
    int iterations;      /* Newton iterations */
    int lin_iterations;  /* total BiCGSTAB iterations */
    int solv_fail;
    int nhist;           /* entries written below (initial residual first) */
    int hist_cap;        /* capacity of the history arrays (may be 0) */
    double *hist_max, *hist_l2, *hist_inc;
    int *hist_lin_it;
    double ms_load, ms_setup, ms_solve, ms_resid, ms_total;  /* device time per phase */
} adh_sw2_report;

typedef struct adh_sw2_hip adh_sw2_hip;

const char *adh_sw2_hip_last_error(void);
void adh_sw2_hip_default_opts(adh_sw2_opts *o);

/* Creates device-resident mesh, adjacency, sparsity pattern and work arrays.
 * Pattern = the reference's SPARSE_VECT insertion order (spv_load,
 * src/solver/spv_load.c:5-139, driven by fe_sw2_load's element loops). */
int adh_sw2_hip_create(const adh_sw2_mesh *mesh, const adh_sw2_opts *opts, adh_sw2_hip **out);
void adh_sw2_hip_destroy(adh_sw2_hip *h);
int adh_sw2_hip_set_opts(adh_sw2_hip *h, const adh_sw2_opts *opts);
int adh_sw2_hip_set_dt(adh_sw2_hip *h, double dt);

/* Geometry as used (host copies; any pointer may be NULL). */
int adh_sw2_hip_geometry(adh_sw2_hip *h, double *djac, double *grad, double *djac1d, double *nrml1d);
/* Off-diagonal sparsity: nnz blocks; row_ptr[N+1], col[nnz] in insertion order. */
int adh_sw2_hip_pattern_nnz(adh_sw2_hip *h);
int adh_sw2_hip_pattern(adh_sw2_hip *h, int *row_ptr, int *col);

/* SSW_2D state (head, vel, old_*, older_*); vel arrays interleaved (x,y). */
int adh_sw2_hip_set_state(adh_sw2_hip *h, const double *head, const double *vel, const double *old_head,
                          const double *old_vel, const double *older_head, const double *older_vel);
int adh_sw2_hip_get_state(adh_sw2_hip *h, double *head, double *vel, double *old_head, double *old_vel,
                          double *older_head, double *older_vel);

/* fe_sw2 state shift older <- old <- current (src/fe/sw2/fe_sw2.c:42-50). */
int adh_sw2_hip_shift(adh_sw2_hip *h);
/* fe_sw2_init (src/fe/sw2/fe_sw2_init.c:24-92) incl. fe_sw2_wdflag; bc_mask[3N] optional. */
int adh_sw2_hip_init(adh_sw2_hip *h, int *bc_mask, int *wd_flag);
/* adh_sw2_hip_init without any device allocation or free: the bc_mask is
 * staged in a fixed per-GPU area of the library (24 MiB, models of up to
 * 2,097,152 nodes; larger ones return ADH_SW2_EINVAL).  No counterpart in the
 * reference: for callers that cannot allocate device memory, e.g. inside HIP
 * graph capture or a serving loop that runs fe_sw2_init once per Newton
 * attempt.  Calls that overlap on one GPU race on the staging area. */
int adh_sw2_hip_init_noalloc(adh_sw2_hip *h, int *bc_mask, int *wd_flag);
/* fe_sw2_resid (src/fe/sw2/fe_sw2_resid.c:29-216); residual[3N] optional. */
int adh_sw2_hip_resid(adh_sw2_hip *h, double *residual);
/* fe_sw2_load (src/fe/sw2/fe_sw2_load.c:34-269) incl. the diagonal floor;
 * diagonal[9N], offdiag[9*nnz] and residual[3N] (after the floor's residual
 * tweak) optional.  The refined forms (precision >= 1, block Jacobi) keep the
 * off-diagonal blocks in FP32: offdiag then holds their FP32 values. */
int adh_sw2_hip_load(adh_sw2_hip *h, double *diagonal, double *offdiag, double *residual);
/* Element Jacobians as assembled (after Dirichlet edits): [81*E] in block order
 * (row node i, col node j) -> 9 values; for parity tests.  The refined forms
 * (precision >= 1) keep them in FP32: the values are the FP32 roundings of
 * the FP64 form's. */
int adh_sw2_hip_elem_jacobians(adh_sw2_hip *h, double *em2d, double *em1d);
/* fe_sw2_inc (src/fe/sw2/fe_sw2_inc.c:24-57); sol NULL = the device solution. */
int adh_sw2_hip_inc(adh_sw2_hip *h, const double *sol);
/* solv_linear_sys_setup (src/solver/solv_linear_sys_setup.c:15-74): Jacobi
 * scaling in place + preconditioner set-up; scale_vect[3N] optional.  x[3N]
 * is the caller's solution array on entry, the shift the solve adds back to
 * its result (solv_return_old.c:23-25); NULL = zero, what fe_newton passes
 * (fe_newton.c:470).  Once per system: a second call before the next load or
 * set_system returns ADH_SW2_EINVAL (the system is already scaled). */
int adh_sw2_hip_linear_sys_setup(adh_sw2_hip *h, const double *x, double *scale_vect);
/* solv_linear_sys_solve (src/solver/solv_linear_sys_solve.c:16-69):
 * BiCGSTAB (src/solver/solv_bcgstab.c:39-301) on the scaled system;
 * *verdict = YES/NO, sol[3N] optional, *lin_it optional. */
int adh_sw2_hip_linear_sys_solve(adh_sw2_hip *h, double *sol, int *lin_it, int *verdict);
/* y[3N] = M^-1 v[3N] with the preconditioner of the last set-up, as
 * solv_prec (src/solver/solv_prec.c:11-66) applies it inside BiCGSTAB:
 * the profile LU (solv_blk_solve.c:5-62) or the block ILU(0); for tests. */
int adh_sw2_hip_precond_apply(adh_sw2_hip *h, const double *v, double *y);
/* The node colouring of the block ILU(0) (preconditioner 3; no counterpart
 * in the reference): color[N] in 0..*ncolor-1, no two nodes that share a
 * matrix block alike.  Available after a set-up with preconditioner 3. */
int adh_sw2_hip_ilu_colors(adh_sw2_hip *h, int *color, int *ncolor);
/* Loads an explicit (unscaled) system: diagonal[9N], offdiag[9*nnz] in the
 * pattern above, rhs[3N]; sol is zeroed.  For solver-only use and tests. */
int adh_sw2_hip_set_system(adh_sw2_hip *h, const double *diagonal, const double *offdiag, const double *rhs);
/* fe_newton with the SW2 callbacks (src/fe/fe_newton.c:214-912):
 * init, residual, [load, setup, solve, inc, residual, line search]*.
 * *verdict = YES/NO; on NO the state is reset by init (as :883) and the
 * caller cuts dt (tc_scale). rep optional. */
int adh_sw2_hip_newton(adh_sw2_hip *h, adh_sw2_report *rep, int *verdict);
/* One time step as the reference's time loop runs it (fe_main -> fe_sw2 ->
 * fe_sw2_solve, retried with dt*0.25 on failure, src/main/adh_func.c:535-549). */
int adh_sw2_hip_timestep(adh_sw2_hip *h, adh_sw2_report *rep, int *verdict, int max_retries);

/* Node wind values SWIND.stress, x and y per node ([2N]), as the reference's
 * meteor_update sets them before each time-step attempt from the wind
 * station series (src/meteor/meteor_update.c:36-43, called at
 * src/main/adh_func.c:545).  Only for a model created with wind = 1. */
int adh_sw2_hip_set_wind(adh_sw2_hip *h, const double *stress_xy);

/* Time-series values the residuals read, re-averaged by the caller over
 * [t, t + dt] before every time-step attempt, as the reference does with
 * sseries_setall_ts_valuesAVG (src/main/adh_func.c:542, src/structs/sseries.c:618)
 * before fe_sw2_boundary_resid reads them through sseries_get_value
 * (src/fe/sw2/fe_sw2_boundary_resid.c:212-214, 372) and fe_sw2_body_resid its
 * rain source (src/fe/sw2/fe_sw2_body_resid.c:728-754): e1d_val[B] as in
 * adh_sw2_mesh.e1d_val (NB VEL flux / NB OTW tail elevation), e2d_src[E] as
 * in adh_sw2_mesh.e2d_src.  Either pointer may be NULL (left as it is). */
int adh_sw2_hip_set_series(adh_sw2_hip *h, const double *e1d_val, const double *e2d_src);

/* ---- Ensembles: n independent models advanced together -------------------
 * No counterpart in the reference, whose fe_newton runs one model at a time
 * (src/fe/fe_newton.c:214); this is the additive batched entry beside the
 * single-model drop-in.  The models are laid side by side in one device mesh
 * (model i's nodes start on a multiple of 256), so every element, assembly,
 * scaling and BiCGSTAB launch covers all models, while reductions, Krylov
 * scalars, Newton decisions, line searches and dt cuts stay per model.  A
 * model in an ensemble computes exactly what it computes alone.
 * The models must share gravity, tau_temporal, tau_pg, the drying limits and
 * viscosity; meshes, materials, boundary conditions, states and dt may
 * differ.  Ensembles use BiCGSTAB with the 3x3 block-Jacobi preconditioner
 * (opts NULL or preconditioner 2) in any precision form, or with the profile
 * LU (preconditioner 1): each model's nodes cut into opts.nblock blocks as
 * that model alone would be, all models' bands factored by the tiled band LU
 * in the same launches (their storage must fit in device memory together:
 * set-up fails with ADH_SW2_EUNSUPPORTED otherwise).  The handle works with
 * set_opts and destroy; the single-model state, init and step calls are not
 * meant for it. */
int adh_sw2_hip_batch_create(const adh_sw2_mesh *meshes, int n, const adh_sw2_opts *opts, adh_sw2_hip **out);
/* number of models (ADH_SW2_EINVAL for a single-model handle) */
int adh_sw2_hip_batch_size(adh_sw2_hip *h);
/* state of model i, same layout as adh_sw2_hip_set_state over its own nodes */
int adh_sw2_hip_batch_set_state(adh_sw2_hip *h, int i, const double *head, const double *vel,
                                const double *old_head, const double *old_vel, const double *older_head,
                                const double *older_vel);
int adh_sw2_hip_batch_get_state(adh_sw2_hip *h, int i, double *head, double *vel);
int adh_sw2_hip_batch_set_dt(adh_sw2_hip *h, int i, double dt);
/* current dt of model i (after any cuts the last time step made) */
int adh_sw2_hip_batch_get_dt(adh_sw2_hip *h, int i, double *dt);
/* node wind values of model i (as adh_sw2_hip_set_wind over its own nodes) */
int adh_sw2_hip_batch_set_wind(adh_sw2_hip *h, int i, const double *stress_xy);
/* series values of model i (as adh_sw2_hip_set_series over its own edges and
 * triangles) */
int adh_sw2_hip_batch_set_series(adh_sw2_hip *h, int i, const double *e1d_val, const double *e2d_src);
/* nonlinear tolerance of model i (the reference's NTL is one absolute value
 * per model); models without one use opts.tol_nonlin */
int adh_sw2_hip_batch_set_tol(adh_sw2_hip *h, int i, double tol_nonlin);
/* max-norm of every model's opening Newton residual (init, residual, norms
 * as fe_newton starts); maxn holds n entries.  Applies init to every model,
 * as adh_sw2_hip_init does. */
int adh_sw2_hip_batch_residual_norm(adh_sw2_hip *h, double *maxn);
/* One time step of every selected model (which[i] != 0; which NULL = all),
 * each as adh_sw2_hip_timestep takes it: shift, fe_newton, and after a
 * failure dt * 0.25 and another attempt, up to max_retries, for the failed
 * models only (the others wait).  Unselected models are left untouched and
 * report YES.  reps (optional) and verdicts hold n entries.  A model whose
 * dt would fall below DBL_EPSILON stops with verdict NO (the single-model
 * call returns an error there). */
int adh_sw2_hip_batch_timestep(adh_sw2_hip *h, const int *which, adh_sw2_report *reps, int *verdicts,
                               int max_retries);

/* Reads the device linear system as it currently stands (after load it is
 * unscaled, after linear_sys_setup it is Jacobi-scaled); any pointer may be NULL. */
int adh_sw2_hip_get_system(adh_sw2_hip *h, double *diagonal, double *offdiag, double *rhs);

/* Average device duration (ms) of one launch of a phase kernel, timed with
 * HIP events on the library's stream over `reps` back-to-back launches on
 * the model's current data: 0 = SpMV + block-Jacobi apply + dot (the BiCGSTAB
 * K1 kernel, in the configured precision), 1 = finite-difference element
 * Jacobian, 2 = element residual, 3 = ordered block assembly, 4 = a plain
 * copy of the 81E-double element-matrix slab into a scratch buffer (read +
 * write of 1296 B per triangle; calibrates memory counters).  Clobbers the
 * Krylov work vectors; no counterpart in the reference (measurement hook). */
int adh_sw2_hip_time_kernel(adh_sw2_hip *h, int which, int reps, double *avg_ms);

/* Launch shapes of the four element list kernels (FD Jacobian over the wet
 * and the general triangle lists, residual over the same lists): block size
 * (64, 128 or 256 threads) and grid size of the grid-stride launch.  They
 * change the schedule only, never a result.  The *_ms fields are outputs of
 * adh_sw2_hip_autotune (average launch time of the chosen shape) and are
 * ignored by adh_sw2_hip_set_tuning.  No counterpart in the reference. */
typedef struct {
    int block[4];      /* 0 Jacobian wet, 1 Jacobian general, 2 residual wet, 3 residual general */
    int grid[4];
    int layout[4];     /* Jacobian lists: 0 = a lane per (triangle, column, sign), 1 = waves that
                          each take one column of 32 triangles (both signs), 2 = as 1 with the
                          triangles' inputs staged in LDS, 3 = waves that each take all nine
                          columns of 32 triangles from one LDS staging; 0 for the residual lists */
    double ms[4];
} adh_sw2_tuning;
int adh_sw2_hip_get_tuning(adh_sw2_hip *h, adh_sw2_tuning *out);
int adh_sw2_hip_set_tuning(adh_sw2_hip *h, const adh_sw2_tuning *t);
/* Times every candidate shape of each list kernel on the handle's current
 * state and triangle lists (HIP events on the library's stream, one warm-up
 * and `reps` timed launches per shape), keeps the fastest and returns it.
 * For an ensemble the lists cover every model.  Leaves the state unchanged. */
int adh_sw2_hip_autotune(adh_sw2_hip *h, int reps, adh_sw2_tuning *out);

/* Correctly rounded log used by the friction law, evaluated on the host. */
double adh_sw2_hip_cr_log(double x);
/* Evaluates the device's log (cr_log) and sqrt on n inputs -- a check that
 * the GPU's rounding matches the host's; no counterpart in the reference. */
int adh_sw2_hip_probe_math(const double *x, int n, double *log_out, double *sqrt_out, double *div_out);

#ifdef __cplusplus
}
#endif
#endif

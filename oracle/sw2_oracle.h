/*
 * CPU oracle for the AdH SW2 Newton step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's legacy serial SW2 path
 * (fe_sw2_init / fe_sw2_resid / fe_sw2_load / solv_linear_sys_setup /
 * solv_linear_sys_solve / fe_newton).  Every function cites the reference
 * file:line it follows.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the timed CPU baseline; the product path never links it.
 *
 * Arithmetic is kept in the reference's expression order and compiled with
 * -O2 -ffp-contract=off (no FMA contraction), so that it reproduces what the
 * reference computes on the same doubles.
 */
#ifndef SW2_ORACLE_H
#define SW2_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary-condition flags (reference src/include/define.h:375-413) */
#define O_BCT_OUTFLOW 4
#define O_BCT_PRS_DIR 7
#define O_BCT_VEL_DIR 8
#define O_BCT_PRS_NEU 9
#define O_BCT_VEL_NEU 10
#define O_BCT_VEL_PRS_DIR 16

/* turbulence closure per material (reference smat.c:98-110) */
#define O_TURB_NONE 0
#define O_TURB_EVS 1
#define O_TURB_EEV 2

typedef struct {
    int nnodes, nelems2d, nelems1d, nmat;
    /* nodes */
    const double *x, *y, *z;        /* [N] */
    const int *node_bc;             /* [N] bc flag of the node's node-string, 0 = none */
    const double *node_bc_val;      /* [3N] Dirichlet u,v,h values (may be NULL) */
    /* 2D triangles */
    const int *e2d;                 /* [3E] node ids (counter-clockwise) */
    const int *e2d_mat;             /* [E] material id */
    const double *e2d_ks;           /* [E] equivalent roughness height of the element string */
    const int *e2d_src_flag;        /* [E] 1 = BCT_WATER_SOURCE string */
    const double *e2d_src;          /* [E] water source value (m/s) */
    /* 1D boundary edges */
    const int *e1d;                 /* [2B] */
    const int *e1d_adj;             /* [B] adjacent triangle (outward normal test) */
    const int *e1d_bc;              /* [B] O_BCT_VEL_NEU / O_BCT_PRS_NEU / O_BCT_OUTFLOW */
    const double *e1d_val;          /* [B] series value: flux series for VEL_NEU, tail elevation for PRS_NEU */
    /* materials */
    const int *mat_turb;            /* [M] */
    const int *mat_eev_mode;        /* [M] */
    const double *mat_eev_coef;     /* [M] */
    const double *mat_ev;           /* [3M] xx, yy, xy */
    const double *mat_cor_speed;    /* [M] coriolis angular speed 2*Omega*sin(lat) */
    /* scalars */
    double g, dt, tau_temporal, tau_pg, drying_lower, drying_upper, viscosity;
    /* winds (MP WND, SERIES WIND): wind = 0 switches them off */
    int wind;
    const double *mat_windatt;      /* [M] attenuation (MP WND ATT, default 1.0) */
    double density;                 /* MP RHO */
} o_desc;

typedef struct {
    double *head, *vel, *old_head, *old_vel, *older_head, *older_vel; /* vel arrays are AoS x,y */
} o_state;

typedef struct {
    double tol_nonlin, inc_nonlin;
    int max_nonlin_it, max_lin_it, max_cuts;
    int prec;       /* 1 = skyline LU over node blocks (reference), 2 = 3x3 block Jacobi */
    int nblock;     /* number of node blocks (reference OP BLK) */
    double lin_rtol, lin_atol; /* BiCGSTAB stop: ||r||inf <= atol + rtol*||M^-1 b||inf (1e-10, 1e-5) */
    int hist_cap;
} o_newton_opts;

typedef struct {
    int converged, iterations, lin_iterations_total, solv_fail;
    int nhist;
    double *hist_max, *hist_l2, *hist_inc;   /* [hist_cap] after each Newton iteration; [0] = initial */
    int *hist_lin_it;
    double time_load, time_setup, time_solve, time_resid;
} o_newton_report;

/* opaque model */
typedef struct o_model o_model;

o_model *o_create(const o_desc *d);
void o_destroy(o_model *m);
/* geometry the oracle derived: djac[E], grad[6E], djac1d[B], nrml1d[2B] */
void o_get_geometry(const o_model *m, double *djac, double *grad, double *djac1d, double *nrml1d);
/* sparsity pattern in the reference's insertion order: row_ptr[N+1], col[nnz] (off-diagonal) */
int o_pattern_nnz(const o_model *m);
void o_get_pattern(const o_model *m, int *row_ptr, int *col);

void o_set_state(o_model *m, const double *head, const double *vel, const double *old_head,
                 const double *old_vel, const double *older_head, const double *older_vel);
void o_get_state(const o_model *m, double *head, double *vel, double *old_head, double *old_vel,
                 double *older_head, double *older_vel);
void o_set_dt(o_model *m, double dt);
/* node wind values SWIND.stress (x, y per node) as meteor_update sets them
   (src/meteor/meteor_update.c:36-43) */
void o_set_wind(o_model *m, const double *stress_xy);
/* series values averaged over [t, t+dt] (sseries_setall_ts_valuesAVG,
   src/main/adh_func.c:542): edge values [B], rain sources [E]; NULL keeps */
void o_set_series(o_model *m, const double *e1d_val, const double *e2d_src);

void o_wdflag(o_model *m, int *wd_flag_out);
void o_init(o_model *m, int *bc_mask_out);
void o_body_resid(o_model *m, int ie, double perturbation, int perturb_node, int perturb_var,
                  int perturb_sign, double *elem_rhs9);
void o_boundary_resid(o_model *m, int ie, double perturbation, int perturb_node, int perturb_var,
                      int perturb_sign, double *elem_rhs6);
void o_resid(o_model *m, double *residual_out);
void o_load(o_model *m, double *diag_out, double *offd_out);
void o_elem_jacobian(o_model *m, int ie, double *mat81);
void o_linear_setup(o_model *m, const o_newton_opts *o, double *scale_out);
int  o_linear_solve(o_model *m, const o_newton_opts *o, double *sol_out, int *lin_it);
void o_inc(o_model *m, const double *sol);
void o_shift(o_model *m);
int  o_newton(o_model *m, const o_newton_opts *o, o_newton_report *r);
void o_get_residual(const o_model *m, double *r);
void o_get_matrix(const o_model *m, double *diag, double *offd);
void o_set_matrix(o_model *m, const double *diag, const double *offd, const double *rhs);
void o_get_rhs(const o_model *m, double *b);
/* generic BiCGSTAB on the current (already scaled) system with block-Jacobi or skyline LU;
   used to pin the Krylov recurrences */
int  o_bcgstab_only(o_model *m, const o_newton_opts *o, double *x_out, int *lin_it);
double o_log(double x);
void o_set_log_mode(int mode);
double o_drag_coef(double depth, double ks);

#ifdef __cplusplus
}
#endif
#endif

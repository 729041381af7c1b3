/*
 * CPU oracle for the AdH SW2 Newton step -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, in the reference's own arithmetic order, the legacy serial SW2
 * path.  Each function cites the reference file:line it follows (paths are
 * relative to the reference's src/).  This file is the checker: tests and the
 * bench's CPU-baseline leg load it; the HIP product never does.
 *
 * Pinning (what checks each function against something other than itself) is
 * listed in tests/test_oracle_*.py; functions without such a check are
 * flagged "parity unpinned" at their definition.
 *
 * Scope of the restatement: triangles only (NDONTRI), SW2 physics with
 * EVS/EEV eddy viscosity, bed friction from the equivalent roughness height,
 * Coriolis, rainfall source strings, wind surface stress (stress law 0, one
 * station), wet/dry clipping, SUPG, node-string Dirichlet flags, and the 1D
 * boundary types NB VEL, NB OTW and NB OUT.  Baroclinic density, lids, wave
 * stress, vorticity, hydraulic structures and quadrilaterals are not
 * represented (their terms are exactly zero for the decks this oracle is used
 * with; see the notes inline).
 *
 * The wind term, the drag law's clamped branch and the wet/dry path with
 * eddy viscosity are pinned by the reference's wind decks: angle_wind's
 * transect error (at the 6000 s run length its golden figure fixes) and
 * wet_dry's three error figures reproduce to 1e-8 and 1e-10
 * (tests/test_qa_golden.py).
 */
#include "sw2_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <quadmath.h>

/* constants, reference src/include/define.h:27-49 (literal spellings kept) */
#define SMALL DBL_EPSILON
#define NOT_QUITE_SMALL FLT_EPSILON
#define SOLV_TOL DBL_EPSILON
#define one_2 0.5e+0
#define one_3 0.333333333333333333333333
#define one_6 0.166666666666666666666667
#define one_12 0.083333333333333333333333
#define one_30 0.033333333333333333333333
#define one_60 0.016666666666666666666667
#define PERTURB_NONE -1
#define PERTURB_U 1
#define PERTURB_V 2
#define PERTURB_H 5
/* reference macro.h:844-846: unparenthesised on purpose */
#define MIN(a, b) (a < b ? a : b)
#define MAX(a, b) (a > b ? a : b)

typedef struct { double x_eq, y_eq, c_eq; } dof3;
typedef struct { double x, y; } v2;
typedef struct { double x, y, z; } v3;

/* SPARSE_VECT (reference structs/ssolve.h:16-21): row list in insertion order */
typedef struct { int size, cap; int *index; double *value; } spv;
/* BAND_VECT (reference structs/ssolve.h:23-29) */
typedef struct { int begin, end, size; double *value; } bandv;

struct o_model {
    o_desc d;
    int N, E, B;
    /* geometry */
    double *djac, *grad;      /* [E], [6E] */
    double *djac1d, *nrml1d;  /* [B], [2B] */
    /* state (reference SSW_2D, vel AoS) */
    double *head, *vel, *old_head, *old_vel, *older_head, *older_vel;
    int *wd_flag;
    int *bc_mask;
    double dt;
    /* linear system (reference SSUPER_MODEL) */
    double *residual, *sol, *diagonal, *scale_vect;
    spv *matrix;
    int pattern_built;
    /* skyline profile */
    bandv *rows, *cols;
    int prof_size;
    int *node_block;
    int prof_nblock;
    /* block-Jacobi inverses */
    double *bjinv;
    /* node wind values (x, y per node) */
    double *wind;
    /* series values the residuals read (owned copies; d.e1d_val / d.e2d_src point here) */
    double *val1d, *src2d;
};

/* ---------------------------------------------------------------------- */
/* geometry                                                                */
/* ---------------------------------------------------------------------- */

/* reference elem/element_djac_nrml_gradPhi.c:193-274 (2D part only) */
static void tri_geometry(double x1, double x2, double x3, double y1, double y2, double y3,
                         double *djac, double *g) {
    double s1x = x2 - x1, s1y = y2 - y1;
    double s2x = x3 - x1, s2y = y3 - y1;
    double dj = (s1x * s2y) - (s2x * s1y);
    if (fabs(dj) > SMALL) {
        double inv = 1. / dj; /* svect2d_scale_replace_array(grad, 1./djac) */
        g[0] = (s1y - s2y) * inv;
        g[1] = (s2x - s1x) * inv;
        g[2] = s2y * inv;
        g[3] = -s2x * inv;
        g[4] = -s1y * inv;
        g[5] = s1x * inv;
    } else {
        int k;
        for (k = 0; k < 6; k++) g[k] = -9999999.9;
    }
    *djac = dj * (1.0 / 2.0); /* SCALE_FOR_MASTER_TRI_VOLUME element_djac_nrml_gradPhi.c:14 */
}

/* reference elem/element_djac_nrml_gradPhi.c:24-36 and structs/selem_1d.c:118-160 */
static void edge_geometry(const o_desc *d, int ie, double *djac, double *nrml) {
    int n1 = d->e1d[2 * ie], n2 = d->e1d[2 * ie + 1];
    double x1 = d->x[n1], x2 = d->x[n2], y1 = d->y[n1], y2 = d->y[n2];
    *djac = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
    double delx = x2 - x1, dely = y2 - y1;
    double distance = sqrt(delx * delx + dely * dely);
    nrml[0] = dely / distance;
    nrml[1] = -delx / distance;
    const int *t = d->e2d + 3 * d->e1d_adj[ie];
    double x = one_3 * (d->x[t[0]] + d->x[t[1]] + d->x[t[2]]);
    double y = one_3 * (d->y[t[0]] + d->y[t[1]] + d->y[t[2]]);
    delx = x - x1;
    dely = y - y1;
    distance = delx * nrml[0] + dely * nrml[1];
    if (distance > 0.) {
        nrml[0] *= -1.;
        nrml[1] *= -1.;
    }
}

/* ---------------------------------------------------------------------- */
/* friction, reference friction/fr_bedshstr_drag_coef.c:18-49             */
/* ---------------------------------------------------------------------- */
/* log mode 0: the platform libm log, as the reference calls it.
   log mode 1: correctly rounded log, computed in binary128 (libquadmath, error
   below 2^-112) and rounded once to double; used when comparing with the HIP
   path, whose device log is correctly rounded. */
static int g_log_mode = 0;
void o_set_log_mode(int mode) { g_log_mode = mode; }
double o_log(double x) {
    if (g_log_mode == 1) return (double)logq((__float128)x);
    return log(x);
}

double o_drag_coef(double depth, double roughness_height) {
    double depth_rh, bta, phi, q;
    depth = MAX(depth, 1.E-8);
    roughness_height = MAX(roughness_height, 1.E-8);
    if (roughness_height > 29.7 * depth)
        depth_rh = 1. / 29.7;
    else
        depth_rh = depth / roughness_height;
    bta = 29.7 * depth_rh;
    /* pow(q, 2.) of a correctly rounded pow is the rounded square q*q */
    q = (.4 * bta) / ((bta + 1) * (o_log(bta + 1) - 1) + 1);
    phi = q * q;
    return 2. * phi;
}

/* reference fe/sw2/fe_sw2_get_roughness.c:28-52 (bed shear branch only) */
static double get_roughness(const o_desc *d, int ie, double depth) {
    double ks = d->e2d_ks[ie];
    if (ks > SMALL) return o_drag_coef(depth, ks);
    return 0.0;
}

/* ---------------------------------------------------------------------- */
/* element integrals, reference fe/fe_elem2d_integrations.c / 1d          */
/* ---------------------------------------------------------------------- */

/* :60-62 */
static double int_tri_f(double djac, double c, const double *f) {
    return (c * djac * (f[0] + f[1] + f[2]) / 3.0);
}
/* :119-121 */
static double int_tri_f_f(double djac, double c, const double *f) {
    return (c * one_6 * djac * (f[0] * f[0] + f[0] * f[1] + f[1] * f[1] + f[0] * f[2] + f[1] * f[2] + f[2] * f[2]));
}
/* :159-164 */
static void int_tri_phi(double djac, double c, double *integral) {
    double t1 = c * djac * one_3;
    integral[0] += t1;
    integral[1] += t1;
    integral[2] += t1;
}
/* :196-202 */
static void int_tri_phi_f(double djac, double c, const double *f, double *integral) {
    double f_sum = f[0] + f[1] + f[2];
    double t1 = djac * c * one_12;
    integral[0] += t1 * (f_sum + f[0]);
    integral[1] += t1 * (f_sum + f[1]);
    integral[2] += t1 * (f_sum + f[2]);
}
/* :250-255 */
static void int_tri_phi_f_lump(double djac, double c, const double *f, double *integral) {
    double t1 = djac * c * one_3;
    integral[0] += t1 * f[0];
    integral[1] += t1 * f[1];
    integral[2] += t1 * f[2];
}
/* :281-287 */
static void int_tri_phi_f_g(double djac, double c, const double *f, const double *g, double *integral) {
    double g_sum = 2 * (g[0] + g[1] + g[2]);
    double t1 = c * djac * one_60;
    integral[0] += t1 * (f[0] * (g_sum + 4 * g[0]) + f[1] * (g_sum - g[2]) + f[2] * (g_sum - g[1]));
    integral[1] += t1 * (f[1] * (g_sum + 4 * g[1]) + f[2] * (g_sum - g[0]) + f[0] * (g_sum - g[2]));
    integral[2] += t1 * (f[2] * (g_sum + 4 * g[2]) + f[1] * (g_sum - g[0]) + f[0] * (g_sum - g[1]));
}
/* :352-367 */
static void int_tri_dphi_f_f(const v2 *gs, double djac, double c, const double *f, double *ix, double *iy) {
    double int_f_f = int_tri_f_f(djac, c, f);
    int i;
    if (ix) for (i = 0; i < 3; i++) ix[i] += gs[i].x * int_f_f;
    if (iy) for (i = 0; i < 3; i++) iy[i] += gs[i].y * int_f_f;
}
/* :422-427 */
static void int_tri_gradPhi_dot_vbar(const v2 *gs, double djac, double c, v2 vbar, double *integral) {
    double t1 = c * djac;
    integral[0] += t1 * (vbar.x * gs[0].x + vbar.y * gs[0].y);
    integral[1] += t1 * (vbar.x * gs[1].x + vbar.y * gs[1].y);
    integral[2] += t1 * (vbar.x * gs[2].x + vbar.y * gs[2].y);
}
/* :546-561 */
static void int_tri_gradPhi_dot_f_v(const v2 *gs, double djac, double c, const double *f, const v2 *v, double *integral) {
    double p1, p2, p3, tx, ty, con;
    p1 = (2 * f[0] + f[1] + f[2]);
    p2 = (f[0] + 2 * f[1] + f[2]);
    p3 = (f[0] + f[1] + 2 * f[2]);
    tx = (p1 * v[0].x + p2 * v[1].x + p3 * v[2].x);
    ty = (p1 * v[0].y + p2 * v[1].y + p3 * v[2].y);
    con = djac * one_12 * c;
    integral[0] += con * (gs[0].x * tx + gs[0].y * ty);
    integral[1] += con * (gs[1].x * tx + gs[1].y * ty);
    integral[2] += con * (gs[2].x * tx + gs[2].y * ty);
}
/* :832-850 */
static void int_tri_dphi_f_g_h(const v2 *gs, double djac, double c, const double *f, const double *g,
                               const double *h, double *ix, double *iy) {
    double con, ff1, ff2, ff3, ff4, ff5, ff6, gg1, gg2, gg3;
    ff1 = (3 * f[0] + f[1] + f[2]);
    ff2 = (2 * f[0] + 2 * f[1] + f[2]);
    ff3 = (2 * f[0] + f[1] + 2 * f[2]);
    ff4 = (f[0] + 3 * f[1] + f[2]);
    ff5 = (f[0] + 2 * f[1] + 2 * f[2]);
    ff6 = (f[0] + f[1] + 3 * f[2]);
    gg1 = (2 * ff1 * g[0] + ff2 * g[1] + ff3 * g[2]);
    gg2 = (ff2 * g[0] + 2 * ff4 * g[1] + ff5 * g[2]);
    gg3 = (ff3 * g[0] + ff5 * g[1] + 2 * ff6 * g[2]);
    con = djac * one_60 * c * (gg1 * h[0] + gg2 * h[1] + gg3 * h[2]);
    ix[0] += con * gs[0].x;
    ix[1] += con * gs[1].x;
    ix[2] += con * gs[2].x;
    iy[0] += con * gs[0].y;
    iy[1] += con * gs[1].y;
    iy[2] += con * gs[2].y;
}
/* :864-884 */
static void int_tri_dphi_f_f_h(const v2 *gs, double djac, double c, const double *f, const double *h,
                               double *ix, double *iy) {
    double p1, p2, p3, con;
    double f1_2 = f[0] * f[0], f2_2 = f[1] * f[1], f3_2 = f[2] * f[2];
    p1 = (3 * f1_2 + 2 * f[0] * f[1] + f2_2 + (2 * f[0] + f[1]) * f[2] + f3_2);
    p2 = (f1_2 + 2 * f[0] * f[1] + 3 * f2_2 + (f[0] + 2 * f[1]) * f[2] + f3_2);
    p3 = (f1_2 + f[0] * f[1] + f2_2 + 2 * (f[0] + f[1]) * f[2] + 3 * f3_2);
    con = djac * one_30 * c * (p1 * h[0] + p2 * h[1] + p3 * h[2]);
    int i;
    if (ix) for (i = 0; i < 3; i++) ix[i] += con * gs[i].x;
    if (iy) for (i = 0; i < 3; i++) iy[i] += con * gs[i].y;
}
/* :1016-1031 */
static void int_tri_phi_h_df(double djac, double c, const double *h, v2 df, double *ix, double *iy) {
    double con, sum_h = h[0] + h[1] + h[2];
    con = djac * one_12 * c * df.x;
    ix[0] += con * (sum_h + h[0]);
    ix[1] += con * (sum_h + h[1]);
    ix[2] += con * (sum_h + h[2]);
    con = djac * one_12 * c * df.y;
    iy[0] += con * (sum_h + h[0]);
    iy[1] += con * (sum_h + h[1]);
    iy[2] += con * (sum_h + h[2]);
}
/* fe_elem1d_integrations.c:31-35 */
static void int_line_phi_h_v_dot_n(double djac, double c, const double *h, const v2 *v, v2 nrml, double *integral) {
    double t1 = djac * one_12 * c;
    integral[0] += t1 * (nrml.x * ((3 * h[0] + h[1]) * v[0].x + (h[0] + h[1]) * v[1].x) + nrml.y * ((3 * h[0] + h[1]) * v[0].y + (h[0] + h[1]) * v[1].y));
    integral[1] += t1 * (nrml.x * ((h[0] + h[1]) * v[0].x + (h[0] + 3 * h[1]) * v[1].x) + nrml.y * ((h[0] + h[1]) * v[0].y + (h[0] + 3 * h[1]) * v[1].y));
}
/* :45-49 */
static void int_line_phi_h_h(double djac, double c, double h1, double h2, double *integral) {
    double t1 = djac * one_12 * c;
    integral[0] += t1 * (3 * h1 * h1 + 2 * h1 * h2 + h2 * h2);
    integral[1] += t1 * (h1 * h1 + 2 * h1 * h2 + 3 * h2 * h2);
}
/* :73-77 */
static void int_line_phi(double djac, double c, double *integral) {
    double t1 = one_2 * djac * c;
    integral[0] += t1;
    integral[1] += t1;
}
/* :87-91 */
static void int_line_phi_f(double djac, double c, const double *f, double *integral) {
    double t1 = one_6 * djac * c;
    integral[0] += t1 * (2 * f[0] + f[1]);
    integral[1] += t1 * (f[0] + 2 * f[1]);
}
/* :115-119 */
static void int_line_phi_f_g_h(double djac, double c, double f1, double f2, double g1, double g2, double h1,
                               double h2, double *integral) {
    double t1 = one_60 * djac * c;
    integral[0] += t1 * ((3 * (4 * f1 + f2) * g1 + (3 * f1 + 2 * f2) * g2) * h1 + ((3 * f1 + 2 * f2) * g1 + (2 * f1 + 3 * f2) * g2) * h2);
    integral[1] += t1 * (((3 * f1 + 2 * f2) * g1 + (2 * f1 + 3 * f2) * g2) * h1 + ((2 * f1 + 3 * f2) * g1 + 3 * (f1 + 4 * f2) * g2) * h2);
}

/* ---------------------------------------------------------------------- */
/* wet/dry wrapper, reference fe/fe_sw2_wet_dry_wrapper.c                  */
/* ---------------------------------------------------------------------- */

enum wd_func { WD_INTEGRATE_F, WD_CONT_TEMPORAL, WD_CONVECTION, WD_GLS_CONVECTION, WD_PRESSURE,
               WD_BODYFORCE, WD_BOUNDARY_PRESSURE };

/* integrands, reference fe/sw2/fe_sw2_wet_dry_integrations.c */
static void wd_apply(enum wd_func fn, const v3 *x, const double *h, const v2 *gp, const v2 *v, const v2 *v_wd,
                     const double *f_wd, double djac, const double *vars, dof3 *rhs) {
    double dt = vars[0], c = vars[1];
    int i;
    switch (fn) {
    case WD_INTEGRATE_F: { /* :95-116 */
        double t1 = (h[0] + h[1] + h[2]) / 3.0, t2 = 0.0, t3 = 0.0;
        if (v_wd) {
            t2 = (v_wd[0].x + v_wd[1].x + v_wd[2].x) / 3.0;
            t3 = (v_wd[0].y + v_wd[1].y + v_wd[2].y) / 3.0;
        }
        rhs[0].c_eq += djac * t1;
        rhs[0].x_eq += djac * t2;
        rhs[0].y_eq += djac * t3;
        break;
    }
    case WD_CONT_TEMPORAL: { /* :143-160 */
        double rc[3] = {0., 0., 0.};
        int_tri_phi_f(djac, c, h, rc);
        for (i = 0; i < 3; i++) rhs[i].c_eq += rc[i];
        break;
    }
    case WD_CONVECTION: { /* :200-229 */
        double rc[3] = {0., 0., 0.}, rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
        double uu[3], vv[3];
        c *= dt;
        for (i = 0; i < 3; i++) { uu[i] = v_wd[i].x; vv[i] = v_wd[i].y; }
        int_tri_gradPhi_dot_f_v(gp, djac, c, h, v_wd, rc);
        int_tri_dphi_f_g_h(gp, djac, c, uu, vv, h, ry, rx);
        int_tri_dphi_f_f_h(gp, djac, c, uu, h, rx, NULL);
        int_tri_dphi_f_f_h(gp, djac, c, vv, h, NULL, ry);
        for (i = 0; i < 3; i++) {
            rhs[i].c_eq += rc[i];
            rhs[i].x_eq += rx[i];
            rhs[i].y_eq += ry[i];
        }
        break;
    }
    case WD_GLS_CONVECTION: { /* :361-381 */
        double h_avg = one_3 * (f_wd[0] + f_wd[1] + f_wd[2]);
        double u_avg = one_3 * (v_wd[0].x + v_wd[1].x + v_wd[2].x);
        double v_avg = one_3 * (v_wd[0].y + v_wd[1].y + v_wd[2].y);
        double dhdx = gp[0].x * f_wd[0] + gp[1].x * f_wd[1] + gp[2].x * f_wd[2];
        double dhdy = gp[0].y * f_wd[0] + gp[1].y * f_wd[1] + gp[2].y * f_wd[2];
        double dudx = gp[0].x * v_wd[0].x + gp[1].x * v_wd[1].x + gp[2].x * v_wd[2].x;
        double dudy = gp[0].y * v_wd[0].x + gp[1].y * v_wd[1].x + gp[2].y * v_wd[2].x;
        double dvdx = gp[0].x * v_wd[0].y + gp[1].x * v_wd[1].y + gp[2].x * v_wd[2].y;
        double dvdy = gp[0].y * v_wd[0].y + gp[1].y * v_wd[1].y + gp[2].y * v_wd[2].y;
        rhs[0].c_eq += djac * dt * c * (u_avg * dhdx + h_avg * dudx + v_avg * dhdy + h_avg * dvdy);
        rhs[0].x_eq += djac * dt * c * (u_avg * dudx + v_avg * dudy);
        rhs[0].y_eq += djac * dt * c * (u_avg * dvdx + v_avg * dvdy);
        break;
    }
    case WD_PRESSURE: { /* :412-425 */
        double rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
        int_tri_dphi_f_f(gp, djac, c, h, rx, ry);
        for (i = 0; i < 3; i++) { rhs[i].x_eq += rx[i]; rhs[i].y_eq += ry[i]; }
        break;
    }
    case WD_BODYFORCE: { /* :470-482 */
        double rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
        int_tri_phi_h_df(djac, c, h, *v, rx, ry);
        for (i = 0; i < 3; i++) { rhs[i].x_eq += rx[i]; rhs[i].y_eq += ry[i]; }
        break;
    }
    case WD_BOUNDARY_PRESSURE: { /* :511-537 */
        int nd1 = 0, nd2 = 0, iedge;
        double delx, dely, r[2];
        for (iedge = 0; iedge < 3; iedge++) {
            if (iedge == 0) { nd1 = 0; nd2 = 1; }
            else if (iedge == 1) { nd1 = 1; nd2 = 2; }
            else { nd1 = 2; nd2 = 0; }
            delx = x[nd1].x - x[nd2].x;
            dely = x[nd2].y - x[nd1].y;
            r[0] = 0.; r[1] = 0.;
            int_line_phi_h_h(1., c, h[nd1], h[nd2], r);
            rhs[nd1].x_eq += dely * r[0];
            rhs[nd2].x_eq += dely * r[1];
            rhs[nd1].y_eq += delx * r[0];
            rhs[nd2].y_eq += delx * r[1];
        }
        break;
    }
    }
}

/* reference fe/fe_sw2_wet_dry_wrapper.c:106-306 */
static double wd_wrapper(dof3 *elem_rhs, const v3 *x, const double *h, const v2 *grad_phi, const v2 *v,
                         const v2 *v_wet_dry, const double *f_wet_dry, double djac, int redistribute,
                         const double *vars, enum wd_func fn) {
    int i, is_dry, total_vtx = 0, passcode = 0, only_vtx;
    for (i = 0; i < 3; i++) {
        is_dry = (h[i] <= 0.0) ? 1 : 0;
        total_vtx += is_dry;
        passcode += (is_dry << i);
    }
    if (total_vtx == 3) return 0;
    else if (total_vtx == 0) {
        wd_apply(fn, x, h, grad_phi, v, v_wet_dry, f_wet_dry, djac, vars, elem_rhs);
        return 1.0;
    }
    if (total_vtx == 1)
        wd_apply(fn, x, h, grad_phi, v, v_wet_dry, f_wet_dry, djac, vars, elem_rhs);
    else
        passcode = (7 - passcode);
    only_vtx = passcode / 2;

    double xsi, new_h[3] = {0.0, 0.0, 0.0};
    v3 new_x[3];
    double new_f[3] = {0.0, 0.0, 0.0};
    v2 new_v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (i = 0; i < 3; i++) {
        if (i != only_vtx) {
            xsi = h[i] / (h[i] - h[only_vtx]);
            new_x[i].x = x[i].x + xsi * (x[only_vtx].x - x[i].x);   /* FIXED_POS macro.h:756 */
            new_x[i].y = x[i].y + xsi * (x[only_vtx].y - x[i].y);
            new_x[i].z = x[i].z;
            if (v_wet_dry) {
                new_v[i].x = v_wet_dry[i].x + xsi * (v_wet_dry[only_vtx].x - v_wet_dry[i].x);
                new_v[i].y = v_wet_dry[i].y + xsi * (v_wet_dry[only_vtx].y - v_wet_dry[i].y);
            }
            if (f_wet_dry) new_f[i] = f_wet_dry[i] + xsi * (f_wet_dry[only_vtx] - f_wet_dry[i]);
        } else {
            new_x[i] = x[i];
            new_h[i] = h[i];
            if (f_wet_dry) new_f[i] = f_wet_dry[i];
            if (v_wet_dry) { new_v[i].x = v_wet_dry[i].x; new_v[i].y = v_wet_dry[i].y; }
        }
    }
    double elem_djac = djac;
    { /* TRI_AREA macro.h:773 */
        double v0x = new_x[1].x - new_x[0].x, v0y = new_x[1].y - new_x[0].y;
        double v1x = new_x[2].x - new_x[0].x, v1y = new_x[2].y - new_x[0].y;
        djac = (v0x * v1y - v0y * v1x) / 2.0;
    }
    if (djac / elem_djac < SMALL) return (double)(total_vtx % 2);
    v2 new_grad[3];
    if (grad_phi) { /* GRAD_SHAPE macro.h:779 */
        new_grad[0].x = 0.5 * (new_x[1].y - new_x[2].y) / djac;
        new_grad[1].x = 0.5 * (new_x[2].y - new_x[0].y) / djac;
        new_grad[2].x = 0.5 * (new_x[0].y - new_x[1].y) / djac;
        new_grad[0].y = 0.5 * (new_x[2].x - new_x[1].x) / djac;
        new_grad[1].y = 0.5 * (new_x[0].x - new_x[2].x) / djac;
        new_grad[2].y = 0.5 * (new_x[1].x - new_x[0].x) / djac;
    }
    dof3 new_rhs[3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    wd_apply(fn, new_x, new_h, grad_phi ? new_grad : NULL, v, v_wet_dry ? new_v : NULL,
             f_wet_dry ? new_f : NULL, djac, vars, new_rhs);
    double factor = (total_vtx >> 1) - (total_vtx % 2);
    if (redistribute) { /* RE_DISTRIBUTE macro.h:789-805 */
        int i99;
        for (i99 = 0; i99 < 3; i99++) {
            if (i99 == only_vtx) continue;
            double a0x = new_x[i99].x - x[only_vtx].x, a0y = new_x[i99].y - x[only_vtx].y;
            double a1x = x[i99].x - x[only_vtx].x, a1y = x[i99].y - x[only_vtx].y;
            double l0 = a0x * a0x + a0y * a0y, l1 = a1x * a1x + a1y * a1y;
            double xsi99 = sqrt(l0 / l1);
            new_rhs[only_vtx].c_eq += (1.0 - xsi99) * new_rhs[i99].c_eq;
            new_rhs[only_vtx].x_eq += (1.0 - xsi99) * new_rhs[i99].x_eq;
            new_rhs[only_vtx].y_eq += (1.0 - xsi99) * new_rhs[i99].y_eq;
            new_rhs[i99].c_eq *= xsi99;
            new_rhs[i99].x_eq *= xsi99;
            new_rhs[i99].y_eq *= xsi99;
        }
    }
    for (i = 0; i < 3; i++) {
        elem_rhs[i].c_eq += factor * new_rhs[i].c_eq;
        elem_rhs[i].x_eq += factor * new_rhs[i].x_eq;
        elem_rhs[i].y_eq += factor * new_rhs[i].y_eq;
    }
    factor *= djac / elem_djac;
    return factor + ((factor < 0) ? 1.0 : 0.0);
}

/* reference fe/fe_sw2_wet_dry_wrapper.c:31-73 */
static double wd_factor(const v3 *x, const double *h, double djac) {
    int i, is_dry, total_vtx = 0, passcode = 0, only_vtx;
    double elem_djac, factor;
    for (i = 0; i < 3; i++) {
        is_dry = (h[i] <= 0.0) ? 1 : 0;
        total_vtx += is_dry;
        passcode += (is_dry << i);
    }
    if (total_vtx == 3) return 0.0;
    else if (total_vtx == 0) return 1.0;
    if (total_vtx != 1) passcode = (7 - passcode);
    only_vtx = passcode / 2;
    double xsi;
    v2 new_x[3];
    for (i = 0; i < 3; i++) {
        if (i != only_vtx) {
            xsi = h[i] / (h[i] - h[only_vtx]);
            new_x[i].x = x[i].x + xsi * (x[only_vtx].x - x[i].x);
            new_x[i].y = x[i].y + xsi * (x[only_vtx].y - x[i].y);
        } else {
            new_x[i].x = x[i].x;
            new_x[i].y = x[i].y;
        }
    }
    elem_djac = djac;
    {
        double v0x = new_x[1].x - new_x[0].x, v0y = new_x[1].y - new_x[0].y;
        double v1x = new_x[2].x - new_x[0].x, v1y = new_x[2].y - new_x[0].y;
        djac = (v0x * v1y - v0y * v1x) / 2.0;
    }
    if (djac / elem_djac < SMALL) return (double)(total_vtx % 2);
    factor = (total_vtx >> 1) - (total_vtx % 2);
    factor *= djac / elem_djac;
    return factor + ((factor < 0) ? 1.0 : 0.0);
}

/* reference fe/sw2/fe_sw2_wet_dry_integrations.c:42-61 */
static void wd_average(const v3 *x, const double *depth, const v2 *v_wd, const double *f_wd, double djac,
                       double *f_wd_avg, v2 *v_wd_avg) {
    dof3 averages[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    double vars[2] = {1., 1.};
    double factor = wd_wrapper(averages, x, depth, NULL, NULL, v_wd, f_wd, djac, 0, vars, WD_INTEGRATE_F);
    *f_wd_avg = 0.;
    v_wd_avg->x = 0.;
    v_wd_avg->y = 0.;
    if (factor > 0.0) {
        double wet_area = djac * factor;
        *f_wd_avg = averages[0].c_eq / wet_area;
        v_wd_avg->x = averages[0].x_eq / wet_area;
        v_wd_avg->y = averages[0].y_eq / wet_area;
    }
}

/* reference fe/sw2/fe_sw2_body_resid.c:1288-1354 (triangle branch) */
static void sw2_temporal(const v3 *nds, double djac, double dll, const double *h, const v2 *vel, double wd_factor,
                         double dt_factor, dof3 *elem_rhs, int wd_flag) {
    int i;
    double con = dt_factor * wd_factor;
    dof3 rhs[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    double uh[3], vh[3], rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
    for (i = 0; i < 3; i++) {
        uh[i] = vel[i].x * MAX(h[i], dll);
        vh[i] = vel[i].y * MAX(h[i], dll);
    }
    double vars[2] = {1.0, dt_factor};
    wd_wrapper(rhs, nds, h, NULL, NULL, NULL, NULL, djac, 1, vars, WD_CONT_TEMPORAL);
    int_tri_phi_f_lump(djac, con, uh, rx);
    int_tri_phi_f_lump(djac, con, vh, ry);
    for (i = 0; i < 3; i++) {
        if (wd_flag != 2) elem_rhs[i].c_eq += rhs[i].c_eq;
        elem_rhs[i].x_eq += rx[i];
        elem_rhs[i].y_eq += ry[i];
    }
}

/* reference fe/sw2/fe_sw2_get_EEVF.c:33-60 */
static void get_eevf(int mode, double coef, double g, double dll, double h_fric, double roughness, double area,
                     v2 gu, v2 gv, double avg_vmag, int wd_flag, double *ev_st, double *ev_tr) {
    double ev_min = 0., vmin;
    vmin = 0.1 * sqrt(g * dll);
    if (mode == 1) {
        *ev_st = 0.0;
        *ev_tr = coef * 2.0 * h_fric * sqrt(roughness) * avg_vmag;
        ev_min = (coef / 0.5) * vmin * sqrt(area) / 40.;
        *ev_tr = MAX(*ev_tr, ev_min);
    } else if (mode == 2) {
        *ev_st = 1.3 * h_fric * sqrt(roughness) * avg_vmag;
        *ev_tr = coef * 0.92 * h_fric * sqrt(roughness) * avg_vmag;
        ev_min = (coef / 0.5) * vmin * sqrt(area) / 40.;
        *ev_tr = MAX(*ev_tr, ev_min);
    } else if (mode == 3) {
        *ev_st = 0.;
        *ev_tr = coef * coef * area * sqrt(gu.x * gu.x + gv.y * gv.y + 0.5 * (gu.y + gv.x) * (gu.y + gv.x));
        ev_min = (coef / 0.2) * vmin * sqrt(area) / 40.;
        *ev_tr = MAX(*ev_tr, ev_min);
    }
    if (wd_flag == 1 || wd_flag == 2) {
        *ev_st = 0.0;
        *ev_tr = MAX(*ev_tr, 5. * ev_min);
    }
}

/* ---------------------------------------------------------------------- */
/* body residual, reference fe/sw2/fe_sw2_body_resid.c:45-1267            */
/* ---------------------------------------------------------------------- */
static void body_resid(o_model *m, dof3 *elem_rhs, int ie, double perturbation, int perturb_node,
                       int perturb_var, int perturb_sign) {
    const o_desc *d = &m->d;
    int i;
    int PRESSURE_FLAG = 1;
    const int *nd = d->e2d + 3 * ie;
    v2 grad_shp[3];
    for (i = 0; i < 3; i++) { grad_shp[i].x = m->grad[6 * ie + 2 * i]; grad_shp[i].y = m->grad[6 * ie + 2 * i + 1]; }
    double alpha = d->tau_temporal, dt = m->dt, djac = m->djac[ie], g = d->g;
    double dll = d->drying_lower, dul = d->drying_upper;
    int imat = d->e2d_mat[ie];
    int wd_flag = m->wd_flag[ie];
    double t1;

    /* :95-105 grid variables */
    double node_z[3];
    v3 elem_nds[3];
    for (i = 0; i < 3; i++) {
        elem_nds[i].x = d->x[nd[i]]; elem_nds[i].y = d->y[nd[i]]; elem_nds[i].z = d->z[nd[i]];
        node_z[i] = elem_nds[i].z;
    }
    /* :111-127 independent variables and perturbation */
    double elem_head[3];
    v2 elem_vel[3];
    for (i = 0; i < 3; i++) {
        elem_head[i] = m->head[nd[i]];
        elem_vel[i].x = m->vel[2 * nd[i]];
        elem_vel[i].y = m->vel[2 * nd[i] + 1];
    }
    if (perturb_var == PERTURB_H) {
        elem_head[perturb_node] += perturb_sign * perturbation;
    } else if (perturb_var == PERTURB_U) {
        PRESSURE_FLAG = 0;
        elem_vel[perturb_node].x += perturb_sign * perturbation;
    } else if (perturb_var == PERTURB_V) {
        PRESSURE_FLAG = 0;
        elem_vel[perturb_node].y += perturb_sign * perturbation;
    }
    /* :133-140 dependent variables */
    double elem_old_head[3], elem_older_head[3];
    v2 elem_old_vel[3], elem_older_vel[3];
    for (i = 0; i < 3; i++) {
        elem_old_head[i] = m->old_head[nd[i]];
        elem_older_head[i] = m->older_head[nd[i]];
        elem_old_vel[i].x = m->old_vel[2 * nd[i]]; elem_old_vel[i].y = m->old_vel[2 * nd[i] + 1];
        elem_older_vel[i].x = m->older_vel[2 * nd[i]]; elem_older_vel[i].y = m->older_vel[2 * nd[i] + 1];
    }
    /* baroclinic density is off (:142-149) and lids are absent (:152-174): both arrays are zero */
    double elem_density[3] = {0., 0., 0.};
    double elem_pressure[3] = {0., 0., 0.};

    double elem_u[3], elem_v[3], elem_old_u[3], elem_old_v[3];
    for (i = 0; i < 3; i++) {
        elem_u[i] = elem_vel[i].x; elem_v[i] = elem_vel[i].y;
        elem_old_u[i] = elem_old_vel[i].x; elem_old_v[i] = elem_old_vel[i].y;
    }
    /* :186-211 averages and gradients (triangle) */
    double area = djac;
    double elem_h_avg = int_tri_f(1., 1., elem_head);
    double elem_h_old_avg = int_tri_f(1., 1., elem_old_head);
    double elem_u_avg = int_tri_f(1., 1., elem_u);
    double elem_v_avg = int_tri_f(1., 1., elem_v);
    double elem_u_old_avg = int_tri_f(1., 1., elem_old_u);
    double elem_v_old_avg = int_tri_f(1., 1., elem_old_v);
    double elem_density_avg = int_tri_f(1., 1., elem_density);
    (void)elem_h_old_avg;
    v2 elem_grad_z = {0., 0.}, elem_grad_head = {0., 0.}, elem_grad_prs = {0., 0.};
    v2 elem_grad_u = {0., 0.}, elem_grad_v = {0., 0.}, elem_grad_wse;
    for (i = 0; i < 3; i++) { elem_grad_z.x += node_z[i] * grad_shp[i].x; elem_grad_z.y += node_z[i] * grad_shp[i].y; }
    for (i = 0; i < 3; i++) { elem_grad_head.x += elem_head[i] * grad_shp[i].x; elem_grad_head.y += elem_head[i] * grad_shp[i].y; }
    for (i = 0; i < 3; i++) { elem_grad_prs.x += elem_pressure[i] * grad_shp[i].x; elem_grad_prs.y += elem_pressure[i] * grad_shp[i].y; }
    for (i = 0; i < 3; i++) { /* grad2d_phi_dot_v fe_elem_utilities.c:685-692 */
        elem_grad_u.x += elem_vel[i].x * grad_shp[i].x; elem_grad_u.y += elem_vel[i].x * grad_shp[i].y;
        elem_grad_v.x += elem_vel[i].y * grad_shp[i].x; elem_grad_v.y += elem_vel[i].y * grad_shp[i].y;
    }
    elem_grad_wse.x = elem_grad_z.x + elem_grad_head.x;
    elem_grad_wse.y = elem_grad_z.y + elem_grad_head.y;

    double elem_vel_avg_mag = sqrt(elem_u_avg * elem_u_avg + elem_v_avg * elem_v_avg);
    double vst_i = elem_u_avg / MAX(elem_vel_avg_mag, 1.E-8);
    double vst_j = elem_v_avg / MAX(elem_vel_avg_mag, 1.E-8);
    (void)elem_u_old_avg; (void)elem_v_old_avg;

    /* :236-252 wet/dry factors */
    double factor, factor_old, factor_older;
    double elem_h_wd_avg = elem_h_avg, elem_h_old_wd_avg;
    v2 elem_vel_wd_avg, elem_old_vel_wd_avg;
    wd_average(elem_nds, elem_head, elem_vel, elem_head, djac, &elem_h_wd_avg, &elem_vel_wd_avg);
    wd_average(elem_nds, elem_old_head, elem_old_vel, elem_old_head, djac, &elem_h_old_wd_avg, &elem_old_vel_wd_avg);
    factor = wd_factor(elem_nds, elem_head, djac);
    factor_old = wd_factor(elem_nds, elem_old_head, djac);
    factor_older = wd_factor(elem_nds, elem_older_head, djac);
    double elem_vel_wd_avg_mag = sqrt(elem_vel_wd_avg.x * elem_vel_wd_avg.x + elem_vel_wd_avg.y * elem_vel_wd_avg.y);

    /* :256-258 friction variables */
    double h_fric = MAX(elem_h_wd_avg, SMALL);
    double roughness = get_roughness(d, ie, h_fric);

    /* :261-289 eddy viscosity */
    double ev_st = 0., ev_tr = 0.;
    if (d->mat_turb[imat] == O_TURB_EVS) {
        double exx = d->mat_ev[3 * imat], eyy = d->mat_ev[3 * imat + 1];
        exx += d->viscosity;
        eyy += d->viscosity;
        ev_st = fabs(exx - eyy);
        ev_tr = MIN(exx, eyy);
    } else if (d->mat_turb[imat] == O_TURB_EEV) {
        get_eevf(d->mat_eev_mode[imat], d->mat_eev_coef[imat], g, dll, h_fric, roughness, area, elem_grad_u,
                 elem_grad_v, elem_vel_avg_mag, wd_flag, &ev_st, &ev_tr);
    }

    /* :340-348 */
    dof3 rhs[3];
    for (i = 0; i < 3; i++) { elem_rhs[i].x_eq = 0.; elem_rhs[i].y_eq = 0.; elem_rhs[i].c_eq = 0.; }
    double rhs_c_eq[3], rhs_x_eq[3], rhs_y_eq[3];
    double vars[2];

    /* :360-419 shock capturing on partially/fully dry elements */
    if (wd_flag == 1 || wd_flag == 2) {
        double elem_u_grad_mag = sqrt(elem_grad_u.x * elem_grad_u.x + elem_grad_u.y * elem_grad_u.y + SMALL);
        double elem_v_grad_mag = sqrt(elem_grad_v.x * elem_grad_v.x + elem_grad_v.y * elem_grad_v.y + SMALL);
        double sh_cap_coef_u = one_2 * dul * djac * elem_u_grad_mag;
        double sh_cap_coef_v = one_2 * dul * djac * elem_v_grad_mag;
        v2 gun, gvn, norm;
        gun.x = elem_grad_u.x / elem_u_grad_mag; gun.y = elem_grad_u.y / elem_u_grad_mag;
        gvn.x = elem_grad_v.x / elem_v_grad_mag; gvn.y = elem_grad_v.y / elem_v_grad_mag;
        for (i = 0; i < 3; i++) {
            rhs[i].c_eq = dt * one_2 * g * djac * ((dll * elem_grad_wse.x) * grad_shp[i].x + (dll * elem_grad_wse.y) * grad_shp[i].y);
            norm.x = gun.x * elem_grad_u.x + gun.y * elem_grad_u.y;
            norm.y = gvn.x * elem_grad_v.x + gvn.y * elem_grad_v.y;
            rhs[i].x_eq = sh_cap_coef_u * dt * djac * norm.x * (gun.x * grad_shp[i].x + gun.y * grad_shp[i].y);
            rhs[i].y_eq = sh_cap_coef_v * dt * djac * norm.y * (gvn.x * grad_shp[i].x + gvn.y * grad_shp[i].y);
        }
        for (i = 0; i < 3; i++) {
            elem_rhs[i].c_eq += rhs[i].c_eq;
            elem_rhs[i].x_eq += rhs[i].x_eq;
            elem_rhs[i].y_eq += rhs[i].y_eq;
        }
    }

    /* :434-461 advection */
    for (i = 0; i < 3; i++) { rhs[i].x_eq = 0.; rhs[i].y_eq = 0.; rhs[i].c_eq = 0.; }
    {
        v2 elem_vel_wd[3];
        for (i = 0; i < 3; i++) elem_vel_wd[i] = elem_vel[i];
        vars[0] = dt;
        vars[1] = -1;
        wd_wrapper(rhs, elem_nds, elem_head, grad_shp, elem_vel, elem_vel_wd, NULL, djac, 1, vars, WD_CONVECTION);
    }
    for (i = 0; i < 3; i++) {
        elem_rhs[i].c_eq += rhs[i].c_eq;
        elem_rhs[i].x_eq += rhs[i].x_eq;
        elem_rhs[i].y_eq += rhs[i].y_eq;
    }

    /* :480-490 temporal */
    sw2_temporal(elem_nds, djac, dll, elem_head, elem_vel, factor, (1. + alpha / 2.), elem_rhs, -3);
    sw2_temporal(elem_nds, djac, dll, elem_old_head, elem_old_vel, factor_old, (-1.) * (1. + alpha), elem_rhs, wd_flag);
    sw2_temporal(elem_nds, djac, dll, elem_older_head, elem_older_vel, factor_older, (alpha / 2.), elem_rhs, wd_flag);

    /* :504-539 diffusion */
    double temp = elem_h_wd_avg;
    if (elem_h_wd_avg < dul) temp = dul;
    for (i = 0; i < 3; i++) { rhs_x_eq[i] = 0.; rhs_y_eq[i] = 0.; }
    {
        v2 dfx, dfy;
        dfx.x = 2. * ev_st * vst_i * (vst_i * elem_grad_u.x + vst_j * elem_grad_u.y) + 2. * ev_tr * elem_grad_u.x;
        dfx.y = ev_st * vst_j * (vst_j * elem_grad_u.y + vst_i * elem_grad_u.x + vst_i * elem_grad_v.x + vst_j * elem_grad_v.y) + ev_tr * (elem_grad_u.y + elem_grad_v.x);
        dfy.x = ev_st * vst_i * (vst_j * elem_grad_u.y + vst_i * elem_grad_u.x + vst_i * elem_grad_v.x + vst_j * elem_grad_v.y) + ev_tr * (elem_grad_u.y + elem_grad_v.x);
        dfy.y = 2. * ev_st * vst_j * (vst_j * elem_grad_v.y + vst_i * elem_grad_v.x) + 2. * ev_tr * elem_grad_v.y;
        int_tri_gradPhi_dot_vbar(grad_shp, djac, temp, dfx, rhs_x_eq);
        int_tri_gradPhi_dot_vbar(grad_shp, djac, temp, dfy, rhs_y_eq);
    }
    t1 = dt * factor;
    for (i = 0; i < 3; i++) {
        elem_rhs[i].x_eq += t1 * rhs_x_eq[i];
        elem_rhs[i].y_eq += t1 * rhs_y_eq[i];
    }

    /* :557-594 friction */
    double resistance = roughness / (2.0 * g);
    double fx[3], fy[3];
    for (i = 0; i < 3; i++) {
        t1 = sqrt(elem_vel[i].x * elem_vel[i].x + elem_vel[i].y * elem_vel[i].y);
        fx[i] = elem_vel[i].x * t1;
        fy[i] = elem_vel[i].y * t1;
    }
    for (i = 0; i < 3; i++) { rhs_x_eq[i] = 0.; rhs_y_eq[i] = 0.; }
    t1 = dt * g * resistance * factor;
    int_tri_phi_f(djac, t1, fx, rhs_x_eq);
    int_tri_phi_f(djac, t1, fy, rhs_y_eq);
    for (i = 0; i < 3; i++) {
        elem_rhs[i].x_eq += rhs_x_eq[i];
        elem_rhs[i].y_eq += rhs_y_eq[i];
    }
    v2 elem_bed_stress;
    elem_bed_stress.x = resistance * elem_vel_wd_avg.x * elem_vel_wd_avg_mag;
    elem_bed_stress.y = resistance * elem_vel_wd_avg.y * elem_vel_wd_avg_mag;

    /* :639-652 coriolis */
    for (i = 0; i < 3; i++) { rhs_x_eq[i] = 0.; rhs_y_eq[i] = 0.; }
    double coriolis_speed = d->mat_cor_speed[imat];
    t1 = dt * factor * coriolis_speed;
    int_tri_phi_f_g(djac, -t1, elem_head, elem_v, rhs_x_eq);
    int_tri_phi_f_g(djac, +t1, elem_head, elem_u, rhs_y_eq);
    for (i = 0; i < 3; i++) {
        elem_rhs[i].x_eq += rhs_x_eq[i];
        elem_rhs[i].y_eq += rhs_y_eq[i];
    }

    /* :674-719 wind and wave stress on fully wet triangles.  Without winds
       (and there are never waves here) the contribution elem_rhs -= dt * 0.0
       is an exact no-op and is skipped. */
    v2 total_stress = {0., 0.};
    if (wd_flag == 0 && d->wind) {
        /* swind_elem2d_get_local_stress, src/structs/smeteor.c:241-308, stress law 0
           (the node value is the shear); pow(windatt, 2.0) == windatt * windatt
           for the attenuations the decks use (1.0) */
        v2 wind_stress = {0., 0.};
        for (i = 0; i < 3; i++) {
            double shear_x = 0.0, shear_y = 0.0;
            if (d->density > 0.0) {
                shear_x = m->wind[2 * nd[i]];
                shear_y = m->wind[2 * nd[i] + 1];
            }
            wind_stress.x += shear_x;
            wind_stress.y += shear_y;
        }
        double watt = d->mat_windatt ? d->mat_windatt[imat] : 1.0;
        wind_stress.x *= (watt * watt) / (3 * d->density);
        wind_stress.y *= (watt * watt) / (3 * d->density);
        total_stress.x = wind_stress.x + 0.0;   /* + wave stress */
        total_stress.y = wind_stress.y + 0.0;
        for (i = 0; i < 3; i++) { rhs_x_eq[i] = 0.; rhs_y_eq[i] = 0.; }
        /* integrate_triangle_phi, fe_elem2d_integrations.c:159-164 */
        t1 = total_stress.x * djac * one_3;
        for (i = 0; i < 3; i++) rhs_x_eq[i] += t1;
        t1 = total_stress.y * djac * one_3;
        for (i = 0; i < 3; i++) rhs_y_eq[i] += t1;
        for (i = 0; i < 3; i++) {
            elem_rhs[i].x_eq -= dt * rhs_x_eq[i];
            elem_rhs[i].y_eq -= dt * rhs_y_eq[i];
        }
    }

    /* :728-754 rainfall */
    double elem_water_source = 0.;
    if (d->e2d_src_flag[ie]) {
        elem_water_source = d->e2d_src[ie];
        for (i = 0; i < 3; i++) rhs_c_eq[i] = 0.;
        int_tri_phi(djac, elem_water_source, rhs_c_eq);
        t1 = dt * factor;
        for (i = 0; i < 3; i++) elem_rhs[i].c_eq -= t1 * rhs_c_eq[i];
    }

    /* :761-862 pressure, bed slope, boundary pressure */
    if (PRESSURE_FLAG) {
        for (i = 0; i < 3; i++) { rhs[i].x_eq = 0.; rhs[i].y_eq = 0.; rhs[i].c_eq = 0.; }
        t1 = -one_2 * g * dt;
        vars[0] = 1; vars[1] = t1;
        wd_wrapper(rhs, elem_nds, elem_head, grad_shp, NULL, NULL, NULL, djac, 1, vars, WD_PRESSURE);
        for (i = 0; i < 3; i++) {
            elem_rhs[i].x_eq = elem_rhs[i].x_eq + rhs[i].x_eq;
            elem_rhs[i].y_eq = elem_rhs[i].y_eq + rhs[i].y_eq;
            elem_rhs[i].c_eq = elem_rhs[i].c_eq + rhs[i].c_eq;
        }
        for (i = 0; i < 3; i++) { rhs[i].x_eq = 0.; rhs[i].y_eq = 0.; rhs[i].c_eq = 0.; }
        t1 = g * dt;
        vars[0] = 1; vars[1] = t1;
        wd_wrapper(rhs, elem_nds, elem_head, NULL, &elem_grad_z, NULL, NULL, djac, 1, vars, WD_BODYFORCE);
        for (i = 0; i < 3; i++) {
            elem_rhs[i].x_eq = elem_rhs[i].x_eq + rhs[i].x_eq;
            elem_rhs[i].y_eq = elem_rhs[i].y_eq + rhs[i].y_eq;
            elem_rhs[i].c_eq = elem_rhs[i].c_eq + rhs[i].c_eq;
        }
        for (i = 0; i < 3; i++) { rhs[i].x_eq = 0.; rhs[i].y_eq = 0.; rhs[i].c_eq = 0.; }
        t1 = one_2 * dt * g;
        vars[0] = 1; vars[1] = t1;
        wd_wrapper(rhs, elem_nds, elem_head, NULL, NULL, NULL, NULL, djac, 1, vars, WD_BOUNDARY_PRESSURE);
        for (i = 0; i < 3; i++) {
            elem_rhs[i].x_eq = elem_rhs[i].x_eq + rhs[i].x_eq;
            elem_rhs[i].y_eq = elem_rhs[i].y_eq + rhs[i].y_eq;
            elem_rhs[i].c_eq = elem_rhs[i].c_eq + rhs[i].c_eq;
        }
    }

    /* :987-1232 SUPG */
    if (fabs(elem_h_wd_avg) >= SMALL) {
        double c_eq, x_eq, y_eq;
        double g_factor = (1 + elem_density_avg) * g;
        double denom = MAX(elem_h_wd_avg, NOT_QUITE_SMALL);
        /* fe_get_supg_tau_sw fe_get_supg_tau.c:43-80 with le = sqrt(area*factor_old) (:289) */
        double tau;
        {
            double le = sqrt(area * factor_old);
            double celerity_squared = g_factor * denom;
            if (denom <= 0.0) celerity_squared = 0.;
            double spectral_radius = sqrt(elem_vel_wd_avg.x * elem_vel_wd_avg.x + elem_vel_wd_avg.y * elem_vel_wd_avg.y + celerity_squared + SMALL);
            tau = d->tau_pg * le / spectral_radius;
        }
        dof3 dfdt[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
        dof3 advection[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
        double dhdt, dudt, dvdt;
        vars[0] = 1; vars[1] = 1;
        factor_old = wd_wrapper(dfdt, elem_nds, elem_old_head, NULL, NULL, elem_old_vel, elem_old_head, djac, 0, vars, WD_INTEGRATE_F);
        dudt = -dfdt[0].x_eq; dvdt = -dfdt[0].y_eq; dhdt = -dfdt[0].c_eq;
        for (i = 0; i < 3; i++) { dfdt[i].x_eq = 0.; dfdt[i].y_eq = 0.; dfdt[i].c_eq = 0.; }
        vars[0] = 1; vars[1] = 1;
        factor = wd_wrapper(dfdt, elem_nds, elem_head, NULL, NULL, elem_vel, elem_head, djac, 0, vars, WD_INTEGRATE_F);
        dudt += dfdt[0].x_eq; dvdt += dfdt[0].y_eq; dhdt += dfdt[0].c_eq;
        vars[0] = dt; vars[1] = 1;
        factor = wd_wrapper(advection, elem_nds, elem_head, grad_shp, NULL, elem_vel, elem_head, djac, 0, vars, WD_GLS_CONVECTION);
        double x_pressure = g_factor * (elem_grad_head.x + elem_grad_z.x) * dt * factor + elem_grad_prs.x * dt * factor;
        double y_pressure = g_factor * (elem_grad_head.y + elem_grad_z.y) * dt * factor + elem_grad_prs.y * dt * factor;
        double x_friction = g * elem_bed_stress.x * dt * factor;
        double y_friction = g * elem_bed_stress.y * dt * factor;
        double x_coriolis = -elem_vel_wd_avg.y * coriolis_speed;
        double y_coriolis = elem_vel_wd_avg.x * coriolis_speed;
        c_eq = dhdt + advection[0].c_eq + factor * elem_water_source * djac * dt;
        x_eq = dudt + advection[0].x_eq + (x_pressure + x_friction / denom) * djac + (x_coriolis - total_stress.x / denom) * djac * factor * dt;
        y_eq = dvdt + advection[0].y_eq + (y_pressure + y_friction / denom) * djac + (y_coriolis - total_stress.y / denom) * djac * factor * dt;
        /* :1126-1139: the second set of assignments is the one that stands */
        for (i = 0; i < 3; i++) {
            rhs_c_eq[i] = (grad_shp[i].x * (tau * elem_vel_wd_avg.x * c_eq + tau * elem_h_wd_avg * x_eq) +
                           grad_shp[i].y * (tau * elem_vel_wd_avg.y * c_eq + tau * elem_h_wd_avg * y_eq));
            rhs_x_eq[i] = (grad_shp[i].x * (tau * g_factor * elem_h_wd_avg * c_eq + tau * elem_vel_wd_avg.x * elem_h_wd_avg * x_eq) +
                           grad_shp[i].y * (tau * elem_vel_wd_avg.y * elem_h_wd_avg * x_eq));
            rhs_y_eq[i] = (grad_shp[i].y * (tau * g_factor * elem_h_wd_avg * c_eq + tau * elem_vel_wd_avg.y * elem_h_wd_avg * y_eq) +
                           grad_shp[i].x * (tau * elem_vel_wd_avg.x * elem_h_wd_avg * y_eq));
        }
        for (i = 0; i < 3; i++) {
            elem_rhs[i].c_eq += rhs_c_eq[i];
            elem_rhs[i].x_eq += rhs_x_eq[i];
            elem_rhs[i].y_eq += rhs_y_eq[i];
        }
    }

    /* :1239-1256 Dirichlet */
    for (i = 0; i < 3; i++) {
        int bc = d->node_bc ? d->node_bc[nd[i]] : 0;
        if (bc == O_BCT_PRS_DIR) {
            elem_rhs[i].c_eq = 0.0;
        } else if (bc == O_BCT_VEL_DIR) {
            elem_rhs[i].x_eq = 0.0;
            elem_rhs[i].y_eq = 0.0;
        } else if (bc == O_BCT_VEL_PRS_DIR) {
            elem_rhs[i].c_eq = 0.0;
            elem_rhs[i].x_eq = 0.0;
            elem_rhs[i].y_eq = 0.0;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* boundary residual, reference fe/sw2/fe_sw2_boundary_resid.c:72-905     */
/* ---------------------------------------------------------------------- */
static void boundary_resid(o_model *m, dof3 *elem_rhs, int ie, double perturbation, int perturb_node,
                           int perturb_var, int perturb_sign) {
    const o_desc *d = &m->d;
    int i;
    const int *nd = d->e1d + 2 * ie;
    v2 nrml = {m->nrml1d[2 * ie], m->nrml1d[2 * ie + 1]};
    double dt = m->dt, djac = m->djac1d[ie], g = d->g;
    int PRESSURE_FLAG = 1;
    if (perturb_var == PERTURB_NONE) perturb_sign = 0;

    double elem_head[2];
    for (i = 0; i < 2; i++) elem_head[i] = m->head[nd[i]];
    if (perturb_var == PERTURB_H) elem_head[perturb_node] += perturb_sign * perturbation;
    v2 elem_vel[2];
    for (i = 0; i < 2; i++) { elem_vel[i].x = m->vel[2 * nd[i]]; elem_vel[i].y = m->vel[2 * nd[i] + 1]; }
    if (perturb_var == PERTURB_U) {
        elem_vel[perturb_node].x += perturb_sign * perturbation;
        PRESSURE_FLAG = 0;
    }
    if (perturb_var == PERTURB_V) {
        elem_vel[perturb_node].y += perturb_sign * perturbation;
        PRESSURE_FLAG = 0;
    }
    double u[2], v[2];
    for (i = 0; i < 2; i++) { u[i] = elem_vel[i].x; v[i] = elem_vel[i].y; }
    double h_avg = one_2 * (elem_head[0] + elem_head[1]);
    if (h_avg < SMALL) h_avg = SMALL;
    double elem_node_z[2];
    elem_node_z[0] = d->z[nd[0]] + 0.0; /* + bed_displacement, zero without sediment */
    elem_node_z[1] = d->z[nd[1]] + 0.0;

    for (i = 0; i < 2; i++) { elem_rhs[i].x_eq = 0.; elem_rhs[i].y_eq = 0.; elem_rhs[i].c_eq = 0.; }

    int bc = d->e1d_bc[ie];
    if (bc == O_BCT_VEL_NEU) { /* :212-219, :562-604 */
        double flux = -1.0 * d->e1d_val[ie];
        double rc[2] = {0., 0.}, rx[2] = {0., 0.}, ry[2] = {0., 0.};
        int_line_phi(djac, 1., rc);
        if (flux < 0) {
            double u_star = 0., v_star = 0.;
            if (h_avg > 0) {
                u_star = flux * nrml.x / h_avg;
                v_star = flux * nrml.y / h_avg;
            }
            int_line_phi(djac, u_star, rx);
            int_line_phi(djac, v_star, ry);
        } else {
            int_line_phi_f(djac, 1., u, rx);
            int_line_phi_f(djac, 1., v, ry);
        }
        double t1 = dt * flux;
        for (i = 0; i < 2; i++) {
            elem_rhs[i].c_eq += t1 * rc[i];
            elem_rhs[i].x_eq += t1 * rx[i];
            elem_rhs[i].y_eq += t1 * ry[i];
        }
    } else if (bc == O_BCT_OUTFLOW || bc == O_BCT_PRS_NEU) {
        if (bc == O_BCT_PRS_NEU) { /* :364-391 */
            double tail_elev = d->e1d_val[ie];
            double elem_depth[2];
            elem_depth[0] = tail_elev - elem_node_z[0];
            elem_depth[1] = tail_elev - elem_node_z[0]; /* node 0's bed for both, as the reference */
            if (elem_depth[0] < SMALL || elem_depth[1] < SMALL || elem_head[0] < SMALL || elem_head[1] < SMALL) return;
            (void)elem_depth;
        }
        /* fe_sw2_1d_implicit_flow :637-687 */
        {
            double rc[2] = {0., 0.}, rx[2] = {0., 0.}, ry[2] = {0., 0.};
            double u1 = elem_vel[0].x, u2 = elem_vel[1].x;
            double v1 = elem_vel[0].y, v2_ = elem_vel[1].y;
            double h1 = elem_head[0], h2 = elem_head[1];
            double ur1 = elem_vel[0].x, ur2 = elem_vel[1].x;
            double vr1 = elem_vel[0].y, vr2 = elem_vel[1].y;
            double dir0 = elem_vel[0].x * nrml.x + elem_vel[0].y * nrml.y;
            double dir1 = elem_vel[1].x * nrml.x + elem_vel[1].y * nrml.y;
            if (dir0 < SMALL) { ur1 = 0.; vr1 = 0.; }
            if (dir1 < SMALL) { ur2 = 0.; vr2 = 0.; }
            int_line_phi_h_v_dot_n(djac, 1., elem_head, elem_vel, nrml, rc);
            int_line_phi_f_g_h(djac, nrml.x, ur1, ur2, u1, u2, h1, h2, rx);
            int_line_phi_f_g_h(djac, nrml.y, vr1, vr2, v1, v2_, h1, h2, ry);
            int_line_phi_f_g_h(djac, nrml.y, ur1, ur2, v1, v2_, h1, h2, rx);
            int_line_phi_f_g_h(djac, nrml.x, vr1, vr2, u1, u2, h1, h2, ry);
            for (i = 0; i < 2; i++) {
                elem_rhs[i].c_eq += dt * rc[i];
                elem_rhs[i].x_eq += dt * rx[i];
                elem_rhs[i].y_eq += dt * ry[i];
            }
        }
        if (bc == O_BCT_PRS_NEU && PRESSURE_FLAG) { /* :381-384, :773-798 */
            double tail_elev = d->e1d_val[ie];
            double elem_depth[2];
            elem_depth[0] = tail_elev - elem_node_z[0];
            elem_depth[1] = tail_elev - elem_node_z[0];
            double gg[2] = {g, -g};
            const double *hh[2] = {elem_depth, elem_head};
            int k;
            for (k = 0; k < 2; k++) {
                double r[2] = {0., 0.};
                int_line_phi_h_h(djac, 1., hh[k][0], hh[k][1], r);
                double t1 = dt * one_2 * gg[k];
                for (i = 0; i < 2; i++) {
                    elem_rhs[i].x_eq += t1 * nrml.x * r[i];
                    elem_rhs[i].y_eq += t1 * nrml.y * r[i];
                }
            }
        }
    }
    /* :483-498 Dirichlet */
    for (i = 0; i < 2; i++) {
        int nbc = d->node_bc ? d->node_bc[nd[i]] : 0;
        if (nbc == O_BCT_PRS_DIR) elem_rhs[i].c_eq = 0.0;
        else if (nbc == O_BCT_VEL_DIR) { elem_rhs[i].x_eq = 0.0; elem_rhs[i].y_eq = 0.0; }
        else if (nbc == O_BCT_VEL_PRS_DIR) { elem_rhs[i].c_eq = 0.0; elem_rhs[i].x_eq = 0.0; elem_rhs[i].y_eq = 0.0; }
    }
}

/* ---------------------------------------------------------------------- */
/* model set-up                                                            */
/* ---------------------------------------------------------------------- */
static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) abort();
    return p;
}

o_model *o_create(const o_desc *d) {
    o_model *m = (o_model *)xcalloc(1, sizeof(o_model));
    int i;
    m->d = *d;
    m->N = d->nnodes; m->E = d->nelems2d; m->B = d->nelems1d;
    m->djac = xcalloc(m->E, sizeof(double));
    m->grad = xcalloc(6 * (size_t)m->E, sizeof(double));
    m->djac1d = xcalloc(m->B, sizeof(double));
    m->nrml1d = xcalloc(2 * (size_t)m->B, sizeof(double));
    for (i = 0; i < m->E; i++) {
        const int *t = d->e2d + 3 * i;
        tri_geometry(d->x[t[0]], d->x[t[1]], d->x[t[2]], d->y[t[0]], d->y[t[1]], d->y[t[2]], m->djac + i, m->grad + 6 * i);
    }
    for (i = 0; i < m->B; i++) edge_geometry(d, i, m->djac1d + i, m->nrml1d + 2 * i);
    size_t N = m->N;
    m->head = xcalloc(N, 8); m->vel = xcalloc(2 * N, 8);
    m->old_head = xcalloc(N, 8); m->old_vel = xcalloc(2 * N, 8);
    m->older_head = xcalloc(N, 8); m->older_vel = xcalloc(2 * N, 8);
    m->wd_flag = xcalloc(m->E, sizeof(int));
    m->bc_mask = xcalloc(3 * N, sizeof(int));
    m->residual = xcalloc(3 * N, 8);
    m->sol = xcalloc(3 * N, 8);
    m->diagonal = xcalloc(9 * N, 8);
    m->scale_vect = xcalloc(3 * N, 8);
    for (i = 0; i < 3 * m->N; i++) m->scale_vect[i] = 1.0; /* fe_main.c:263 */
    m->matrix = xcalloc(N, sizeof(spv));
    m->node_block = xcalloc(N, sizeof(int));
    m->bjinv = xcalloc(9 * N, 8);
    m->wind = xcalloc(2 * N, 8);
    m->val1d = xcalloc(m->B > 0 ? m->B : 1, 8);
    m->src2d = xcalloc(m->E > 0 ? m->E : 1, 8);
    if (d->e1d_val && m->B) memcpy(m->val1d, d->e1d_val, 8 * (size_t)m->B);
    if (d->e2d_src && m->E) memcpy(m->src2d, d->e2d_src, 8 * (size_t)m->E);
    m->d.e1d_val = m->val1d;
    m->d.e2d_src = m->src2d;
    m->dt = d->dt;
    return m;
}

void o_destroy(o_model *m) {
    int i;
    if (!m) return;
    for (i = 0; i < m->N; i++) { free(m->matrix[i].index); free(m->matrix[i].value); }
    for (i = 0; i < m->prof_size; i++) { free(m->rows[i].value); free(m->cols[i].value); }
    free(m->rows); free(m->cols);
    free(m->djac); free(m->grad); free(m->djac1d); free(m->nrml1d);
    free(m->head); free(m->vel); free(m->old_head); free(m->old_vel); free(m->older_head); free(m->older_vel);
    free(m->wd_flag); free(m->bc_mask); free(m->residual); free(m->sol); free(m->diagonal); free(m->scale_vect);
    free(m->matrix); free(m->node_block); free(m->bjinv); free(m->wind); free(m->val1d); free(m->src2d);
    free(m);
}

void o_get_geometry(const o_model *m, double *djac, double *grad, double *djac1d, double *nrml1d) {
    if (djac) memcpy(djac, m->djac, sizeof(double) * m->E);
    if (grad) memcpy(grad, m->grad, sizeof(double) * 6 * m->E);
    if (djac1d) memcpy(djac1d, m->djac1d, sizeof(double) * m->B);
    if (nrml1d) memcpy(nrml1d, m->nrml1d, sizeof(double) * 2 * m->B);
}

void o_set_dt(o_model *m, double dt) { m->dt = dt; }

void o_set_wind(o_model *m, const double *stress_xy) { memcpy(m->wind, stress_xy, 16 * (size_t)m->N); }

/* series values averaged over the coming step, sseries_setall_ts_valuesAVG
   (src/main/adh_func.c:542, src/structs/sseries.c:618), read back by
   sseries_get_value in the residuals */
void o_set_series(o_model *m, const double *e1d_val, const double *e2d_src) {
    if (e1d_val && m->B) memcpy(m->val1d, e1d_val, 8 * (size_t)m->B);
    if (e2d_src && m->E) memcpy(m->src2d, e2d_src, 8 * (size_t)m->E);
}

void o_set_state(o_model *m, const double *head, const double *vel, const double *old_head, const double *old_vel,
                 const double *older_head, const double *older_vel) {
    size_t N = m->N;
    memcpy(m->head, head, 8 * N); memcpy(m->vel, vel, 16 * N);
    memcpy(m->old_head, old_head, 8 * N); memcpy(m->old_vel, old_vel, 16 * N);
    memcpy(m->older_head, older_head, 8 * N); memcpy(m->older_vel, older_vel, 16 * N);
}

void o_get_state(const o_model *m, double *head, double *vel, double *old_head, double *old_vel, double *older_head,
                 double *older_vel) {
    size_t N = m->N;
    if (head) memcpy(head, m->head, 8 * N);
    if (vel) memcpy(vel, m->vel, 16 * N);
    if (old_head) memcpy(old_head, m->old_head, 8 * N);
    if (old_vel) memcpy(old_vel, m->old_vel, 16 * N);
    if (older_head) memcpy(older_head, m->older_head, 8 * N);
    if (older_vel) memcpy(older_vel, m->older_vel, 16 * N);
}

/* reference fe/sw2/fe_sw2.c:42-50 */
void o_shift(o_model *m) {
    int i;
    for (i = 0; i < m->N; i++) {
        m->older_head[i] = m->old_head[i];
        m->older_vel[2 * i] = m->old_vel[2 * i];
        m->older_vel[2 * i + 1] = m->old_vel[2 * i + 1];
        m->old_head[i] = m->head[i];
        m->old_vel[2 * i] = m->vel[2 * i];
        m->old_vel[2 * i + 1] = m->vel[2 * i + 1];
    }
}

/* reference fe/sw2/fe_sw2_wdflag.c:24-75 */
void o_wdflag(o_model *m, int *out) {
    const o_desc *d = &m->d;
    int i, ie, e_wet;
    int *wd_tmp = xcalloc(m->N, sizeof(int));
    for (ie = 0; ie < m->E; ie++) {
        e_wet = 0;
        for (i = 0; i < 3; i++) if (m->head[d->e2d[3 * ie + i]] > 0.0) e_wet++;
        if (e_wet != 3) for (i = 0; i < 3; i++) wd_tmp[d->e2d[3 * ie + i]] = 1;
    }
    for (ie = 0; ie < m->E; ie++) {
        const int *t = d->e2d + 3 * ie;
        m->wd_flag[ie] = 0;
        for (i = 0; i < 3; i++) if (wd_tmp[t[i]] != 0) m->wd_flag[ie] = 1;
        if (m->head[t[0]] < 0 && m->head[t[1]] < 0 && m->head[t[2]] < 0 &&
            m->old_head[t[0]] < 0 && m->old_head[t[1]] < 0 && m->old_head[t[2]] < 0)
            m->wd_flag[ie] = 2;
    }
    free(wd_tmp);
    if (out) memcpy(out, m->wd_flag, sizeof(int) * m->E);
}

/* reference fe/sw2/fe_sw2_init.c:24-92 */
void o_init(o_model *m, int *bc_mask_out) {
    const o_desc *d = &m->d;
    int i;
    for (i = 0; i < m->N; i++) {
        m->vel[2 * i] = m->old_vel[2 * i];
        m->vel[2 * i + 1] = m->old_vel[2 * i + 1];
        m->head[i] = m->old_head[i];
    }
    for (i = 0; i < m->N; i++) {
        int s = 3 * i;
        int bc = d->node_bc ? d->node_bc[i] : 0;
        m->bc_mask[s] = -3; m->bc_mask[s + 1] = -3; m->bc_mask[s + 2] = -3;
        if (bc == O_BCT_VEL_DIR) {
            m->bc_mask[s] = 1; m->bc_mask[s + 1] = 1;
            m->vel[2 * i] = d->node_bc_val[s]; m->vel[2 * i + 1] = d->node_bc_val[s + 1];
        } else if (bc == O_BCT_PRS_DIR) {
            m->bc_mask[s + 2] = 1;
            m->head[i] = d->node_bc_val[s + 2];
        } else if (bc == O_BCT_VEL_PRS_DIR) {
            m->bc_mask[s] = 1; m->bc_mask[s + 1] = 1; m->bc_mask[s + 2] = 1;
            m->vel[2 * i] = d->node_bc_val[s]; m->vel[2 * i + 1] = d->node_bc_val[s + 1];
            m->head[i] = d->node_bc_val[s + 2];
        }
    }
    o_wdflag(m, NULL);
    if (bc_mask_out) memcpy(bc_mask_out, m->bc_mask, sizeof(int) * 3 * m->N);
}

void o_body_resid(o_model *m, int ie, double pert, int pn, int pv, int ps, double *out) {
    dof3 r[3];
    body_resid(m, r, ie, pert, pn, pv, ps);
    memcpy(out, r, sizeof(r));
}

void o_boundary_resid(o_model *m, int ie, double pert, int pn, int pv, int ps, double *out) {
    dof3 r[2];
    boundary_resid(m, r, ie, pert, pn, pv, ps);
    memcpy(out, r, sizeof(r));
}

/* reference fe/sw2/fe_sw2_resid.c:29-216 */
void o_resid(o_model *m, double *out) {
    const o_desc *d = &m->d;
    int ie, i, j;
    dof3 elem_rhs[3];
    memset(m->residual, 0, sizeof(double) * 3 * m->N);
    for (ie = 0; ie < m->E; ie++) {
        body_resid(m, elem_rhs, ie, 0, 0, PERTURB_NONE, 0);
        for (i = 0; i < 3; i++) {
            int n = d->e2d[3 * ie + i];
            if (m->old_head[n] <= 0.) {
                elem_rhs[i].x_eq = 0.0;
                elem_rhs[i].y_eq = 0.0;
                m->vel[2 * n] = 0.;
                m->vel[2 * n + 1] = 0.;
            }
        }
        for (i = 0; i < 3; i++) {
            j = 3 * d->e2d[3 * ie + i];
            m->residual[j] -= elem_rhs[i].x_eq;
            m->residual[j + 1] -= elem_rhs[i].y_eq;
            m->residual[j + 2] -= elem_rhs[i].c_eq;
        }
    }
    for (ie = 0; ie < m->B; ie++) {
        boundary_resid(m, elem_rhs, ie, 0, 0, PERTURB_NONE, 0);
        for (i = 0; i < 2; i++) {
            int n = d->e1d[2 * ie + i];
            if (m->old_head[n] <= 0.) {
                elem_rhs[i].x_eq = 0.0;
                elem_rhs[i].y_eq = 0.0;
                m->vel[2 * n] = 0.;
                m->vel[2 * n + 1] = 0.;
            }
        }
        for (i = 0; i < 2; i++) {
            j = 3 * d->e1d[2 * ie + i];
            m->residual[j] -= elem_rhs[i].x_eq;
            m->residual[j + 1] -= elem_rhs[i].y_eq;
            m->residual[j + 2] -= elem_rhs[i].c_eq;
        }
    }
    if (out) memcpy(out, m->residual, sizeof(double) * 3 * m->N);
}

void o_get_residual(const o_model *m, double *r) { memcpy(r, m->residual, sizeof(double) * 3 * m->N); }

/* ---------------------------------------------------------------------- */
/* Jacobian, reference fe/sw2/fe_sw2_load.c                                */
/* ---------------------------------------------------------------------- */

/* NUM_DIFF_EPSILON macro.h:669-673 */
static double num_diff_eps(double var, double perturbation) {
    double diff_ep = fabs(var) * (1. + perturbation);
    diff_ep -= fabs(var);
    if (diff_ep < perturbation) diff_ep = perturbation;
    return diff_ep;
}

/* element Jacobian for 2D (dim 2) or 1D (dim 1) elements: mat_u/v/h[row*nn+col]
   (reference fe_sw2_load.c:293-483 and fe_elem_utilities.c:518-525) */
static void elem_jac(o_model *m, int dim, int ie, dof3 *mu, dof3 *mv, dof3 *mh) {
    const o_desc *d = &m->d;
    int nn = dim == 2 ? 3 : 2;
    const int *nd = dim == 2 ? d->e2d + 3 * ie : d->e1d + 2 * ie;
    double perturbation = sqrt(SMALL);
    dof3 P[3], M[3];
    int i, k;
    for (i = 0; i < nn * nn; i++) {
        mu[i].x_eq = mu[i].y_eq = mu[i].c_eq = 0.;
        mv[i] = mu[i]; mh[i] = mu[i];
    }
    /* perturb_sw2_head :293-355 */
    for (i = 0; i < nn; i++) {
        double eps = num_diff_eps(m->head[nd[i]], perturbation), eps2 = 2.0 * eps;
        if (dim == 1) { boundary_resid(m, P, ie, eps, i, PERTURB_H, +1); boundary_resid(m, M, ie, eps, i, PERTURB_H, -1); }
        else { body_resid(m, P, ie, eps, i, PERTURB_H, +1); body_resid(m, M, ie, eps, i, PERTURB_H, -1); }
        for (k = 0; k < nn; k++) {
            mh[i + k * nn].x_eq = (P[k].x_eq - M[k].x_eq) / eps2;
            mh[i + k * nn].y_eq = (P[k].y_eq - M[k].y_eq) / eps2;
            mh[i + k * nn].c_eq = (P[k].c_eq - M[k].c_eq) / eps2;
        }
    }
    /* perturb_sw2_vel :377-483 */
    for (i = 0; i < nn; i++) {
        double eu = num_diff_eps(m->vel[2 * nd[i]], perturbation), eu2 = 2.0 * eu;
        double ev = num_diff_eps(m->vel[2 * nd[i] + 1], perturbation), ev2 = 2.0 * ev;
        dof3 PU[3], MU[3], PV[3], MV[3];
        if (dim == 1) {
            boundary_resid(m, PU, ie, eu, i, PERTURB_U, +1); boundary_resid(m, MU, ie, eu, i, PERTURB_U, -1);
            boundary_resid(m, PV, ie, ev, i, PERTURB_V, +1); boundary_resid(m, MV, ie, ev, i, PERTURB_V, -1);
        } else {
            body_resid(m, PU, ie, eu, i, PERTURB_U, +1); body_resid(m, MU, ie, eu, i, PERTURB_U, -1);
            body_resid(m, PV, ie, ev, i, PERTURB_V, +1); body_resid(m, MV, ie, ev, i, PERTURB_V, -1);
        }
        for (k = 0; k < nn; k++) {
            mu[i + k * nn].x_eq = (PU[k].x_eq - MU[k].x_eq) / eu2;
            mu[i + k * nn].y_eq = (PU[k].y_eq - MU[k].y_eq) / eu2;
            mu[i + k * nn].c_eq = (PU[k].c_eq - MU[k].c_eq) / eu2;
            mv[i + k * nn].x_eq = (PV[k].x_eq - MV[k].x_eq) / ev2;
            mv[i + k * nn].y_eq = (PV[k].y_eq - MV[k].y_eq) / ev2;
            mv[i + k * nn].c_eq = (PV[k].c_eq - MV[k].c_eq) / ev2;
        }
    }
}

/* reference fe/fe_assign_elem_mat_db.c:30-61 */
static void assign_mom_db(int inode, int nn, dof3 *mu, dof3 *mv, dof3 *mh) {
    int j;
    for (j = nn * inode; j < nn * inode + nn; j++) {
        mu[j].x_eq = 0.; mu[j].y_eq = 0.;
        mv[j].x_eq = 0.; mv[j].y_eq = 0.;
        mh[j].x_eq = 0.; mh[j].y_eq = 0.;
    }
    for (j = inode; j < nn * nn; j += nn) {
        mu[j].x_eq = 0.; mu[j].y_eq = 0.; mu[j].c_eq = 0.;
        mv[j].x_eq = 0.; mv[j].y_eq = 0.; mv[j].c_eq = 0.;
    }
    mu[inode * (nn + 1)].x_eq = 1.;
    mv[inode * (nn + 1)].y_eq = 1.;
}

/* reference fe/fe_assign_elem_mat_db.c:196-219 */
static void assign_dacont_db(int inode, int nn, dof3 *mu, dof3 *mv, dof3 *mh) {
    int j;
    for (j = nn * inode; j < nn * inode + nn; j++) {
        mu[j].c_eq = 0.; mv[j].c_eq = 0.; mh[j].c_eq = 0.;
    }
    for (j = inode; j < nn * nn; j += nn) {
        mh[j].x_eq = 0.; mh[j].y_eq = 0.; mh[j].c_eq = 0.;
    }
    mh[inode * (nn + 1)].c_eq = 1.;
}

/* element Jacobian with the Dirichlet edits of fe_sw2_load.c:100-120 */
static void elem_jac_db(o_model *m, int dim, int ie, dof3 *mu, dof3 *mv, dof3 *mh) {
    const o_desc *d = &m->d;
    int nn = dim == 2 ? 3 : 2, i;
    const int *nd = dim == 2 ? d->e2d + 3 * ie : d->e1d + 2 * ie;
    elem_jac(m, dim, ie, mu, mv, mh);
    for (i = 0; i < nn; i++) {
        /* node_string = grid->node[i].string: the LOCAL index i, as the reference reads it (:103, :165) */
        int node_bc = d->node_bc ? d->node_bc[i] : 0;
        if (m->old_head[nd[i]] <= 0.) assign_mom_db(i, nn, mu, mv, mh);
        if (node_bc == O_BCT_PRS_DIR) assign_dacont_db(i, nn, mu, mv, mh);
        else if (node_bc == O_BCT_VEL_DIR) assign_mom_db(i, nn, mu, mv, mh);
        else if (node_bc == O_BCT_VEL_PRS_DIR) { assign_dacont_db(i, nn, mu, mv, mh); assign_mom_db(i, nn, mu, mv, mh); }
    }
}

void o_elem_jacobian(o_model *m, int ie, double *out81) {
    dof3 mu[9], mv[9], mh[9];
    int i, j;
    elem_jac_db(m, 2, ie, mu, mv, mh);
    /* out[(i*3+j)*9 + k] in the reference's block value order (fe_assemble_matrix.c:200-208) */
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) {
            int idx = 3 * i + j;
            double *b = out81 + 9 * idx;
            b[0] = mu[idx].x_eq; b[1] = mv[idx].x_eq; b[2] = mh[idx].x_eq;
            b[3] = mu[idx].y_eq; b[4] = mv[idx].y_eq; b[5] = mh[idx].y_eq;
            b[6] = mu[idx].c_eq; b[7] = mv[idx].c_eq; b[8] = mh[idx].c_eq;
        }
}

/* spv_load, reference solver/spv_load.c:82-112 (p == 3), growth by SPV_BLOCK (define.h:457) */
static void spv_add(spv *sv, int ind, const double *value) {
    int i = 0, k;
    while (i < sv->size && sv->index[i] != ind) i++;
    if (i == sv->size) {
        if (sv->size == sv->cap) {
            sv->cap += 75;
            sv->index = realloc(sv->index, sizeof(int) * sv->cap);
            sv->value = realloc(sv->value, sizeof(double) * 9 * sv->cap);
        }
        sv->size++;
        sv->index[i] = ind;
        for (k = 0; k < 9; k++) sv->value[9 * i + k] = 0.0;
    }
    for (k = 0; k < 9; k++) sv->value[9 * i + k] += value[k];
}

/* reference fe/fe_assemble_matrix.c:166-224 */
static void assemble(o_model *m, int nn, const int *nodes, const dof3 *mu, const dof3 *mv, const dof3 *mh) {
    int i, j, k;
    double values[9];
    for (i = 0; i < nn; i++) {
        int ifmap = nodes[i];
        for (j = 0; j < nn; j++) {
            int jfmap = nodes[j];
            int index = nn * i + j;
            values[0] = mu[index].x_eq; values[1] = mv[index].x_eq; values[2] = mh[index].x_eq;
            values[3] = mu[index].y_eq; values[4] = mv[index].y_eq; values[5] = mh[index].y_eq;
            values[6] = mu[index].c_eq; values[7] = mv[index].c_eq; values[8] = mh[index].c_eq;
            if (jfmap != ifmap) spv_add(m->matrix + ifmap, jfmap, values);
            else for (k = 0; k < 9; k++) m->diagonal[9 * ifmap + k] += values[k];
        }
    }
}

/* reference fe/sw2/fe_sw2_load.c:34-269 */
void o_load(o_model *m, double *diag_out, double *offd_out) {
    const o_desc *d = &m->d;
    int ie, i, j;
    dof3 mu[9], mv[9], mh[9];
    /* init_adh_matrix tools/tl_matrix.c:28-39 */
    memset(m->diagonal, 0, sizeof(double) * 9 * m->N);
    for (i = 0; i < m->N; i++) memset(m->matrix[i].value, 0, sizeof(double) * 9 * m->matrix[i].size);
    for (ie = 0; ie < m->E; ie++) {
        elem_jac_db(m, 2, ie, mu, mv, mh);
        assemble(m, 3, d->e2d + 3 * ie, mu, mv, mh);
    }
    for (ie = 0; ie < m->B; ie++) {
        elem_jac_db(m, 1, ie, mu, mv, mh);
        assemble(m, 2, d->e1d + 2 * ie, mu, mv, mh);
    }
    m->pattern_built = 1;
    /* :218-236 diagonal floor */
    for (i = 0; i < m->N; i++) {
        j = i * 9;
        if (fabs(m->diagonal[j]) < NOT_QUITE_SMALL) {
            m->diagonal[j] = NOT_QUITE_SMALL;
            if (fabs(m->residual[3 * i]) < NOT_QUITE_SMALL) m->residual[3 * i] = -NOT_QUITE_SMALL * m->vel[2 * i];
        }
        j = j + 4;
        if (fabs(m->diagonal[j]) < NOT_QUITE_SMALL) {
            m->diagonal[j] = NOT_QUITE_SMALL;
            if (fabs(m->residual[3 * i + 1]) < NOT_QUITE_SMALL) m->residual[3 * i + 1] = -NOT_QUITE_SMALL * m->vel[2 * i + 1];
        }
        j = j + 4;
        if (fabs(m->diagonal[j]) < NOT_QUITE_SMALL) m->diagonal[j] = NOT_QUITE_SMALL;
    }
    o_get_matrix(m, diag_out, offd_out);
}

int o_pattern_nnz(const o_model *m) {
    int i, n = 0;
    for (i = 0; i < m->N; i++) n += m->matrix[i].size;
    return n;
}

void o_get_pattern(const o_model *m, int *row_ptr, int *col) {
    int i, k, p = 0;
    row_ptr[0] = 0;
    for (i = 0; i < m->N; i++) {
        for (k = 0; k < m->matrix[i].size; k++) col[p++] = m->matrix[i].index[k];
        row_ptr[i + 1] = p;
    }
}

void o_get_matrix(const o_model *m, double *diag, double *offd) {
    int i, p = 0;
    if (diag) memcpy(diag, m->diagonal, sizeof(double) * 9 * m->N);
    if (offd)
        for (i = 0; i < m->N; i++) {
            memcpy(offd + 9 * (size_t)p, m->matrix[i].value, sizeof(double) * 9 * m->matrix[i].size);
            p += m->matrix[i].size;
        }
}

void o_set_matrix(o_model *m, const double *diag, const double *offd, const double *rhs) {
    int i, p = 0;
    memcpy(m->diagonal, diag, sizeof(double) * 9 * m->N);
    for (i = 0; i < m->N; i++) {
        memcpy(m->matrix[i].value, offd + 9 * (size_t)p, sizeof(double) * 9 * m->matrix[i].size);
        p += m->matrix[i].size;
    }
    if (rhs) memcpy(m->residual, rhs, sizeof(double) * 3 * m->N);
}

void o_get_rhs(const o_model *m, double *b) { memcpy(b, m->residual, sizeof(double) * 3 * m->N); }

/* ---------------------------------------------------------------------- */
/* linear algebra, reference src/solver                                    */
/* ---------------------------------------------------------------------- */

/* spv_jacobi 3-DOF branch, reference solver/spv_jacobi.c:222-316 */
static void jacobi_scale(o_model *m) {
    int i, j, k, n3 = 3 * m->N;
    double small = SOLV_TOL, small_factor = 1.0 / sqrt(small);
    double *sv = m->scale_vect, *D = m->diagonal;
    for (i = 0; i < m->N; i++) {
        for (j = 0; j < 3; j++) {
            double s = 0.0;
            for (k = 0; k < 3; k++) if (fabs(D[9 * i + 3 * j + k]) > s) s = fabs(D[9 * i + 3 * j + k]);
            sv[3 * i + j] = s;
        }
    }
    for (i = 0; i < m->N; i++) {
        spv *r = m->matrix + i;
        double dn[3] = {0.0, 0.0, 0.0};
        int b;
        /* spv_norm solver/spv_norm.c:109-160: column-by-column within each block */
        for (b = 0; b < r->size; b++)
            for (k = 0; k < 3; k++)
                for (j = 0; j < 3; j++) {
                    double t = fabs(r->value[9 * b + 3 * j + k]);
                    if (t > dn[j]) dn[j] = t;
                }
        for (j = 0; j < 3; j++) if (dn[j] > sv[3 * i + j]) sv[3 * i + j] = dn[j];
    }
    for (i = 0; i < n3; i++) {
        if (sv[i] < small) sv[i] = small_factor;
        else sv[i] = 1.0 / sqrt(fabs(sv[i]));
    }
    for (i = 0; i < n3; i++) {
        double factor = sv[i];
        m->residual[i] *= factor;
        m->sol[i] /= factor;
    }
    for (i = 0; i < m->N; i++) {
        double rf[3] = {sv[3 * i], sv[3 * i + 1], sv[3 * i + 2]};
        for (j = 0; j < 3; j++)
            for (k = 0; k < 3; k++) D[9 * i + 3 * j + k] *= rf[j] * rf[k];
        spv *r = m->matrix + i;
        int b;
        for (b = 0; b < r->size; b++) { /* spv_vscale solver/spv_vscale.c:66-90 */
            int c = 3 * r->index[b];
            for (j = 0; j < 3; j++)
                for (k = 0; k < 3; k++) r->value[9 * b + 3 * j + k] *= rf[j] * sv[c + k];
        }
    }
}

/* bv_load, reference solver/bv_load.c:6-57 (final band content only) */
static void band_put(bandv *bv, double entry, int index) {
    if (index < bv->begin) {
        int new_len = bv->end - index, i;
        if (new_len > bv->size) {
            bv->value = realloc(bv->value, sizeof(double) * (new_len + 10));
            for (i = bv->size; i < new_len + 10; i++) bv->value[i] = 0.0;
            bv->size = new_len + 10;
        }
        int shift = bv->begin - index;
        for (i = bv->end - bv->begin - 1; i >= 0; i--) bv->value[i + shift] = bv->value[i];
        for (i = 0; i < shift; i++) bv->value[i] = 0.0;
        bv->begin = index;
    }
    if (index >= bv->end) {
        int new_len = index - bv->begin + 1, i;
        if (new_len > bv->size) {
            bv->value = realloc(bv->value, sizeof(double) * (new_len + 10));
            for (i = bv->size; i < new_len + 10; i++) bv->value[i] = 0.0;
            bv->size = new_len + 10;
        }
        for (i = bv->end - bv->begin; i < new_len; i++) bv->value[i] = 0.0;
        bv->end = index + 1;
    }
    bv->value[index - bv->begin] = entry;
}

/* bv_dot, reference solver/bv_dot.c:6-40 */
static double band_dot(const bandv *v1, const bandv *v2, int ib, int ie) {
    int i, isum;
    double value;
    if (ib < v1->begin) ib = v1->begin;
    if (ib < v2->begin) ib = v2->begin;
    if (ie > v1->end) ie = v1->end;
    if (ie > v2->end) ie = v2->end;
    const double *a1 = v1->value + (ib - v1->begin), *a2 = v2->value + (ib - v2->begin);
    isum = ie - ib;
    for (i = 0, value = 0.0; i < isum; i++) value += a1[i] * a2[i];
    return value;
}

/* node blocks, reference node/node_order.c:64-104 (serial) */
static void set_node_blocks(o_model *m, int nblock) {
    int iblk, ilast = 0, i, inode = 0, blk_inc = m->N / nblock;
    if (blk_inc < 1) blk_inc = 1;
    for (iblk = 0; iblk < nblock; iblk++) {
        int ifirst = ilast;
        if (iblk == nblock - 1) ilast = m->N;
        else ilast += blk_inc;
        for (i = ifirst; i < ilast; i++, inode++) m->node_block[inode] = iblk;
    }
    m->prof_nblock = nblock;
}

/* solv_blk_set_profile / solv_blk_load_profile / solv_blk_factor_profile,
   reference solver/solv_blk_profile.c:5-79, solv_blk_load_profile.c:16-233 */
static void profile_factor(o_model *m) {
    int neq = 3 * m->N, i, j, r, c;
    if (m->prof_size < neq) {
        m->rows = realloc(m->rows, sizeof(bandv) * neq);
        m->cols = realloc(m->cols, sizeof(bandv) * neq);
        for (i = m->prof_size; i < neq; i++) {
            m->rows[i].value = NULL; m->rows[i].size = 0;
            m->cols[i].value = NULL; m->cols[i].size = 0;
        }
        m->prof_size = neq;
    }
    for (i = 0; i < neq; i++) {
        m->rows[i].begin = i; m->rows[i].end = i;
        m->cols[i].begin = i; m->cols[i].end = i;
    }
    for (i = 0; i < m->N; i++) {
        int my_block = m->node_block[i];
        spv *row = m->matrix + i;
        for (j = 0; j < row->size; j++) {
            int other = row->index[j];
            if (m->node_block[other] != my_block) continue;
            const double *v = row->value + 9 * j;
            if (other < i) {
                for (r = 0; r < 3; r++)
                    for (c = 0; c < 3; c++) band_put(m->rows + 3 * i + r, v[3 * r + c], 3 * other + c);
            } else {
                for (r = 0; r < 3; r++)
                    for (c = 0; c < 3; c++) band_put(m->cols + 3 * other + c, v[3 * r + c], 3 * i + r);
            }
        }
        const double *D = m->diagonal + 9 * i;
        int s = 3 * i;
        band_put(m->rows + s + 1, D[3], s + 0);
        band_put(m->rows + s + 2, D[6], s + 0);
        band_put(m->rows + s + 2, D[7], s + 1);
        band_put(m->cols + s + 0, D[0], s + 0);
        band_put(m->cols + s + 1, D[1], s + 0);
        band_put(m->cols + s + 2, D[2], s + 0);
        band_put(m->cols + s + 1, D[4], s + 1);
        band_put(m->cols + s + 2, D[5], s + 1);
        band_put(m->cols + s + 2, D[8], s + 2);
    }
    bandv *rows = m->rows, *cols = m->cols;
    int ip, k;
    for (ip = 0; ip < neq; ip++) {
        for (j = rows[ip].begin, k = 0; j < rows[ip].end; j++, k++) {
            rows[ip].value[k] -= band_dot(rows + ip, cols + j, 0, j);
            double pivot = cols[j].value[cols[j].end - cols[j].begin - 1];
            if (fabs(pivot) > SMALL) rows[ip].value[k] /= pivot;
            else rows[ip].value[k] /= SMALL;
        }
        for (i = cols[ip].begin, k = 0; i < cols[ip].end; i++, k++) cols[ip].value[k] -= band_dot(rows + i, cols + ip, 0, i);
    }
}

/* solv_blk_solve, reference solver/solv_blk_solve.c:5-62 */
static void profile_solve(o_model *m, double *sol) {
    int neq = 3 * m->N, k, i, j, iend;
    bandv *rows = m->rows, *cols = m->cols;
    for (k = 0; k < neq; k++) {
        iend = (k > rows[k].end) ? rows[k].end : k;
        for (i = rows[k].begin, j = 0; i < iend; i++, j++) sol[k] -= rows[k].value[j] * sol[i];
    }
    for (k = neq - 1; k > -1; k--) {
        sol[k] /= cols[k].value[cols[k].end - cols[k].begin - 1];
        iend = (k > cols[k].end) ? cols[k].end : k;
        double scale;
        for (i = cols[k].begin, scale = sol[k], j = 0; i < iend; i++, j++) sol[i] -= cols[k].value[j] * scale;
    }
}

/* 3x3 block-Jacobi: inverse of each diagonal block by Gauss-Jordan with
   partial pivoting.  This is the HIP path's preconditioner, restated here
   for Krylov-iterate comparisons; it has no counterpart in the reference. */
static void bj_setup(o_model *m) {
    int n, r, c, k;
    for (n = 0; n < m->N; n++) {
        double a[3][6];
        for (r = 0; r < 3; r++) {
            for (c = 0; c < 3; c++) { a[r][c] = m->diagonal[9 * n + 3 * r + c]; a[r][3 + c] = (r == c) ? 1.0 : 0.0; }
        }
        for (k = 0; k < 3; k++) {
            int p = k;
            for (r = k + 1; r < 3; r++) if (fabs(a[r][k]) > fabs(a[p][k])) p = r;
            if (p != k) for (c = 0; c < 6; c++) { double t = a[k][c]; a[k][c] = a[p][c]; a[p][c] = t; }
            double piv = a[k][k];
            if (fabs(piv) < SMALL) piv = (piv < 0) ? -SMALL : SMALL;
            double ip = 1.0 / piv;
            for (c = 0; c < 6; c++) a[k][c] *= ip;
            for (r = 0; r < 3; r++) {
                if (r == k) continue;
                double f = a[r][k];
                for (c = 0; c < 6; c++) a[r][c] -= f * a[k][c];
            }
        }
        for (r = 0; r < 3; r++) for (c = 0; c < 3; c++) m->bjinv[9 * n + 3 * r + c] = a[r][3 + c];
    }
}

static void bj_apply(const o_model *m, const double *in, double *out) {
    int n;
    for (n = 0; n < m->N; n++) {
        const double *A = m->bjinv + 9 * n;
        const double *x = in + 3 * n;
        out[3 * n] = A[0] * x[0] + A[1] * x[1] + A[2] * x[2];
        out[3 * n + 1] = A[3] * x[0] + A[4] * x[1] + A[5] * x[2];
        out[3 * n + 2] = A[6] * x[0] + A[7] * x[1] + A[8] * x[2];
    }
}

/* solv_amult 3-DOF branch, reference solver/solv_amult.c:92-129 + spv_dot.c:59-83 */
static void amult(const o_model *m, const double *v, double *av) {
    int i, j;
    for (i = 0; i < m->N; i++) {
        const double *D = m->diagonal + 9 * i;
        const double *x = v + 3 * i;
        av[3 * i] = D[0] * x[0] + D[1] * x[1] + D[2] * x[2];
        av[3 * i + 1] = D[3] * x[0] + D[4] * x[1] + D[5] * x[2];
        av[3 * i + 2] = D[6] * x[0] + D[7] * x[1] + D[8] * x[2];
    }
    for (i = 0; i < m->N; i++) {
        const spv *r = m->matrix + i;
        double p0 = 0.0, p1 = 0.0, p2 = 0.0;
        for (j = 0; j < r->size; j++) {
            const double *a = r->value + 9 * j;
            const double *x = v + 3 * r->index[j];
            p0 += a[0] * x[0] + a[1] * x[1] + a[2] * x[2];
            p1 += a[3] * x[0] + a[4] * x[1] + a[5] * x[2];
            p2 += a[6] * x[0] + a[7] * x[1] + a[8] * x[2];
        }
        av[3 * i] += p0;
        av[3 * i + 1] += p1;
        av[3 * i + 2] += p2;
    }
}

static void prec_apply(o_model *m, int prec, const double *in, double *out) {
    int n3 = 3 * m->N;
    if (prec == 2) { bj_apply(m, in, out); return; }
    memcpy(out, in, sizeof(double) * n3);
    if (prec == 1) profile_solve(m, out);
}

/* math_blas_wrappers.c:44-106 (non-BLAS loops) and math_simple.c:361 */
static double vdot(int n, const double *x, const double *y) {
    double v = 0.0;
    int i;
    for (i = 0; i < n; i++) v += x[i] * y[i];
    return v;
}
static void vaxpy(int n, double a, const double *x, double *y) {
    int i;
    for (i = 0; i < n; i++) y[i] += a * x[i];
}
static double vinf(int n, const double *x) {
    double v = 0.0;
    int i;
    for (i = 0; i < n; i++) if (fabs(x[i]) > v) v = fabs(x[i]);
    return v;
}

/* solv_bcgstab, reference solver/solv_bcgstab.c:39-301 (rho update at :246 kept) */
static int bcgstab(o_model *m, const o_newton_opts *o, int *lin_it) {
    int n = 3 * m->N, it = 0, i;
    double *r = xcalloc(n, 8), *p = xcalloc(n, 8), *Ap = xcalloc(n, 8), *Mp = xcalloc(n, 8), *q = xcalloc(n, 8);
    double *s = xcalloc(n, 8), *As = xcalloc(n, 8), *Ms = xcalloc(n, 8), *u0 = xcalloc(n, 8);
    double *b = m->residual, *u = m->sol;
    double alpha, omega, beta, gamma, rho, rhop, rnorm, bnorm;
    memcpy(u0, u, 8 * (size_t)n);
    memset(u, 0, 8 * (size_t)n);
    memcpy(p, b, 8 * (size_t)n);
    memset(b, 0, 8 * (size_t)n);
    prec_apply(m, o->prec, p, b);            /* b = M^-1 b */
    memset(p, 0, 8 * (size_t)n);
    memcpy(u, b, 8 * (size_t)n);             /* x0 = M^-1 b */
    amult(m, u, Mp);
    prec_apply(m, o->prec, Mp, Ap);
    memcpy(r, b, 8 * (size_t)n);
    vaxpy(n, -1.0, Ap, r);
    memset(Mp, 0, 8 * (size_t)n);
    memset(Ap, 0, 8 * (size_t)n);
    memcpy(q, r, 8 * (size_t)n);
    rnorm = vinf(n, r);
    bnorm = vinf(n, b);
    rhop = 1.0; alpha = 1.0; omega = 1.0;
    rho = vdot(n, r, q);
    while (rnorm > (o->lin_atol + bnorm * o->lin_rtol) && it < o->max_lin_it) {
        it++;
        beta = (rho * alpha) / (rhop * omega);
        vaxpy(n, -omega, Ap, p);
        for (i = 0; i < n; i++) p[i] *= beta;
        vaxpy(n, 1.0, r, p);
        amult(m, p, Mp);
        prec_apply(m, o->prec, Mp, Ap);
        gamma = vdot(n, q, Ap);
        alpha = rho / gamma;
        memcpy(s, r, 8 * (size_t)n);
        vaxpy(n, -alpha, Ap, s);
        amult(m, s, Ms);
        prec_apply(m, o->prec, Ms, As);
        omega = vdot(n, As, s);
        gamma = vdot(n, As, As);
        omega = omega / gamma;
        rhop = rho;
        rho = -1.0 * omega * vdot(n, As, q);
        vaxpy(n, alpha, p, u);
        vaxpy(n, omega, s, u);
        memcpy(r, s, 8 * (size_t)n);
        vaxpy(n, -omega, As, r);
        rnorm = vinf(n, r);
    }
    int ok = (it == o->max_lin_it) ? -3 : 1;
    /* solv_return_old solver/solv_return_old.c:23-30 */
    for (i = 0; i < n; i++) u[i] += u0[i];
    for (i = 0; i < n; i++) u[i] *= m->scale_vect[i];
    free(r); free(p); free(Ap); free(Mp); free(q); free(s); free(As); free(Ms); free(u0);
    if (lin_it) *lin_it = it;
    return ok;
}

/* solv_linear_sys_setup prec 1 (solv_linear_sys_setup.c:44-64) */
void o_linear_setup(o_model *m, const o_newton_opts *o, double *scale_out) {
    jacobi_scale(m);
    if (o->prec == 1) {
        if (m->prof_nblock != o->nblock) set_node_blocks(m, o->nblock);
        profile_factor(m);
    } else if (o->prec == 2) {
        bj_setup(m);
    }
    if (scale_out) memcpy(scale_out, m->scale_vect, sizeof(double) * 3 * m->N);
}

int o_linear_solve(o_model *m, const o_newton_opts *o, double *sol_out, int *lin_it) {
    int ok = bcgstab(m, o, lin_it);
    if (sol_out) memcpy(sol_out, m->sol, sizeof(double) * 3 * m->N);
    return ok;
}

int o_bcgstab_only(o_model *m, const o_newton_opts *o, double *x_out, int *lin_it) {
    int i;
    for (i = 0; i < 3 * m->N; i++) { m->sol[i] = 0.0; m->scale_vect[i] = 1.0; }
    if (o->prec == 1) { if (m->prof_nblock != o->nblock) set_node_blocks(m, o->nblock); profile_factor(m); }
    else if (o->prec == 2) bj_setup(m);
    int ok = bcgstab(m, o, lin_it);
    if (x_out) memcpy(x_out, m->sol, sizeof(double) * 3 * m->N);
    return ok;
}

/* reference fe/sw2/fe_sw2_inc.c:47-54 */
void o_inc(o_model *m, const double *sol) {
    int i;
    const double *s = sol ? sol : m->sol;
    for (i = 0; i < m->N; i++) {
        m->vel[2 * i] += 1. * s[3 * i];
        m->vel[2 * i + 1] += 1. * s[3 * i + 1];
        m->head[i] += 1. * s[3 * i + 2];
    }
}

/* get_residual_norms fe_newton.c:40-151 with include_node_in_residual :177-187 */
static void norms(o_model *m, double *maxn, double *l2, double *incn) {
    int i, j;
    double pl2 = 0., pmax = 0., pinc = 0.;
    for (i = 0; i < m->N; i++) {
        int inc = MIN(m->head[i], m->old_head[i]) > 0;
        for (j = 0; j < 3; j++) {
            int idof = 3 * i + j;
            if (fabs(m->sol[idof]) > pinc && inc) pinc = fabs(m->sol[idof]);
            if (fabs(m->residual[idof]) > pmax && inc) pmax = fabs(m->residual[idof]);
            pl2 += m->residual[idof] * m->residual[idof];
        }
    }
    *maxn = pmax;
    *l2 = sqrt(pl2 / (m->N * 3));
    *incn = pinc;
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* fe_newton, reference fe/fe_newton.c:214-912 (serial, non-PETSc, SW2) */
int o_newton(o_model *m, const o_newton_opts *o, o_newton_report *rep) {
    int it = 0, i, n3 = 3 * m->N, solv_flag = 1, keep;
    double maxn, l2, incn, old_l2, t0;
    if (rep) { rep->nhist = 0; rep->lin_iterations_total = 0; rep->time_load = rep->time_setup = rep->time_solve = rep->time_resid = 0.; }
    o_init(m, NULL);
    t0 = now_s();
    o_resid(m, NULL);
    if (rep) rep->time_resid += now_s() - t0;
    norms(m, &maxn, &l2, &incn);
    if (rep && rep->nhist < o->hist_cap) {
        rep->hist_max[rep->nhist] = maxn; rep->hist_l2[rep->nhist] = l2; rep->hist_inc[rep->nhist] = incn;
        rep->hist_lin_it[rep->nhist] = 0; rep->nhist++;
    }
    do {
        int lin_it = 0;
        t0 = now_s();
        o_load(m, NULL, NULL);
        if (rep) rep->time_load += now_s() - t0;
        for (i = 0; i < n3; i++) m->sol[i] = 0.0;
        t0 = now_s();
        o_linear_setup(m, o, NULL);
        if (rep) rep->time_setup += now_s() - t0;
        t0 = now_s();
        solv_flag = bcgstab(m, o, &lin_it);
        if (rep) { rep->time_solve += now_s() - t0; rep->lin_iterations_total += lin_it; }
        o_inc(m, NULL);
        for (i = 0; i < n3; i++) m->sol[i] = -m->sol[i];
        old_l2 = l2;
        t0 = now_s();
        o_resid(m, NULL);
        if (rep) rep->time_resid += now_s() - t0;
        norms(m, &maxn, &l2, &incn);
        if (isnan(l2) || isinf(l2)) solv_flag = -3;
        int cuts = 0;
        while (maxn > o->tol_nonlin && l2 > old_l2 && cuts < o->max_cuts) {
            for (i = 0; i < n3; i++) m->sol[i] *= 0.5;
            o_inc(m, NULL);
            o_resid(m, NULL);
            norms(m, &maxn, &l2, &incn);
            cuts++;
        }
        it++;
        if (rep && rep->nhist < o->hist_cap) {
            rep->hist_max[rep->nhist] = maxn; rep->hist_l2[rep->nhist] = l2; rep->hist_inc[rep->nhist] = incn;
            rep->hist_lin_it[rep->nhist] = lin_it; rep->nhist++;
        }
        keep = (solv_flag == 1) && (it < o->max_nonlin_it) && (maxn < 1.0E+20) && (maxn > o->tol_nonlin && incn > o->inc_nonlin);
    } while (keep);
    int fail = (solv_flag != 1) || ((maxn > o->tol_nonlin) && (incn > o->inc_nonlin)) || (maxn >= 1.0E+20);
    if (rep) { rep->iterations = it; rep->converged = !fail; rep->solv_fail = solv_flag != 1; }
    if (fail) {
        o_init(m, NULL); /* state reset on failure (fe_newton.c:883); dt cut is the caller's */
        return -3;
    }
    return 1;
}

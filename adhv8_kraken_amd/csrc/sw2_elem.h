// SW2 element residuals (2D triangle body + 1D boundary edge), device side.
//
// These are the per-element computations of the reference's
// fe_sw2_body_resid (src/fe/sw2/fe_sw2_body_resid.c:45-1267) and
// fe_sw2_boundary_resid (src/fe/sw2/fe_sw2_boundary_resid.c:72-905), with
// the wet/dry clipping of fe_sw2_wet_dry_wrapper (src/fe/fe_sw2_wet_dry_wrapper.c)
// and the P1 integrals of src/fe/fe_elem2d_integrations.c / fe_elem1d_integrations.c.
//
// Every floating-point expression keeps the reference's operation order and
// the file is compiled with -ffp-contract=off, so a residual computed here is
// bit-identical to the reference's on the same doubles (log is correctly
// rounded, see cr_log.h).  Inputs are read from structure-of-arrays node
// fields with coalesced element-index loads; nothing is written to memory
// here -- the kernels in sw2_kernels.hip own all stores.
//
// The functions are __host__ __device__ so the test suite can also run them
// on the CPU against the oracle; the library itself only calls them on the GPU.
#pragma once

#include "cr_log.h"

#ifdef __HIPCC__
#define HD __host__ __device__ inline
#define HDF __host__ __device__ __forceinline__
#define HDM __host__ __device__ __forceinline__
#else
#define HD static inline
#define HDF static inline
#define HDM inline
#endif

namespace sw2 {

// reference src/include/define.h
constexpr double SMALL = 2.220446049250313080847e-16;        // DBL_EPSILON
constexpr double NOT_QUITE_SMALL = 1.1920928955078125e-07;   // FLT_EPSILON
constexpr double one_2 = 0.5e+0;
constexpr double one_3 = 0.333333333333333333333333;
constexpr double one_6 = 0.166666666666666666666667;
constexpr double one_12 = 0.083333333333333333333333;
constexpr double one_30 = 0.033333333333333333333333;
constexpr double one_60 = 0.016666666666666666666667;

constexpr int PERTURB_NONE = -1, PERTURB_U = 1, PERTURB_V = 2, PERTURB_H = 5;
constexpr int BCT_OUTFLOW = 4, BCT_PRS_DIR = 7, BCT_VEL_DIR = 8, BCT_PRS_NEU = 9, BCT_VEL_NEU = 10,
              BCT_VEL_PRS_DIR = 16;
constexpr int TURB_EVS = 1, TURB_EEV = 2;

// the reference's unparenthesised MIN/MAX (src/include/macro.h:844-846): a > b ? a : b
HDF double dmax(double a, double b) { return a > b ? a : b; }
HDF double dmin(double a, double b) { return a < b ? a : b; }

struct d3 { double x_eq, y_eq, c_eq; };
struct v2 { double x, y; };
struct v3 { double x, y, z; };

// Read-only model data for the element kernels (all device pointers).
struct ElemData {
    // nodes (SoA)
    const double *x, *y, *z;
    const double *head, *u, *v;
    const double *old_head, *old_u, *old_v;
    const double *older_head, *older_u, *older_v;
    const int *node_bc;        // bc flag of each node's node string (0 = none)
    const int *first_elem;     // lowest 2D element touching each node (dry-node velocity rule), or null
    // triangles
    const int *e2d;            // [3E]
    const double *djac;        // [E]
    const double *grad;        // [6E] gx0 gy0 gx1 gy1 gx2 gy2
    const int *mat;
    const double *ks;
    const int *src_flag;
    const double *src;
    const int *wd_flag;
    const double *pre;         // [PRE_STRIDE * E] old/older-state terms (elem_old_terms)
    // boundary edges
    const int *e1d;            // [2B]
    const double *djac1d;      // [B]
    const double *nrml1d;      // [2B]
    const int *bc1d;
    const double *val1d;
    // materials
    const int *mat_turb, *mat_eev_mode;
    const double *mat_eev_coef, *mat_ev, *mat_cor;
    // constants
    double g, dt, alpha, tau_pg, dll, dul, visc;
    const crlog::tables *logt;
    // ensembles of independent models in one mesh: per-model time step,
    // looked up through the element's (edge's) model; null for one model
    const double *dt_inst;
    const int *elem_inst, *edge_inst;
    const int *loc_bc_inst;    // [3 * models] node-string flags of each model's nodes 0..2
    // winds: node wind values (x, y per node; null = no wind), attenuation per material, density
    const double *wind;
    const double *mat_windatt;
    double density;
};

HDF double elem_dt(const ElemData &D, int ie) { return D.dt_inst ? D.dt_inst[D.elem_inst[ie]] : D.dt; }
HDF double edge_dt(const ElemData &D, int ie) { return D.dt_inst ? D.dt_inst[D.edge_inst[ie]] : D.dt; }

// Correctly rounded x / b from y = RN(1/b): one multiply and two FMAs
// instead of a full division.  Markstein's correction step (P. Markstein,
// "Computation of elementary functions on the IBM RISC System/6000
// processor", IBM J. Res. Develop. 34(1), 1990, 111-119; J.-M. Muller et al.,
// Handbook of Floating-Point Arithmetic, 2nd ed., 2018, sec. 4.9): with
// q = RN(x y) and the exact remainder r = x - q b (one FMA), RN(q + r y) is
// the correctly rounded quotient for normal quotients; a zero x returns
// x * y, a zero of x's sign.  Valid where quotients stay normal (no subnormal
// or overflowing x / b), which holds at the call sites: sums of residual
// terms and FD differences over 2 eps >= 3e-8.  The host tests compare every
// use against the oracle's divisions.
HDF double div_rn(double x, double b, double y) {
    const double q = x * y;
    const double r = fma(-q, b, x);
    return x == 0.0 ? q : fma(r, y, q);
}
constexpr double RN_ONE_THIRD = 1.0 / 3.0;   // RN(1/3)

// ---------------------------------------------------------------- integrals
// fe_elem2d_integrations.c:60-62 (c * djac * sum / 3.0, the division exact-rounded)
HDF double int_tri_f(double djac, double c, const double *f) {
    return div_rn(c * djac * (f[0] + f[1] + f[2]), 3.0, RN_ONE_THIRD);
}
// :119-121
HDF double int_tri_f_f(double djac, double c, const double *f) {
    return (c * one_6 * djac * (f[0] * f[0] + f[0] * f[1] + f[1] * f[1] + f[0] * f[2] + f[1] * f[2] + f[2] * f[2]));
}
// :196-202
HDF void int_tri_phi_f(double djac, double c, const double *f, double *I) {
    double f_sum = f[0] + f[1] + f[2];
    double t1 = djac * c * one_12;
    I[0] += t1 * (f_sum + f[0]);
    I[1] += t1 * (f_sum + f[1]);
    I[2] += t1 * (f_sum + f[2]);
}
// :250-255
HDF void int_tri_phi_f_lump(double djac, double c, const double *f, double *I) {
    double t1 = djac * c * one_3;
    I[0] += t1 * f[0];
    I[1] += t1 * f[1];
    I[2] += t1 * f[2];
}
// :281-287
HDF void int_tri_phi_f_g(double djac, double c, const double *f, const double *g, double *I) {
    double g_sum = 2 * (g[0] + g[1] + g[2]);
    double t1 = c * djac * one_60;
    I[0] += t1 * (f[0] * (g_sum + 4 * g[0]) + f[1] * (g_sum - g[2]) + f[2] * (g_sum - g[1]));
    I[1] += t1 * (f[1] * (g_sum + 4 * g[1]) + f[2] * (g_sum - g[0]) + f[0] * (g_sum - g[2]));
    I[2] += t1 * (f[2] * (g_sum + 4 * g[2]) + f[1] * (g_sum - g[0]) + f[0] * (g_sum - g[1]));
}
// :422-427
HDF void int_tri_gradPhi_dot_vbar(const v2 *gs, double djac, double c, v2 vb, double *I) {
    double t1 = c * djac;
    I[0] += t1 * (vb.x * gs[0].x + vb.y * gs[0].y);
    I[1] += t1 * (vb.x * gs[1].x + vb.y * gs[1].y);
    I[2] += t1 * (vb.x * gs[2].x + vb.y * gs[2].y);
}
// fe_elem1d_integrations.c:31-35
HDF void int_line_phi_h_v_dot_n(double djac, double c, const double *h, const v2 *v, v2 n, double *I) {
    double t1 = djac * one_12 * c;
    I[0] += t1 * (n.x * ((3 * h[0] + h[1]) * v[0].x + (h[0] + h[1]) * v[1].x) + n.y * ((3 * h[0] + h[1]) * v[0].y + (h[0] + h[1]) * v[1].y));
    I[1] += t1 * (n.x * ((h[0] + h[1]) * v[0].x + (h[0] + 3 * h[1]) * v[1].x) + n.y * ((h[0] + h[1]) * v[0].y + (h[0] + 3 * h[1]) * v[1].y));
}
// :45-49
HDF void int_line_phi_h_h(double djac, double c, double h1, double h2, double *I) {
    double t1 = djac * one_12 * c;
    I[0] += t1 * (3 * h1 * h1 + 2 * h1 * h2 + h2 * h2);
    I[1] += t1 * (h1 * h1 + 2 * h1 * h2 + 3 * h2 * h2);
}
// :73-77
HDF void int_line_phi(double djac, double c, double *I) {
    double t1 = one_2 * djac * c;
    I[0] += t1;
    I[1] += t1;
}
// :87-91
HDF void int_line_phi_f(double djac, double c, const double *f, double *I) {
    double t1 = one_6 * djac * c;
    I[0] += t1 * (2 * f[0] + f[1]);
    I[1] += t1 * (f[0] + 2 * f[1]);
}
// :115-119
HDF void int_line_phi_f_g_h(double djac, double c, double f1, double f2, double g1, double g2, double h1, double h2,
                            double *I) {
    double t1 = one_60 * djac * c;
    I[0] += t1 * ((3 * (4 * f1 + f2) * g1 + (3 * f1 + 2 * f2) * g2) * h1 + ((3 * f1 + 2 * f2) * g1 + (2 * f1 + 3 * f2) * g2) * h2);
    I[1] += t1 * (((3 * f1 + 2 * f2) * g1 + (2 * f1 + 3 * f2) * g2) * h1 + ((2 * f1 + 3 * f2) * g1 + 3 * (f1 + 4 * f2) * g2) * h2);
}

// ---------------------------------------------------------------- friction
// fr_bedshstr_drag_coef (src/friction/fr_bedshstr_drag_coef.c:18-49)
HDF double drag_coef(double depth, double rh, const crlog::tables *T) {
    depth = dmax(depth, 1.E-8);
    rh = dmax(rh, 1.E-8);
    double depth_rh;
    if (rh > 29.7 * depth) depth_rh = 1. / 29.7;
    else depth_rh = depth / rh;
    double bta = 29.7 * depth_rh;
    double q = (.4 * bta) / ((bta + 1) * (crlog::cr_log_t(bta + 1, T) - 1) + 1);
    return 2. * (q * q);  // pow(q, 2.) correctly rounded == q*q
}

// ---------------------------------------------------------------- wet/dry
// integrand selectors for the wet/dry wrapper (fe_sw2_wet_dry_integrations.c)
enum { WD_INTEGRATE_F = 0, WD_CONT_TEMPORAL, WD_CONVECTION, WD_GLS_CONVECTION, WD_PRESSURE, WD_BODYFORCE,
       WD_BOUNDARY_PRESSURE };

template <int FN>
HDF void wd_apply(const v3 *x, const double *h, const v2 *gp, v2 vconst, const v2 *vwd, const double *fwd, double djac,
                  double dt, double c, d3 *rhs) {
    if (FN == WD_INTEGRATE_F) {  // :95-116
        double t1 = div_rn(h[0] + h[1] + h[2], 3.0, RN_ONE_THIRD), t2 = 0.0, t3 = 0.0;
        if (vwd) {
            t2 = div_rn(vwd[0].x + vwd[1].x + vwd[2].x, 3.0, RN_ONE_THIRD);
            t3 = div_rn(vwd[0].y + vwd[1].y + vwd[2].y, 3.0, RN_ONE_THIRD);
        }
        rhs[0].c_eq += djac * t1;
        rhs[0].x_eq += djac * t2;
        rhs[0].y_eq += djac * t3;
    } else if (FN == WD_CONT_TEMPORAL) {  // :143-160
        double rc[3] = {0., 0., 0.};
        int_tri_phi_f(djac, c, h, rc);
        for (int i = 0; i < 3; i++) rhs[i].c_eq += rc[i];
    } else if (FN == WD_CONVECTION) {  // :200-229
        double rc[3] = {0., 0., 0.}, rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
        c *= dt;
        double uu[3] = {vwd[0].x, vwd[1].x, vwd[2].x}, vv[3] = {vwd[0].y, vwd[1].y, vwd[2].y};
        {   // integrate_triangle_gradPhi_dot_f_v :546-561
            double p1 = (2 * h[0] + h[1] + h[2]);
            double p2 = (h[0] + 2 * h[1] + h[2]);
            double p3 = (h[0] + h[1] + 2 * h[2]);
            double tx = (p1 * vwd[0].x + p2 * vwd[1].x + p3 * vwd[2].x);
            double ty = (p1 * vwd[0].y + p2 * vwd[1].y + p3 * vwd[2].y);
            double con = djac * one_12 * c;
            for (int i = 0; i < 3; i++) rc[i] += con * (gp[i].x * tx + gp[i].y * ty);
        }
        {   // integrate_triangle_dphi_f_g_h(..., uu, vv, h, rhs_y, rhs_x) :832-850
            const double *f = uu, *g = vv;
            double ff1 = (3 * f[0] + f[1] + f[2]);
            double ff2 = (2 * f[0] + 2 * f[1] + f[2]);
            double ff3 = (2 * f[0] + f[1] + 2 * f[2]);
            double ff4 = (f[0] + 3 * f[1] + f[2]);
            double ff5 = (f[0] + 2 * f[1] + 2 * f[2]);
            double ff6 = (f[0] + f[1] + 3 * f[2]);
            double gg1 = (2 * ff1 * g[0] + ff2 * g[1] + ff3 * g[2]);
            double gg2 = (ff2 * g[0] + 2 * ff4 * g[1] + ff5 * g[2]);
            double gg3 = (ff3 * g[0] + ff5 * g[1] + 2 * ff6 * g[2]);
            double con = djac * one_60 * c * (gg1 * h[0] + gg2 * h[1] + gg3 * h[2]);
            for (int i = 0; i < 3; i++) ry[i] += con * gp[i].x;
            for (int i = 0; i < 3; i++) rx[i] += con * gp[i].y;
        }
        // integrate_triangle_dphi_f_f_h :864-884, once with uu (x) and once with vv (y)
        for (int pass = 0; pass < 2; pass++) {
            const double *f = pass == 0 ? uu : vv;
            double f1_2 = f[0] * f[0], f2_2 = f[1] * f[1], f3_2 = f[2] * f[2];
            double p1 = (3 * f1_2 + 2 * f[0] * f[1] + f2_2 + (2 * f[0] + f[1]) * f[2] + f3_2);
            double p2 = (f1_2 + 2 * f[0] * f[1] + 3 * f2_2 + (f[0] + 2 * f[1]) * f[2] + f3_2);
            double p3 = (f1_2 + f[0] * f[1] + f2_2 + 2 * (f[0] + f[1]) * f[2] + 3 * f3_2);
            double con = djac * one_30 * c * (p1 * h[0] + p2 * h[1] + p3 * h[2]);
            if (pass == 0) for (int i = 0; i < 3; i++) rx[i] += con * gp[i].x;
            else for (int i = 0; i < 3; i++) ry[i] += con * gp[i].y;
        }
        for (int i = 0; i < 3; i++) {
            rhs[i].c_eq += rc[i];
            rhs[i].x_eq += rx[i];
            rhs[i].y_eq += ry[i];
        }
    } else if (FN == WD_GLS_CONVECTION) {  // :361-381
        double h_avg = one_3 * (fwd[0] + fwd[1] + fwd[2]);
        double u_avg = one_3 * (vwd[0].x + vwd[1].x + vwd[2].x);
        double v_avg = one_3 * (vwd[0].y + vwd[1].y + vwd[2].y);
        double dhdx = gp[0].x * fwd[0] + gp[1].x * fwd[1] + gp[2].x * fwd[2];
        double dhdy = gp[0].y * fwd[0] + gp[1].y * fwd[1] + gp[2].y * fwd[2];
        double dudx = gp[0].x * vwd[0].x + gp[1].x * vwd[1].x + gp[2].x * vwd[2].x;
        double dudy = gp[0].y * vwd[0].x + gp[1].y * vwd[1].x + gp[2].y * vwd[2].x;
        double dvdx = gp[0].x * vwd[0].y + gp[1].x * vwd[1].y + gp[2].x * vwd[2].y;
        double dvdy = gp[0].y * vwd[0].y + gp[1].y * vwd[1].y + gp[2].y * vwd[2].y;
        rhs[0].c_eq += djac * dt * c * (u_avg * dhdx + h_avg * dudx + v_avg * dhdy + h_avg * dvdy);
        rhs[0].x_eq += djac * dt * c * (u_avg * dudx + v_avg * dudy);
        rhs[0].y_eq += djac * dt * c * (u_avg * dvdx + v_avg * dvdy);
    } else if (FN == WD_PRESSURE) {  // :412-425 with integrate_triangle_dphi_f_f :352-367
        double int_f_f = int_tri_f_f(djac, c, h);
        double rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
        for (int i = 0; i < 3; i++) rx[i] += gp[i].x * int_f_f;
        for (int i = 0; i < 3; i++) ry[i] += gp[i].y * int_f_f;
        for (int i = 0; i < 3; i++) { rhs[i].x_eq += rx[i]; rhs[i].y_eq += ry[i]; }
    } else if (FN == WD_BODYFORCE) {  // :470-482 with integrate_triangle_phi_h_df :1016-1031
        double rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
        double sum_h = h[0] + h[1] + h[2];
        double con = djac * one_12 * c * vconst.x;
        for (int i = 0; i < 3; i++) rx[i] += con * (sum_h + h[i]);
        con = djac * one_12 * c * vconst.y;
        for (int i = 0; i < 3; i++) ry[i] += con * (sum_h + h[i]);
        for (int i = 0; i < 3; i++) { rhs[i].x_eq += rx[i]; rhs[i].y_eq += ry[i]; }
    } else if (FN == WD_BOUNDARY_PRESSURE) {  // :511-537
        for (int iedge = 0; iedge < 3; iedge++) {
            int nd1 = iedge, nd2 = iedge == 2 ? 0 : iedge + 1;
            double delx = x[nd1].x - x[nd2].x;
            double dely = x[nd2].y - x[nd1].y;
            double r[2] = {0., 0.};
            int_line_phi_h_h(1., c, h[nd1], h[nd2], r);
            rhs[nd1].x_eq += dely * r[0];
            rhs[nd2].x_eq += dely * r[1];
            rhs[nd1].y_eq += delx * r[0];
            rhs[nd2].y_eq += delx * r[1];
        }
    }
}

// Register-resident selection of element [k] of a 3-array by a runtime index
// (avoids dynamically indexed private arrays, which would live in scratch).
template <typename T>
HDF T pick3(const T *a, int k) { return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]); }

// fe_sw2_wet_dry_wrapper (src/fe/fe_sw2_wet_dry_wrapper.c:106-306).  Flags
// select which optional inputs exist (grad_phi, v_wet_dry, f_wet_dry).
// WET: the caller guarantees h > 0 at all three nodes, so only the fully
// wet branch is compiled (the element kernels run such triangles through a
// variant without the clipping code and its registers).
template <int FN, bool HAS_GRAD, bool HAS_VWD, bool HAS_FWD, bool REDIST, bool WET = false>
HDF double wd_wrapper(d3 *elem_rhs, const v3 *x, const double *h, const v2 *grad_phi, v2 vconst, const v2 *v_wet_dry,
                      const double *f_wet_dry, double djac, double dt, double c) {
    if constexpr (WET) {
        wd_apply<FN>(x, h, grad_phi, vconst, HAS_VWD ? v_wet_dry : nullptr, f_wet_dry, djac, dt, c, elem_rhs);
        return 1.0;
    }
    int total_vtx = 0, passcode = 0;
    for (int i = 0; i < 3; i++) {
        int is_dry = (h[i] <= 0.0) ? 1 : 0;
        total_vtx += is_dry;
        passcode += (is_dry << i);
    }
    if (total_vtx == 3) return 0;
    if (total_vtx == 0) {
        wd_apply<FN>(x, h, grad_phi, vconst, HAS_VWD ? v_wet_dry : nullptr, f_wet_dry, djac, dt, c, elem_rhs);
        return 1.0;
    }
    if (total_vtx == 1) wd_apply<FN>(x, h, grad_phi, vconst, HAS_VWD ? v_wet_dry : nullptr, f_wet_dry, djac, dt, c, elem_rhs);
    else passcode = (7 - passcode);
    int only = passcode / 2;

    double new_h[3] = {0.0, 0.0, 0.0};
    v3 nx[3];
    double nf[3] = {0.0, 0.0, 0.0};
    v2 nv[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    const double ho = pick3(h, only);
    const v3 xo = pick3(x, only);
    const v2 vo = HAS_VWD ? pick3(v_wet_dry, only) : v2{0.0, 0.0};
    const double fo = HAS_FWD ? pick3(f_wet_dry, only) : 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i != only) {
            double xsi = h[i] / (h[i] - ho);
            nx[i].x = x[i].x + xsi * (xo.x - x[i].x);   // FIXED_POS macro.h:756
            nx[i].y = x[i].y + xsi * (xo.y - x[i].y);
            nx[i].z = x[i].z;
            if (HAS_VWD) {
                nv[i].x = v_wet_dry[i].x + xsi * (vo.x - v_wet_dry[i].x);
                nv[i].y = v_wet_dry[i].y + xsi * (vo.y - v_wet_dry[i].y);
            }
            if (HAS_FWD) nf[i] = f_wet_dry[i] + xsi * (fo - f_wet_dry[i]);
        } else {
            nx[i] = x[i];
            new_h[i] = h[i];
            if (HAS_FWD) nf[i] = f_wet_dry[i];
            if (HAS_VWD) nv[i] = v_wet_dry[i];
        }
    }
    double elem_djac = djac;
    {   // TRI_AREA macro.h:773
        double a0x = nx[1].x - nx[0].x, a0y = nx[1].y - nx[0].y;
        double a1x = nx[2].x - nx[0].x, a1y = nx[2].y - nx[0].y;
        djac = (a0x * a1y - a0y * a1x) / 2.0;
    }
    if (djac / elem_djac < SMALL) return (double)(total_vtx % 2);
    v2 ng[3];
    if (HAS_GRAD) {  // GRAD_SHAPE macro.h:779
        ng[0].x = 0.5 * (nx[1].y - nx[2].y) / djac;
        ng[1].x = 0.5 * (nx[2].y - nx[0].y) / djac;
        ng[2].x = 0.5 * (nx[0].y - nx[1].y) / djac;
        ng[0].y = 0.5 * (nx[2].x - nx[1].x) / djac;
        ng[1].y = 0.5 * (nx[0].x - nx[2].x) / djac;
        ng[2].y = 0.5 * (nx[1].x - nx[0].x) / djac;
    }
    d3 nr[3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    wd_apply<FN>(nx, new_h, ng, vconst, HAS_VWD ? nv : nullptr, nf, djac, dt, c, nr);
    double factor = (total_vtx >> 1) - (total_vtx % 2);
    if (REDIST) {  // RE_DISTRIBUTE macro.h:789-805
        d3 no = pick3(nr, only);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i == only) continue;
            double a0x = nx[i].x - xo.x, a0y = nx[i].y - xo.y;
            double a1x = x[i].x - xo.x, a1y = x[i].y - xo.y;
            double l0 = a0x * a0x + a0y * a0y, l1 = a1x * a1x + a1y * a1y;
            double xsi = sqrt(l0 / l1);
            no.c_eq += (1.0 - xsi) * nr[i].c_eq;
            no.x_eq += (1.0 - xsi) * nr[i].x_eq;
            no.y_eq += (1.0 - xsi) * nr[i].y_eq;
            nr[i].c_eq *= xsi;
            nr[i].x_eq *= xsi;
            nr[i].y_eq *= xsi;
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (i == only) nr[i] = no;
    }
    for (int i = 0; i < 3; i++) {
        elem_rhs[i].c_eq += factor * nr[i].c_eq;
        elem_rhs[i].x_eq += factor * nr[i].x_eq;
        elem_rhs[i].y_eq += factor * nr[i].y_eq;
    }
    factor *= djac / elem_djac;
    return factor + ((factor < 0) ? 1.0 : 0.0);
}

// fe_sw2_wet_dry_factor (src/fe/fe_sw2_wet_dry_wrapper.c:31-73)
template <bool WET = false>
HD double wd_factor(const v3 *x, const double *h, double djac) {
    if constexpr (WET) return 1.0;
    int total_vtx = 0, passcode = 0;
    for (int i = 0; i < 3; i++) {
        int is_dry = (h[i] <= 0.0) ? 1 : 0;
        total_vtx += is_dry;
        passcode += (is_dry << i);
    }
    if (total_vtx == 3) return 0.0;
    if (total_vtx == 0) return 1.0;
    if (total_vtx != 1) passcode = (7 - passcode);
    int only = passcode / 2;
    v2 nx[3];
    const double ho = pick3(h, only);
    const v3 xo = pick3(x, only);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i != only) {
            double xsi = h[i] / (h[i] - ho);
            nx[i].x = x[i].x + xsi * (xo.x - x[i].x);
            nx[i].y = x[i].y + xsi * (xo.y - x[i].y);
        } else {
            nx[i].x = x[i].x;
            nx[i].y = x[i].y;
        }
    }
    double elem_djac = djac;
    double a0x = nx[1].x - nx[0].x, a0y = nx[1].y - nx[0].y;
    double a1x = nx[2].x - nx[0].x, a1y = nx[2].y - nx[0].y;
    djac = (a0x * a1y - a0y * a1x) / 2.0;
    if (djac / elem_djac < SMALL) return (double)(total_vtx % 2);
    double factor = (total_vtx >> 1) - (total_vtx % 2);
    factor *= djac / elem_djac;
    return factor + ((factor < 0) ? 1.0 : 0.0);
}

// fe_sw2_wd_average (src/fe/sw2/fe_sw2_wet_dry_integrations.c:42-61)
template <bool WET = false>
HDF void wd_average(const v3 *x, const double *depth, const v2 *v_wd, double djac, double *f_avg, v2 *v_avg) {
    d3 av[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    double factor = wd_wrapper<WD_INTEGRATE_F, false, true, true, false, WET>(av, x, depth, nullptr, v2{0., 0.}, v_wd, depth,
                                                                        djac, 1., 1.);
    *f_avg = 0.;
    v_avg->x = 0.;
    v_avg->y = 0.;
    if (factor > 0.0) {
        double wet_area = djac * factor;
        *f_avg = av[0].c_eq / wet_area;
        v_avg->x = av[0].x_eq / wet_area;
        v_avg->y = av[0].y_eq / wet_area;
    }
}

// fe_sw2_temporal (src/fe/sw2/fe_sw2_body_resid.c:1288-1354, triangle branch)
template <bool WET = false>
HDF void sw2_temporal(const v3 *nds, double djac, double dll, const double *h, const v2 *vel, double wdf,
                      double dt_factor, d3 *elem_rhs, int wd_flag) {
    double con = dt_factor * wdf;
    d3 rhs[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    double uh[3], vh[3], rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
    for (int i = 0; i < 3; i++) {
        uh[i] = vel[i].x * dmax(h[i], dll);
        vh[i] = vel[i].y * dmax(h[i], dll);
    }
    wd_wrapper<WD_CONT_TEMPORAL, false, false, false, true, WET>(rhs, nds, h, nullptr, v2{0., 0.}, nullptr, nullptr, djac,
                                                             1.0, dt_factor);
    int_tri_phi_f_lump(djac, con, uh, rx);
    int_tri_phi_f_lump(djac, con, vh, ry);
    for (int i = 0; i < 3; i++) {
        if (wd_flag != 2) elem_rhs[i].c_eq += rhs[i].c_eq;
        elem_rhs[i].x_eq += rx[i];
        elem_rhs[i].y_eq += ry[i];
    }
}

// Per-element terms of the body residual that depend only on the old and
// older states (fixed during a Newton solve), computed once per solve with
// the same code and consumed by body_resid in the reference's order.
// Layout of the 24 doubles per triangle:
//   [0..8]   t(n) temporal contribution: continuity c0 c1 c2, x0 x1 x2, y0 y1 y2
//   [9..17]  t(n-1) temporal contribution, same layout
//   [18]     wet/dry factor of the old depths (SUPG length scale)
//   [19..21] minus the wet/dry-averaged old u, v, h integrals (SUPG d/dt seeds)
constexpr int PRE_STRIDE = 24;

HDF void sw2_temporal_terms(const v3 *nds, double djac, double dll, const double *h, const v2 *vel, double wdf,
                            double dt_factor, double *out9) {
    double con = dt_factor * wdf;
    d3 rhs[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    double uh[3], vh[3], rx[3] = {0., 0., 0.}, ry[3] = {0., 0., 0.};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        uh[i] = vel[i].x * dmax(h[i], dll);
        vh[i] = vel[i].y * dmax(h[i], dll);
    }
    wd_wrapper<WD_CONT_TEMPORAL, false, false, false, true>(rhs, nds, h, nullptr, v2{0., 0.}, nullptr, nullptr, djac,
                                                             1.0, dt_factor);
    int_tri_phi_f_lump(djac, con, uh, rx);
    int_tri_phi_f_lump(djac, con, vh, ry);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        out9[i] = rhs[i].c_eq;
        out9[3 + i] = rx[i];
        out9[6 + i] = ry[i];
    }
}

HD void elem_old_terms(const ElemData &D, int ie, double *pre) {
    const int nd[3] = {D.e2d[3 * ie], D.e2d[3 * ie + 1], D.e2d[3 * ie + 2]};
    const double djac = D.djac[ie], alpha = D.alpha;
    v3 nds[3];
    double hold[3], holder[3];
    v2 vold[3], volder[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int n = nd[i];
        nds[i].x = D.x[n]; nds[i].y = D.y[n]; nds[i].z = D.z[n];
        hold[i] = D.old_head[n];
        holder[i] = D.older_head[n];
        vold[i].x = D.old_u[n]; vold[i].y = D.old_v[n];
        volder[i].x = D.older_u[n]; volder[i].y = D.older_v[n];
    }
    const double factor_old = wd_factor(nds, hold, djac);
    const double factor_older = wd_factor(nds, holder, djac);
    sw2_temporal_terms(nds, djac, D.dll, hold, vold, factor_old, (-1.) * (1. + alpha), pre);
    sw2_temporal_terms(nds, djac, D.dll, holder, volder, factor_older, (alpha / 2.), pre + 9);
    pre[18] = factor_old;
    d3 dfdt[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    wd_wrapper<WD_INTEGRATE_F, false, true, true, false>(dfdt, nds, hold, nullptr, v2{0., 0.}, vold, hold, djac, 1, 1);
    pre[19] = -dfdt[0].x_eq;
    pre[20] = -dfdt[0].y_eq;
    pre[21] = -dfdt[0].c_eq;
    pre[22] = 0.;
    pre[23] = 0.;
}

// ------------------------------------------------------------ body residual
// fe_sw2_body_resid (src/fe/sw2/fe_sw2_body_resid.c:45-1267), triangles.
// dry_rule: apply the reference's in-loop zeroing of dry-node velocities
// (fe_sw2_resid.c:85-93): a node with old_head <= 0 keeps its velocity only
// for the lowest-index element that touches it.  Old/older-state terms come
// from `pre` (elem_old_terms); the reference's wet/dry average of the old
// state (:248) feeds only debug output and is not formed.
// Where body_resid reads one triangle's inputs: the model arrays themselves
// (GlobalSrc), or a copy a kernel has staged in LDS (sw2_kernels.hip).  Node
// fields are addressed by the triangle's local node 0..2.
struct GlobalSrc {
    static constexpr bool has_rough0 = false;   // see body_resid_src's friction coefficient
    const ElemData &D;
    const int *nd;
    int ie;
    const double *pre_;
    HDM double rough0() const { return 0.0; }
    HDM double x(int i) const { return D.x[nd[i]]; }
    HDM double y(int i) const { return D.y[nd[i]]; }
    HDM double z(int i) const { return D.z[nd[i]]; }
    HDM double head(int i) const { return D.head[nd[i]]; }
    HDM double u(int i) const { return D.u[nd[i]]; }
    HDM double v(int i) const { return D.v[nd[i]]; }
    HDM int node_bc(int i) const { return D.node_bc[nd[i]]; }
    HDM double grad(int k) const { return D.grad[6 * ie + k]; }
    HDM double djac() const { return D.djac[ie]; }
    HDM int mat() const { return D.mat[ie]; }
    HDM int wd_flag() const { return D.wd_flag[ie]; }
    HDM double ks() const { return D.ks[ie]; }
    HDM int src_flag() const { return D.src_flag[ie]; }
    HDM double src() const { return D.src[ie]; }
    HDM double pre(int k) const { return pre_[k]; }
};

template <bool WET = false, typename SRC>
HD void body_resid_src(const ElemData &D, const SRC &S, int ie, double pert, int pnode, int pvar, int psign,
                       bool dry_rule, d3 *elem_rhs, const int *nd);

template <bool WET = false>
HD void body_resid(const ElemData &D, const double *pre, int ie, double pert, int pnode, int pvar, int psign,
                   bool dry_rule, d3 *elem_rhs, const int *ndp = nullptr) {
    // ndp: the triangle's node ids when the caller already holds them
    const int nd[3] = {ndp ? ndp[0] : D.e2d[3 * ie], ndp ? ndp[1] : D.e2d[3 * ie + 1],
                       ndp ? ndp[2] : D.e2d[3 * ie + 2]};
    body_resid_src<WET>(D, GlobalSrc{D, nd, ie, pre}, ie, pert, pnode, pvar, psign, dry_rule, elem_rhs, nd);
}

// The drag coefficient of the triangle at its unperturbed state, by the
// operations body_resid_src performs to reach it (node coordinates and
// depths, wd_average's depth average, the SMALL floor, drag_coef).
template <bool WET, typename SRC>
HD double unperturbed_roughness(const ElemData &D, const SRC &S) {
    v3 nds[3];
    double h[3];
    v2 vel[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        nds[i].x = S.x(i); nds[i].y = S.y(i); nds[i].z = S.z(i);
        h[i] = S.head(i);
        vel[i].x = S.u(i); vel[i].y = S.v(i);
    }
    double h_wd = int_tri_f(1., 1., h);
    v2 vwd;
    wd_average<WET>(nds, h, vel, S.djac(), &h_wd, &vwd);
    const double ks = S.ks();
    return ks > SMALL ? drag_coef(dmax(h_wd, SMALL), ks, D.logt) : 0.0;
}

// The body residual with its inputs read through S (GlobalSrc or a staged
// copy); nd is only used by the dry-node rule and the winds.
template <bool WET, typename SRC>
HD void body_resid_src(const ElemData &D, const SRC &S, int ie, double pert, int pnode, int pvar, int psign,
                       bool dry_rule, d3 *elem_rhs, const int *nd) {
    int PRESSURE_FLAG = 1;
    v2 gs[3];
    for (int i = 0; i < 3; i++) {
        gs[i].x = S.grad(2 * i);
        gs[i].y = S.grad(2 * i + 1);
    }
    const double alpha = D.alpha, dt = elem_dt(D, ie), djac = S.djac(), g = D.g, dll = D.dll, dul = D.dul;
    const int imat = S.mat();
    const int wd_flag = S.wd_flag();
    double t1;

    double node_z[3];
    v3 nds[3];
    double h[3];
    v2 vel[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        nds[i].x = S.x(i); nds[i].y = S.y(i); nds[i].z = S.z(i);
        node_z[i] = nds[i].z;
        h[i] = S.head(i);
        double un = S.u(i), vn = S.v(i);
        if (dry_rule && D.old_head[nd[i]] <= 0. && D.first_elem[nd[i]] != ie) { un = 0.; vn = 0.; }
        vel[i].x = un; vel[i].y = vn;
    }
    if (pvar == PERTURB_U || pvar == PERTURB_V) PRESSURE_FLAG = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i != pnode) continue;
        if (pvar == PERTURB_H) h[i] += psign * pert;
        else if (pvar == PERTURB_U) vel[i].x += psign * pert;
        else if (pvar == PERTURB_V) vel[i].y += psign * pert;
    }
    // baroclinic density and lid pressures are zero for the supported decks
    const double elem_density[3] = {0., 0., 0.};
    const double elem_pressure[3] = {0., 0., 0.};
    double eu[3] = {vel[0].x, vel[1].x, vel[2].x}, ev[3] = {vel[0].y, vel[1].y, vel[2].y};

    const double area = djac;
    const double elem_h_avg = int_tri_f(1., 1., h);
    const double elem_u_avg = int_tri_f(1., 1., eu);
    const double elem_v_avg = int_tri_f(1., 1., ev);
    const double elem_density_avg = int_tri_f(1., 1., elem_density);
    v2 ggz = {0., 0.}, ggh = {0., 0.}, ggp = {0., 0.}, ggu = {0., 0.}, ggv = {0., 0.}, ggw;
    for (int i = 0; i < 3; i++) { ggz.x += node_z[i] * gs[i].x; ggz.y += node_z[i] * gs[i].y; }
    for (int i = 0; i < 3; i++) { ggh.x += h[i] * gs[i].x; ggh.y += h[i] * gs[i].y; }
    for (int i = 0; i < 3; i++) { ggp.x += elem_pressure[i] * gs[i].x; ggp.y += elem_pressure[i] * gs[i].y; }
    for (int i = 0; i < 3; i++) {
        ggu.x += vel[i].x * gs[i].x; ggu.y += vel[i].x * gs[i].y;
        ggv.x += vel[i].y * gs[i].x; ggv.y += vel[i].y * gs[i].y;
    }
    ggw.x = ggz.x + ggh.x;
    ggw.y = ggz.y + ggh.y;
    const double vmag = sqrt(elem_u_avg * elem_u_avg + elem_v_avg * elem_v_avg);
    const double vst_i = elem_u_avg / dmax(vmag, 1.E-8);
    const double vst_j = elem_v_avg / dmax(vmag, 1.E-8);

    double factor;
    const double factor_old = S.pre(18);
    double h_wd = elem_h_avg;
    v2 vwd;
    wd_average<WET>(nds, h, vel, djac, &h_wd, &vwd);
    factor = wd_factor<WET>(nds, h, djac);
    const double vwd_mag = sqrt(vwd.x * vwd.x + vwd.y * vwd.y);

    const double h_fric = dmax(h_wd, SMALL);
    const double ks = S.ks();
    // h_wd depends on the depths alone (wd_average's f_avg), so under a u or v
    // perturbation the drag coefficient is the unperturbed triangle's; a source
    // that carries it (computed once per triangle by unperturbed_roughness, the
    // same operations on the same inputs) saves its log for those columns
    double roughness;
    if constexpr (SRC::has_rough0) {
        if (pvar == PERTURB_U || pvar == PERTURB_V) roughness = S.rough0();
        else roughness = ks > SMALL ? drag_coef(h_fric, ks, D.logt) : 0.0;
    } else {
        roughness = ks > SMALL ? drag_coef(h_fric, ks, D.logt) : 0.0;
    }

    double ev_st = 0., ev_tr = 0.;
    const int turb = D.mat_turb[imat];
    if (turb == TURB_EVS) {
        double exx = D.mat_ev[3 * imat], eyy = D.mat_ev[3 * imat + 1];
        exx += D.visc;
        eyy += D.visc;
        ev_st = fabs(exx - eyy);
        ev_tr = dmin(exx, eyy);
    } else if (turb == TURB_EEV) {  // fe_sw2_get_EEVF (src/fe/sw2/fe_sw2_get_EEVF.c:33-60)
        int mode = D.mat_eev_mode[imat];
        double coef = D.mat_eev_coef[imat];
        double ev_min = 0., vmin = 0.1 * sqrt(g * dll);
        if (mode == 1) {
            ev_st = 0.0;
            ev_tr = coef * 2.0 * h_fric * sqrt(roughness) * vmag;
            ev_min = (coef / 0.5) * vmin * sqrt(area) / 40.;
            ev_tr = dmax(ev_tr, ev_min);
        } else if (mode == 2) {
            ev_st = 1.3 * h_fric * sqrt(roughness) * vmag;
            ev_tr = coef * 0.92 * h_fric * sqrt(roughness) * vmag;
            ev_min = (coef / 0.5) * vmin * sqrt(area) / 40.;
            ev_tr = dmax(ev_tr, ev_min);
        } else if (mode == 3) {
            ev_st = 0.;
            ev_tr = coef * coef * area * sqrt(ggu.x * ggu.x + ggv.y * ggv.y + 0.5 * (ggu.y + ggv.x) * (ggu.y + ggv.x));
            ev_min = (coef / 0.2) * vmin * sqrt(area) / 40.;
            ev_tr = dmax(ev_tr, ev_min);
        }
        if (wd_flag == 1 || wd_flag == 2) {
            ev_st = 0.0;
            ev_tr = dmax(ev_tr, 5. * ev_min);
        }
    }

    for (int i = 0; i < 3; i++) { elem_rhs[i].x_eq = 0.; elem_rhs[i].y_eq = 0.; elem_rhs[i].c_eq = 0.; }
    d3 rhs[3];
    double rcq[3], rxq[3], ryq[3];

    // shock capturing :360-419
    if (wd_flag == 1 || wd_flag == 2) {
        double umag = sqrt(ggu.x * ggu.x + ggu.y * ggu.y + SMALL);
        double vmg = sqrt(ggv.x * ggv.x + ggv.y * ggv.y + SMALL);
        double cu = one_2 * dul * djac * umag;
        double cv = one_2 * dul * djac * vmg;
        v2 gun = {ggu.x / umag, ggu.y / umag}, gvn = {ggv.x / vmg, ggv.y / vmg}, norm;
        for (int i = 0; i < 3; i++) {
            rhs[i].c_eq = dt * one_2 * g * djac * ((dll * ggw.x) * gs[i].x + (dll * ggw.y) * gs[i].y);
            norm.x = gun.x * ggu.x + gun.y * ggu.y;
            norm.y = gvn.x * ggv.x + gvn.y * ggv.y;
            rhs[i].x_eq = cu * dt * djac * norm.x * (gun.x * gs[i].x + gun.y * gs[i].y);
            rhs[i].y_eq = cv * dt * djac * norm.y * (gvn.x * gs[i].x + gvn.y * gs[i].y);
        }
        for (int i = 0; i < 3; i++) {
            elem_rhs[i].c_eq += rhs[i].c_eq;
            elem_rhs[i].x_eq += rhs[i].x_eq;
            elem_rhs[i].y_eq += rhs[i].y_eq;
        }
    }

    // advection :434-461
    for (int i = 0; i < 3; i++) { rhs[i].x_eq = 0.; rhs[i].y_eq = 0.; rhs[i].c_eq = 0.; }
    wd_wrapper<WD_CONVECTION, true, true, false, true, WET>(rhs, nds, h, gs, v2{0., 0.}, vel, nullptr, djac, dt, -1);
    for (int i = 0; i < 3; i++) {
        elem_rhs[i].c_eq += rhs[i].c_eq;
        elem_rhs[i].x_eq += rhs[i].x_eq;
        elem_rhs[i].y_eq += rhs[i].y_eq;
    }

    // temporal :480-490
    sw2_temporal<WET>(nds, djac, dll, h, vel, factor, (1. + alpha / 2.), elem_rhs, -3);
#pragma unroll
    for (int k = 0; k < 2; k++) {   // t(n) and t(n-1) terms, fe_sw2_temporal with the element's wd_flag
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (wd_flag != 2) elem_rhs[i].c_eq += S.pre(9 * k + i);
            elem_rhs[i].x_eq += S.pre(9 * k + 3 + i);
            elem_rhs[i].y_eq += S.pre(9 * k + 6 + i);
        }
    }

    // diffusion :504-539
    {
        double temp = h_wd;
        if (h_wd < dul) temp = dul;
        for (int i = 0; i < 3; i++) { rxq[i] = 0.; ryq[i] = 0.; }
        v2 dfx, dfy;
        dfx.x = 2. * ev_st * vst_i * (vst_i * ggu.x + vst_j * ggu.y) + 2. * ev_tr * ggu.x;
        dfx.y = ev_st * vst_j * (vst_j * ggu.y + vst_i * ggu.x + vst_i * ggv.x + vst_j * ggv.y) + ev_tr * (ggu.y + ggv.x);
        dfy.x = ev_st * vst_i * (vst_j * ggu.y + vst_i * ggu.x + vst_i * ggv.x + vst_j * ggv.y) + ev_tr * (ggu.y + ggv.x);
        dfy.y = 2. * ev_st * vst_j * (vst_j * ggv.y + vst_i * ggv.x) + 2. * ev_tr * ggv.y;
        int_tri_gradPhi_dot_vbar(gs, djac, temp, dfx, rxq);
        int_tri_gradPhi_dot_vbar(gs, djac, temp, dfy, ryq);
        t1 = dt * factor;
        for (int i = 0; i < 3; i++) {
            elem_rhs[i].x_eq += t1 * rxq[i];
            elem_rhs[i].y_eq += t1 * ryq[i];
        }
    }

    // friction :557-594
    const double resistance = roughness / (2.0 * g);
    {
        double fx[3], fy[3];
        for (int i = 0; i < 3; i++) {
            t1 = sqrt(vel[i].x * vel[i].x + vel[i].y * vel[i].y);
            fx[i] = vel[i].x * t1;
            fy[i] = vel[i].y * t1;
        }
        for (int i = 0; i < 3; i++) { rxq[i] = 0.; ryq[i] = 0.; }
        t1 = dt * g * resistance * factor;
        int_tri_phi_f(djac, t1, fx, rxq);
        int_tri_phi_f(djac, t1, fy, ryq);
        for (int i = 0; i < 3; i++) {
            elem_rhs[i].x_eq += rxq[i];
            elem_rhs[i].y_eq += ryq[i];
        }
    }
    v2 bed_stress;
    bed_stress.x = resistance * vwd.x * vwd_mag;
    bed_stress.y = resistance * vwd.y * vwd_mag;

    // coriolis :639-652
    const double cor = D.mat_cor[imat];
    {
        for (int i = 0; i < 3; i++) { rxq[i] = 0.; ryq[i] = 0.; }
        t1 = dt * factor * cor;
        int_tri_phi_f_g(djac, -t1, h, ev, rxq);
        int_tri_phi_f_g(djac, +t1, h, eu, ryq);
        for (int i = 0; i < 3; i++) {
            elem_rhs[i].x_eq += rxq[i];
            elem_rhs[i].y_eq += ryq[i];
        }
    }
    // wind and wave stress (:674-719) on fully wet triangles; without winds
    // (there are never waves here) the update elem_rhs -= dt * 0.0 is an
    // exact no-op and is skipped
    v2 total_stress = {0., 0.};
    if (wd_flag == 0 && D.wind) {
        // swind_elem2d_get_local_stress (src/structs/smeteor.c:241-308), stress
        // law 0: the node value is the shear; windatt * windatt is the
        // reference's pow(windatt, 2.0) for the attenuations decks use (1.0)
        v2 wind_stress = {0., 0.};
        for (int i = 0; i < 3; i++) {
            double shear_x = 0.0, shear_y = 0.0;
            if (D.density > 0.0) {
                shear_x = D.wind[2 * nd[i]];
                shear_y = D.wind[2 * nd[i] + 1];
            }
            wind_stress.x += shear_x;
            wind_stress.y += shear_y;
        }
        const double watt = D.mat_windatt[imat];
        wind_stress.x *= (watt * watt) / (3 * D.density);
        wind_stress.y *= (watt * watt) / (3 * D.density);
        total_stress.x = wind_stress.x + 0.0;   // + wave stress
        total_stress.y = wind_stress.y + 0.0;
        double sx[3] = {0., 0., 0.}, sy[3] = {0., 0., 0.};
        t1 = total_stress.x * djac * one_3;     // integrate_triangle_phi (fe_elem2d_integrations.c:159-164)
        for (int i = 0; i < 3; i++) sx[i] += t1;
        t1 = total_stress.y * djac * one_3;
        for (int i = 0; i < 3; i++) sy[i] += t1;
        for (int i = 0; i < 3; i++) {
            elem_rhs[i].x_eq -= dt * sx[i];
            elem_rhs[i].y_eq -= dt * sy[i];
        }
    }

    // rainfall :728-754
    double src = 0.;
    if (S.src_flag()) {
        src = S.src();
        for (int i = 0; i < 3; i++) rcq[i] = 0.;
        t1 = src * djac * one_3;   // integrate_triangle_phi :159-164
        for (int i = 0; i < 3; i++) rcq[i] += t1;
        t1 = dt * factor;
        for (int i = 0; i < 3; i++) elem_rhs[i].c_eq -= t1 * rcq[i];
    }

    // pressure, bed slope, boundary pressure :761-862
    if (PRESSURE_FLAG) {
        for (int pass = 0; pass < 3; pass++) {
            for (int i = 0; i < 3; i++) { rhs[i].x_eq = 0.; rhs[i].y_eq = 0.; rhs[i].c_eq = 0.; }
            if (pass == 0) {
                t1 = -one_2 * g * dt;
                wd_wrapper<WD_PRESSURE, true, false, false, true, WET>(rhs, nds, h, gs, v2{0., 0.}, nullptr, nullptr, djac, 1, t1);
            } else if (pass == 1) {
                t1 = g * dt;
                wd_wrapper<WD_BODYFORCE, false, false, false, true, WET>(rhs, nds, h, nullptr, ggz, nullptr, nullptr, djac, 1, t1);
            } else {
                t1 = one_2 * dt * g;
                wd_wrapper<WD_BOUNDARY_PRESSURE, false, false, false, true, WET>(rhs, nds, h, nullptr, v2{0., 0.}, nullptr, nullptr,
                                                                            djac, 1, t1);
            }
            for (int i = 0; i < 3; i++) {   // dof3_add_replace_array
                elem_rhs[i].x_eq = elem_rhs[i].x_eq + rhs[i].x_eq;
                elem_rhs[i].y_eq = elem_rhs[i].y_eq + rhs[i].y_eq;
                elem_rhs[i].c_eq = elem_rhs[i].c_eq + rhs[i].c_eq;
            }
        }
    }

    // SUPG :987-1232
    if (fabs(h_wd) >= SMALL) {
        const double g_factor = (1 + elem_density_avg) * g;
        const double denom = dmax(h_wd, NOT_QUITE_SMALL);
        double tau;
        {   // fe_get_supg_tau_sw (src/fe/fe_get_supg_tau.c:43-80), le = sqrt(area*factor_old) (:289)
            double le = sqrt(area * factor_old);
            double cs2 = g_factor * denom;
            if (denom <= 0.0) cs2 = 0.;
            double sr = sqrt(vwd.x * vwd.x + vwd.y * vwd.y + cs2 + SMALL);
            tau = D.tau_pg * le / sr;
        }
        d3 dfdt[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
        d3 adv[3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
        double dudt = S.pre(19), dvdt = S.pre(20), dhdt = S.pre(21);   // -(old-state integrals), :1012-1013
        factor = wd_wrapper<WD_INTEGRATE_F, false, true, true, false, WET>(dfdt, nds, h, nullptr, v2{0., 0.}, vel, h, djac, 1, 1);
        dudt += dfdt[0].x_eq; dvdt += dfdt[0].y_eq; dhdt += dfdt[0].c_eq;
        factor = wd_wrapper<WD_GLS_CONVECTION, true, true, true, false, WET>(adv, nds, h, gs, v2{0., 0.}, vel, h, djac, dt, 1);
        double x_pressure = g_factor * (ggh.x + ggz.x) * dt * factor + ggp.x * dt * factor;
        double y_pressure = g_factor * (ggh.y + ggz.y) * dt * factor + ggp.y * dt * factor;
        double x_friction = g * bed_stress.x * dt * factor;
        double y_friction = g * bed_stress.y * dt * factor;
        double x_coriolis = -vwd.y * cor;
        double y_coriolis = vwd.x * cor;
        double c_eq = dhdt + adv[0].c_eq + factor * src * djac * dt;
        double x_eq = dudt + adv[0].x_eq + (x_pressure + x_friction / denom) * djac + (x_coriolis - total_stress.x / denom) * djac * factor * dt;
        double y_eq = dvdt + adv[0].y_eq + (y_pressure + y_friction / denom) * djac + (y_coriolis - total_stress.y / denom) * djac * factor * dt;
        for (int i = 0; i < 3; i++) {
            rcq[i] = (gs[i].x * (tau * vwd.x * c_eq + tau * h_wd * x_eq) + gs[i].y * (tau * vwd.y * c_eq + tau * h_wd * y_eq));
            rxq[i] = (gs[i].x * (tau * g_factor * h_wd * c_eq + tau * vwd.x * h_wd * x_eq) + gs[i].y * (tau * vwd.y * h_wd * x_eq));
            ryq[i] = (gs[i].y * (tau * g_factor * h_wd * c_eq + tau * vwd.y * h_wd * y_eq) + gs[i].x * (tau * vwd.x * h_wd * y_eq));
        }
        for (int i = 0; i < 3; i++) {
            elem_rhs[i].c_eq += rcq[i];
            elem_rhs[i].x_eq += rxq[i];
            elem_rhs[i].y_eq += ryq[i];
        }
    }

    // Dirichlet :1239-1256
    for (int i = 0; i < 3; i++) {
        int bc = S.node_bc(i);
        if (bc == BCT_PRS_DIR) {
            elem_rhs[i].c_eq = 0.0;
        } else if (bc == BCT_VEL_DIR) {
            elem_rhs[i].x_eq = 0.0;
            elem_rhs[i].y_eq = 0.0;
        } else if (bc == BCT_VEL_PRS_DIR) {
            elem_rhs[i].c_eq = 0.0;
            elem_rhs[i].x_eq = 0.0;
            elem_rhs[i].y_eq = 0.0;
        }
    }
}

// -------------------------------------------------------- boundary residual
// fe_sw2_boundary_resid (src/fe/sw2/fe_sw2_boundary_resid.c:72-510) for
// NB VEL, NB OTW and NB OUT edges.  dry_zero: velocities at nodes with
// old_head <= 0 read as zero (the 2D loop has already zeroed them).
HD void boundary_resid(const ElemData &D, int ie, double pert, int pnode, int pvar, int psign, bool dry_zero,
                       d3 *elem_rhs) {
    const int nd[2] = {D.e1d[2 * ie], D.e1d[2 * ie + 1]};
    const v2 nrml = {D.nrml1d[2 * ie], D.nrml1d[2 * ie + 1]};
    const double dt = edge_dt(D, ie), djac = D.djac1d[ie], g = D.g;
    int PRESSURE_FLAG = 1;
    if (pvar == PERTURB_NONE) psign = 0;
    double h[2];
    v2 vel[2];
    for (int i = 0; i < 2; i++) {
        int n = nd[i];
        h[i] = D.head[n];
        double un = D.u[n], vn = D.v[n];
        if (dry_zero && D.old_head[n] <= 0.) { un = 0.; vn = 0.; }
        vel[i].x = un; vel[i].y = vn;
    }
    if (pvar == PERTURB_U || pvar == PERTURB_V) PRESSURE_FLAG = 0;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        if (i != pnode) continue;
        if (pvar == PERTURB_H) h[i] += psign * pert;
        if (pvar == PERTURB_U) vel[i].x += psign * pert;
        if (pvar == PERTURB_V) vel[i].y += psign * pert;
    }
    double u[2] = {vel[0].x, vel[1].x}, v[2] = {vel[0].y, vel[1].y};
    double h_avg = one_2 * (h[0] + h[1]);
    if (h_avg < SMALL) h_avg = SMALL;
    const double z0 = D.z[nd[0]] + 0.0;
    for (int i = 0; i < 2; i++) { elem_rhs[i].x_eq = 0.; elem_rhs[i].y_eq = 0.; elem_rhs[i].c_eq = 0.; }
    const int bc = D.bc1d[ie];
    if (bc == BCT_VEL_NEU) {  // :212-219, explicit flow :562-604
        double flux = -1.0 * D.val1d[ie];
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
        for (int i = 0; i < 2; i++) {
            elem_rhs[i].c_eq += t1 * rc[i];
            elem_rhs[i].x_eq += t1 * rx[i];
            elem_rhs[i].y_eq += t1 * ry[i];
        }
    } else if (bc == BCT_OUTFLOW || bc == BCT_PRS_NEU) {
        double depth[2] = {0., 0.};
        if (bc == BCT_PRS_NEU) {  // :364-391; node 0's bed for both depths (:375-376)
            double tail = D.val1d[ie];
            depth[0] = tail - z0;
            depth[1] = tail - z0;
            if (depth[0] < SMALL || depth[1] < SMALL || h[0] < SMALL || h[1] < SMALL) return;
        }
        {   // implicit flow :637-687
            double rc[2] = {0., 0.}, rx[2] = {0., 0.}, ry[2] = {0., 0.};
            double u1 = vel[0].x, u2 = vel[1].x, v1 = vel[0].y, v2_ = vel[1].y, h1 = h[0], h2 = h[1];
            double ur1 = vel[0].x, ur2 = vel[1].x, vr1 = vel[0].y, vr2 = vel[1].y;
            double dir0 = vel[0].x * nrml.x + vel[0].y * nrml.y;
            double dir1 = vel[1].x * nrml.x + vel[1].y * nrml.y;
            if (dir0 < SMALL) { ur1 = 0.; vr1 = 0.; }
            if (dir1 < SMALL) { ur2 = 0.; vr2 = 0.; }
            int_line_phi_h_v_dot_n(djac, 1., h, vel, nrml, rc);
            int_line_phi_f_g_h(djac, nrml.x, ur1, ur2, u1, u2, h1, h2, rx);
            int_line_phi_f_g_h(djac, nrml.y, vr1, vr2, v1, v2_, h1, h2, ry);
            int_line_phi_f_g_h(djac, nrml.y, ur1, ur2, v1, v2_, h1, h2, rx);
            int_line_phi_f_g_h(djac, nrml.x, vr1, vr2, u1, u2, h1, h2, ry);
            for (int i = 0; i < 2; i++) {
                elem_rhs[i].c_eq += dt * rc[i];
                elem_rhs[i].x_eq += dt * rx[i];
                elem_rhs[i].y_eq += dt * ry[i];
            }
        }
        if (bc == BCT_PRS_NEU && PRESSURE_FLAG) {  // fe_sw2_1d_pressure :773-798, depth then head
            for (int k = 0; k < 2; k++) {
                const double *hh = k == 0 ? depth : h;
                double gg = k == 0 ? g : -g;
                double r[2] = {0., 0.};
                int_line_phi_h_h(djac, 1., hh[0], hh[1], r);
                double t1 = dt * one_2 * gg;
                for (int i = 0; i < 2; i++) {
                    elem_rhs[i].x_eq += t1 * nrml.x * r[i];
                    elem_rhs[i].y_eq += t1 * nrml.y * r[i];
                }
            }
        }
    }
    for (int i = 0; i < 2; i++) {  // :483-498
        int nbc = D.node_bc[nd[i]];
        if (nbc == BCT_PRS_DIR) elem_rhs[i].c_eq = 0.0;
        else if (nbc == BCT_VEL_DIR) { elem_rhs[i].x_eq = 0.0; elem_rhs[i].y_eq = 0.0; }
        else if (nbc == BCT_VEL_PRS_DIR) { elem_rhs[i].c_eq = 0.0; elem_rhs[i].x_eq = 0.0; elem_rhs[i].y_eq = 0.0; }
    }
}

// NUM_DIFF_EPSILON (src/include/macro.h:669-673)
HDF double num_diff_eps(double var) {
    const double perturbation = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON), exact
    double e = fabs(var) * (1. + perturbation);
    e -= fabs(var);
    if (e < perturbation) e = perturbation;
    return e;
}

// One column of an element Jacobian: the central difference of the element
// residual for the perturbation of variable `var` (0 = u, 1 = v, 2 = h) at
// local node `j` (perturb_sw2_head/vel, src/fe/sw2/fe_sw2_load.c:293-483 and
// elem_matrix_deriv, src/fe/fe_elem_utilities.c:518-525).  col[i*3 + r] is
// d(elem_rhs[i].eq_r)/d(var_j) with eq order x, y, c.
template <int NN, bool WET = false>
HD void jac_column(const ElemData &D, int ie, int j, int var, double *col) {
    int n = NN == 3 ? D.e2d[3 * ie + j] : D.e1d[2 * ie + j];
    double val = var == 0 ? D.u[n] : (var == 1 ? D.v[n] : D.head[n]);
    double eps = num_diff_eps(val), eps2 = 2.0 * eps;
    int pv = var == 0 ? PERTURB_U : (var == 1 ? PERTURB_V : PERTURB_H);
    d3 P[3], M[3];
    if (NN == 3) {
        const double *pre = D.pre + PRE_STRIDE * (size_t)ie;
        body_resid<WET>(D, pre, ie, eps, j, pv, +1, false, P);
        body_resid<WET>(D, pre, ie, eps, j, pv, -1, false, M);
    } else {
        boundary_resid(D, ie, eps, j, pv, +1, false, P);
        boundary_resid(D, ie, eps, j, pv, -1, false, M);
    }
    const double inv = 1.0 / eps2;
    for (int i = 0; i < NN; i++) {
        col[3 * i + 0] = div_rn(P[i].x_eq - M[i].x_eq, eps2, inv);
        col[3 * i + 1] = div_rn(P[i].y_eq - M[i].y_eq, eps2, inv);
        col[3 * i + 2] = div_rn(P[i].c_eq - M[i].c_eq, eps2, inv);
    }
}

// Dirichlet edits of fe_sw2_load.c:100-120 (fe_assign_mom_db_dof3 /
// fe_assign_dacont_db_dof3, src/fe/fe_assign_elem_mat_db.c:30-61, 196-219)
// applied to one column (node j, var) of the element matrix.  The reference
// reads the node string with the LOCAL index i (fe_sw2_load.c:103, :165), so
// the node-string flags come from nodes 0..NN-1 of the mesh; `loc_bc` holds them.
template <int NN>
HDF void jac_column_db(const ElemData &D, const int *nodes, const int *loc_bc, int j, int var, double *col) {
#pragma unroll
    for (int k = 0; k < NN; k++) {
        // mom: zero rows k (x,y) in all columns, zero columns (k,u),(k,v), unit diagonal
        auto mom = [&]() {
            col[3 * k + 0] = 0.;
            col[3 * k + 1] = 0.;
            if (j == k && var < 2) {
#pragma unroll
                for (int i = 0; i < NN; i++) { col[3 * i] = 0.; col[3 * i + 1] = 0.; col[3 * i + 2] = 0.; }
                if (var == 0) col[3 * k + 0] = 1.;
                else col[3 * k + 1] = 1.;
            }
        };
        // dacont: zero row (k,c) in all columns, zero column (k,h), unit diagonal
        auto dacont = [&]() {
            col[3 * k + 2] = 0.;
            if (j == k && var == 2) {
#pragma unroll
                for (int i = 0; i < NN; i++) { col[3 * i] = 0.; col[3 * i + 1] = 0.; col[3 * i + 2] = 0.; }
                col[3 * k + 2] = 1.;
            }
        };
        // the reference's order of edits for local node k
        if (D.old_head[nodes[k]] <= 0.) mom();
        const int nb = loc_bc[k];
        if (nb == BCT_PRS_DIR) dacont();
        else if (nb == BCT_VEL_DIR) mom();
        else if (nb == BCT_VEL_PRS_DIR) { dacont(); mom(); }
    }
}

}  // namespace sw2

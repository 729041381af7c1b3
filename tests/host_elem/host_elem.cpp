// Test-only: the library's element residual / Jacobian-column code compiled
// for the host, so the CPU suite can compare it with the oracle bit for bit
// before any GPU run.  Not part of the product library.
#include <cstring>
#include <vector>
#include "../../adhv8_kraken_amd/csrc/sw2_elem.h"

using namespace sw2;

static crlog::tables T;
static bool ready = false;

extern "C" {

struct HostElem {
    ElemData D;
    std::vector<double> pre;
};

void *he_create(const double *x, const double *y, const double *z, const double *head, const double *u,
                const double *v, const double *oh, const double *ou, const double *ov, const double *ooh,
                const double *oou, const double *oov, const int *node_bc, const int *first_elem, const int *e2d,
                const double *djac, const double *grad, const int *mat, const double *ks, const int *src_flag,
                const double *src, const int *wd_flag, const int *e1d, const double *djac1d, const double *nrml1d,
                const int *bc1d, const double *val1d, const int *mat_turb, const int *mat_eev_mode,
                const double *mat_eev_coef, const double *mat_ev, const double *mat_cor, double g, double dt,
                double alpha, double tau_pg, double dll, double dul, double visc, int nelems) {
    if (!ready) { crlog::make_tables(&T); ready = true; }
    HostElem *h = new HostElem();
    ElemData &D = h->D;
    D.x = x; D.y = y; D.z = z; D.head = head; D.u = u; D.v = v; D.old_head = oh; D.old_u = ou; D.old_v = ov;
    D.older_head = ooh; D.older_u = oou; D.older_v = oov; D.node_bc = node_bc; D.first_elem = first_elem;
    D.e2d = e2d; D.djac = djac; D.grad = grad; D.mat = mat; D.ks = ks; D.src_flag = src_flag; D.src = src;
    D.wd_flag = wd_flag; D.e1d = e1d; D.djac1d = djac1d; D.nrml1d = nrml1d; D.bc1d = bc1d; D.val1d = val1d;
    D.mat_turb = mat_turb; D.mat_eev_mode = mat_eev_mode; D.mat_eev_coef = mat_eev_coef; D.mat_ev = mat_ev;
    D.mat_cor = mat_cor; D.g = g; D.dt = dt; D.alpha = alpha; D.tau_pg = tau_pg; D.dll = dll; D.dul = dul;
    D.visc = visc; D.logt = &T;
    h->pre.assign(PRE_STRIDE * (size_t)nelems, 0.0);
    for (int ie = 0; ie < nelems; ie++) elem_old_terms(D, ie, h->pre.data() + PRE_STRIDE * (size_t)ie);
    D.pre = h->pre.data();
    return h;
}

void he_destroy(void *p) { delete (HostElem *)p; }

// div_rn over n pairs (quotients out), for the division test
void he_div_rn(int n, const double *x, const double *b, double *q) {
    for (int i = 0; i < n; i++) q[i] = div_rn(x[i], b[i], 1.0 / b[i]);
}

// winds: node values (x, y per node), attenuation per material, density
void he_set_wind(void *p, const double *wind, const double *mat_windatt, double density) {
    HostElem *h = (HostElem *)p;
    h->D.wind = wind;
    h->D.mat_windatt = mat_windatt;
    h->D.density = density;
}

void he_body(void *p, int ie, double pert, int pn, int pv, int ps, int dry_rule, double *out) {
    d3 r[3];
    const ElemData &D = ((HostElem *)p)->D;
    body_resid(D, D.pre + PRE_STRIDE * (size_t)ie, ie, pert, pn, pv, ps, dry_rule != 0, r);
    memcpy(out, r, sizeof(r));
}

// the fully wet variant of the element residual (valid when all depths > 0)
void he_body_wet(void *p, int ie, double pert, int pn, int pv, int ps, double *out) {
    d3 r[3];
    const ElemData &D = ((HostElem *)p)->D;
    body_resid<true>(D, D.pre + PRE_STRIDE * (size_t)ie, ie, pert, pn, pv, ps, false, r);
    memcpy(out, r, sizeof(r));
}

void he_boundary(void *p, int ie, double pert, int pn, int pv, int ps, int dry_zero, double *out) {
    d3 r[3];
    boundary_resid(((HostElem *)p)->D, ie, pert, pn, pv, ps, dry_zero != 0, r);
    memcpy(out, r, 2 * sizeof(d3));
}

// element matrix [81] in block order, Dirichlet edits applied
void he_elem_jac(void *p, int ie, int b0, int b1, int b2, double *out) {
    const ElemData &D = ((HostElem *)p)->D;
    const int nodes[3] = {D.e2d[3 * ie], D.e2d[3 * ie + 1], D.e2d[3 * ie + 2]};
    const int lb[3] = {b0, b1, b2};
    for (int j = 0; j < 3; j++)
        for (int var = 0; var < 3; var++) {
            double col[9];
            jac_column<3>(D, ie, j, var, col);
            jac_column_db<3>(D, nodes, lb, j, var, col);
            for (int i = 0; i < 3; i++)
                for (int r = 0; r < 3; r++) out[(3 * i + j) * 9 + 3 * r + var] = col[3 * i + r];
        }
}

double he_cr_log(double x) {
    if (!ready) { crlog::make_tables(&T); ready = true; }
    return crlog::cr_log_t(x, &T);
}
}

// C ABI of the SW2 Newton-step library (include/adh_sw2_hip.h): model
// set-up on the host (geometry, node/element adjacency, the reference's
// insertion-ordered sparsity pattern, assembly slot lists), device memory
// management, and the host side of fe_newton (src/fe/fe_newton.c:214-912),
// which only reads back three norms per Newton iteration and the BiCGSTAB
// flag every opts.check_every iterations.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sw2_ctx.h"

using namespace sw2;

static void drop_graphs(Ctx &c);

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(ADH_SW2_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct adh_sw2_hip {
    Ctx c;
    std::vector<void *> allocs;
    int *scratch;   // this handle's device's copy of `scratch` (adh_sw2_hip_init_noalloc), or null
};

// ------------------------------------------- staging area of init_noalloc
// adh_sw2_hip_init_noalloc stages the DOF boundary mask in this file-scope
// array instead of a per-call device allocation.  Every GPU the library runs
// on holds its own copy (24 MiB) until the process ends.  Host code never
// uses the array's name as an address: create_impl looks up the current
// device's copy once per device with hipGetSymbolAddress and the handle
// keeps that address.  Calls on one GPU share the copy, so overlapping calls
// on one GPU race; each call writes all 3N entries it reads back.
constexpr size_t NOALLOC_MAX_NODES = (size_t)1 << 21;   // 2,097,152 nodes
static __device__ __align__(256) int scratch[3 * NOALLOC_MAX_NODES];
static std::mutex g_scratch_mu;
static std::map<int, int *> g_scratch_addr;   // device ordinal -> address of its `scratch`

// Binds h to the current device's staging area (looked up on the first
// handle created on that device).
static int bind_scratch(adh_sw2_hip *h) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(dev);
    if (it == g_scratch_addr.end()) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ADH_SW2_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        it = g_scratch_addr.emplace(dev, (int *)p).first;
    }
    h->scratch = it->second;
    return 0;
}

extern "C" const char *adh_sw2_hip_last_error(void) { return g_err.c_str(); }

extern "C" void adh_sw2_hip_default_opts(adh_sw2_opts *o) {
    o->tol_nonlin = 1e-8;
    o->inc_nonlin = 1e-14;
    o->max_nonlin_it = 20;
    o->max_lin_it = 500;
    o->max_cuts = 5;
    o->lin_rtol = 1e-5;
    o->lin_atol = 1e-10;
    o->precision = 0;
    o->preconditioner = 2;
    o->check_every = 8;
    o->nblock = 1;
    o->graphs = 0;  // measured: no gain over eager launches on MI355X for single models
    o->krylov = 0;  // BiCGSTAB, as the reference
    o->lu_tiled = 0;
}

static int check_opts(const adh_sw2_opts *o) {
    if (o->precision < 0 || o->precision > 2) return fail(ADH_SW2_EINVAL, "precision must be 0, 1 or 2");
    if (o->preconditioner < 1 || o->preconditioner > 3)
        return fail(ADH_SW2_EUNSUPPORTED, "preconditioner must be 1 (profile LU), 2 (block Jacobi) or 3 (block ILU(0))");
    if (o->krylov != 0 && o->krylov != 1) return fail(ADH_SW2_EINVAL, "krylov must be 0 (BiCGSTAB) or 1 (GMRES)");
    if (o->krylov == 1 && o->precision == 2)
        return fail(ADH_SW2_EUNSUPPORTED, "GMRES runs with FP64 Krylov vectors (precision 0 or 1)");
    if (o->preconditioner == 3 && o->precision == 2)
        return fail(ADH_SW2_EUNSUPPORTED, "the ILU(0) preconditioner runs with FP64 Krylov vectors (precision 0 or 1)");
    if (o->preconditioner == 1 && o->precision != 0)
        return fail(ADH_SW2_EUNSUPPORTED, "the profile-LU preconditioner runs in FP64 (precision 0)");
    if (o->preconditioner == 1 && o->nblock < 1) return fail(ADH_SW2_EINVAL, "nblock must be >= 1");
    if (o->lu_tiled != 0 && o->lu_tiled != 1) return fail(ADH_SW2_EINVAL, "lu_tiled must be 0 or 1");
    return 0;
}

static crlog::tables g_host_tables;
static bool g_host_tables_ready = false;

extern "C" double adh_sw2_hip_cr_log(double x) {
    if (!g_host_tables_ready) {
        crlog::make_tables(&g_host_tables);
        g_host_tables_ready = true;
    }
    return crlog::cr_log_t(x, &g_host_tables);
}

template <typename T>
static int dalloc(adh_sw2_hip *h, T **p, size_t n) {
    void *q = nullptr;
    size_t bytes = sizeof(T) * (n ? n : 1);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(ADH_SW2_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = (T *)q;
    return 0;
}

template <typename T>
static int dupload(adh_sw2_hip *h, T **p, const T *src, size_t n) {
    int rc = dalloc(h, p, n);
    if (rc) return rc;
    if (n && src) HIPCHK(hipMemcpy(*p, src, sizeof(T) * n, hipMemcpyHostToDevice));
    else if (n) HIPCHK(hipMemset(*p, 0, sizeof(T) * n));
    return 0;
}

// Triangle djac and shape gradients, src/elem/element_djac_nrml_gradPhi.c:193-274.
static void tri_geometry(double x1, double x2, double x3, double y1, double y2, double y3, double *djac, double *g) {
    double s1x = x2 - x1, s1y = y2 - y1;
    double s2x = x3 - x1, s2y = y3 - y1;
    double dj = (s1x * s2y) - (s2x * s1y);
    if (fabs(dj) > SMALL) {
        double inv = 1. / dj;
        g[0] = (s1y - s2y) * inv;
        g[1] = (s2x - s1x) * inv;
        g[2] = s2y * inv;
        g[3] = -s2x * inv;
        g[4] = -s1y * inv;
        g[5] = s1x * inv;
    } else {
        for (int k = 0; k < 6; k++) g[k] = -9999999.9;
    }
    *djac = dj * (1.0 / 2.0);
}

// Edge length and outward normal, element_djac_nrml_gradPhi.c:24-36 and
// src/structs/selem_1d.c:118-160.
static void edge_geometry(const adh_sw2_mesh *m, int ie, double *djac, double *nrml) {
    int n1 = m->e1d[2 * ie], n2 = m->e1d[2 * ie + 1];
    double x1 = m->x[n1], x2 = m->x[n2], y1 = m->y[n1], y2 = m->y[n2];
    *djac = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
    double delx = x2 - x1, dely = y2 - y1;
    double dist = sqrt(delx * delx + dely * dely);
    nrml[0] = dely / dist;
    nrml[1] = -delx / dist;
    const int *t = m->e2d + 3 * m->e1d_adj[ie];
    double cx = one_3 * (m->x[t[0]] + m->x[t[1]] + m->x[t[2]]);
    double cy = one_3 * (m->y[t[0]] + m->y[t[1]] + m->y[t[2]]);
    delx = cx - x1;
    dely = cy - y1;
    dist = delx * nrml[0] + dely * nrml[1];
    if (dist > 0.) {
        nrml[0] *= -1.;
        nrml[1] *= -1.;
    }
}

// Element geometry as given by the caller or computed from the coordinates.
static void mesh_geometry(const adh_sw2_mesh *m, std::vector<double> &djac, std::vector<double> &grad,
                          std::vector<double> &djac1d, std::vector<double> &nrml1d) {
    const int E = m->nelems2d, B = m->nelems1d;
    djac.assign(E, 0.0); grad.assign(6 * (size_t)E, 0.0); djac1d.assign(B, 0.0); nrml1d.assign(2 * (size_t)B, 0.0);
    if (m->e2d_djac && m->e2d_grad) {
        memcpy(djac.data(), m->e2d_djac, sizeof(double) * E);
        memcpy(grad.data(), m->e2d_grad, sizeof(double) * 6 * E);
    } else {
        for (int i = 0; i < E; i++) {
            const int *t = m->e2d + 3 * i;
            tri_geometry(m->x[t[0]], m->x[t[1]], m->x[t[2]], m->y[t[0]], m->y[t[1]], m->y[t[2]], &djac[i],
                         &grad[6 * (size_t)i]);
        }
    }
    for (int i = 0; i < B; i++) {
        if (m->e1d_djac && m->e1d_nrml) {
            djac1d[i] = m->e1d_djac[i];
            nrml1d[2 * i] = m->e1d_nrml[2 * i];
            nrml1d[2 * i + 1] = m->e1d_nrml[2 * i + 1];
        } else {
            edge_geometry(m, i, &djac1d[i], &nrml1d[2 * (size_t)i]);
        }
    }
}

static int validate(const adh_sw2_mesh *m) {
    if (!m || m->nnodes <= 0 || m->nelems2d <= 0 || m->nmat <= 0 || m->nelems1d < 0)
        return fail(ADH_SW2_EINVAL, "mesh sizes must be positive");
    if (m->nelems2d > MAX_TRIANGLES) return fail(ADH_SW2_EINVAL, "more than 115,000,000 triangles per device");
    if (!m->x || !m->y || !m->z || !m->e2d || !m->e2d_mat || !m->e2d_ks || !m->mat_turb || !m->mat_ev ||
        !m->mat_cor_speed)
        return fail(ADH_SW2_EINVAL, "missing required mesh array");
    if (m->nelems1d && (!m->e1d || !m->e1d_bc || !m->e1d_val || (!m->e1d_adj && !(m->e1d_djac && m->e1d_nrml))))
        return fail(ADH_SW2_EINVAL, "missing boundary-edge array");
    for (int i = 0; i < 3 * m->nelems2d; i++)
        if (m->e2d[i] < 0 || m->e2d[i] >= m->nnodes) return fail(ADH_SW2_EINVAL, "triangle node id out of range");
    for (int i = 0; i < m->nelems2d; i++)
        if (m->e2d_mat[i] < 0 || m->e2d_mat[i] >= m->nmat) return fail(ADH_SW2_EINVAL, "element material out of range");
    for (int i = 0; i < 2 * m->nelems1d; i++)
        if (m->e1d[i] < 0 || m->e1d[i] >= m->nnodes) return fail(ADH_SW2_EINVAL, "edge node id out of range");
    for (int i = 0; i < m->nelems1d; i++) {
        int b = m->e1d_bc[i];
        if (b != BCT_VEL_NEU && b != BCT_PRS_NEU && b != BCT_OUTFLOW)
            return fail(ADH_SW2_EUNSUPPORTED, "boundary edge type not supported (NB VEL/OTW/OUT only)");
        if (m->e1d_adj && (m->e1d_adj[i] < 0 || m->e1d_adj[i] >= m->nelems2d))
            return fail(ADH_SW2_EINVAL, "edge adjacent element out of range");
    }
    for (int i = 0; i < m->nmat; i++) {
        int t = m->mat_turb[i];
        if (t != 0 && t != TURB_EVS && t != TURB_EEV) return fail(ADH_SW2_EINVAL, "bad turbulence closure");
        if (t == TURB_EEV && (!m->mat_eev_mode || m->mat_eev_mode[i] < 1 || m->mat_eev_mode[i] > 3))
            return fail(ADH_SW2_EINVAL, "EEV mode must be 1, 2 or 3");
    }
    if (m->wind) {
        for (int i = 0; i < m->nmat; i++) {
            if (m->mat_wind_flag && m->mat_wind_flag[i] != 0)
                return fail(ADH_SW2_EUNSUPPORTED, "wind stress law not supported (MP WND STR 0 only)");
            if (m->mat_windatt && (m->mat_windatt[i] < 0.0 || m->mat_windatt[i] > 1.0))
                return fail(ADH_SW2_EINVAL, "wind attenuation must lie in [0, 1]");
        }
    }
    if (m->node_bc) {
        for (int i = 0; i < m->nnodes; i++) {
            int b = m->node_bc[i];
            if (b != 0 && b != BCT_PRS_DIR && b != BCT_VEL_DIR && b != BCT_VEL_PRS_DIR)
                return fail(ADH_SW2_EUNSUPPORTED, "node string type not supported (Dirichlet only)");
            if (b && !m->node_bc_val) return fail(ADH_SW2_EINVAL, "Dirichlet node without values");
        }
    }
    return 0;
}

static int alloc_solver(adh_sw2_hip *h) {
    Ctx &c = h->c;
    int rc;
    size_t n3 = 3 * (size_t)c.Np;
    if ((rc = dalloc(h, &c.kb, n3)) || (rc = dalloc(h, &c.kr, n3)) || (rc = dalloc(h, &c.kp, n3)) ||
        (rc = dalloc(h, &c.kAp, n3)) || (rc = dalloc(h, &c.ks_, n3)) || (rc = dalloc(h, &c.kAs, n3)) ||
        (rc = dalloc(h, &c.kq, n3)) || (rc = dalloc(h, &c.ku, n3)))
        return rc;
    if ((rc = dalloc(h, &c.fb, n3)) || (rc = dalloc(h, &c.fr, n3)) || (rc = dalloc(h, &c.fp, n3)) ||
        (rc = dalloc(h, &c.fAp, n3)) || (rc = dalloc(h, &c.fs, n3)) || (rc = dalloc(h, &c.fAs, n3)) ||
        (rc = dalloc(h, &c.fq, n3)) || (rc = dalloc(h, &c.fu, n3)))
        return rc;
    if ((rc = dalloc(h, &c.diag32, 9 * (size_t)c.Np)) || (rc = dalloc(h, &c.offd32, 9 * c.nsell)) ||
        (rc = dalloc(h, &c.dinv32, 9 * (size_t)c.Np)))
        return rc;
    for (double *v : {c.kb, c.kr, c.kp, c.kAp, c.ks_, c.kAs, c.kq, c.ku}) HIPCHK(hipMemset(v, 0, sizeof(double) * n3));
    for (float *v : {c.fb, c.fr, c.fp, c.fAp, c.fs, c.fAs, c.fq, c.fu}) HIPCHK(hipMemset(v, 0, sizeof(float) * n3));
    // SELL padding blocks are read by the SpMV and must stay zero; set-up writes only real blocks
    HIPCHK(hipMemset(c.offd32, 0, sizeof(float) * 9 * (c.nsell ? c.nsell : 1)));
    HIPCHK(hipMemset(c.diag32, 0, sizeof(float) * 9 * (size_t)c.Np));
    HIPCHK(hipMemset(c.dinv32, 0, sizeof(float) * 9 * (size_t)c.Np));
    return 0;
}

// ninst > 0: the mesh is the union of an ensemble of ninst models (see
// adh_sw2_hip_batch_create); the per-model scalars are sized for them.
static int create_impl(const adh_sw2_mesh *m, const adh_sw2_opts *opts, adh_sw2_hip **out, int ninst) {
    if (!out) return fail(ADH_SW2_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = validate(m);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(ADH_SW2_EHIP, "no HIP device available: the SW2 library runs on the GPU only");
    adh_sw2_hip *h = new adh_sw2_hip();  // value-initialised: all device pointers start null
    Ctx &c = h->c;
    c.N = m->nnodes; c.E = m->nelems2d; c.B = m->nelems1d; c.M = m->nmat;
    c.Np = (c.N + 63) & ~63;
    c.kgraph_prec = -1;
    if (opts) {
        int rc0 = check_opts(opts);
        if (rc0) { delete h; return rc0; }
        c.opts = *opts;
    } else adh_sw2_hip_default_opts(&c.opts);
    if (c.opts.check_every < 1) c.opts.check_every = 1;
    c.dt = m->dt;
    const int N = c.N, E = c.E, B = c.B;

    // ---- geometry (host)
    std::vector<double> djac, grad, djac1d, nrml1d;
    mesh_geometry(m, djac, grad, djac1d, nrml1d);

    // ---- adjacency (ascending element order)
    std::vector<int> n2e_ptr(N + 1, 0), n2b_ptr(N + 1, 0);
    for (int i = 0; i < 3 * E; i++) n2e_ptr[m->e2d[i] + 1]++;
    for (int i = 0; i < 2 * B; i++) n2b_ptr[m->e1d[i] + 1]++;
    for (int n = 0; n < N; n++) { n2e_ptr[n + 1] += n2e_ptr[n]; n2b_ptr[n + 1] += n2b_ptr[n]; }
    std::vector<int> n2e(3 * (size_t)E), n2b(2 * (size_t)B), fill(N);
    for (int n = 0; n < N; n++) fill[n] = n2e_ptr[n];
    for (int e = 0; e < E; e++)
        for (int i = 0; i < 3; i++) n2e[fill[m->e2d[3 * e + i]]++] = 3 * e + i;
    for (int n = 0; n < N; n++) fill[n] = n2b_ptr[n];
    for (int e = 0; e < B; e++)
        for (int i = 0; i < 2; i++) n2b[fill[m->e1d[2 * e + i]]++] = 2 * e + i;
    std::vector<int> first_elem(N, -1);
    for (int n = 0; n < N; n++)
        if (n2e_ptr[n + 1] > n2e_ptr[n]) first_elem[n] = n2e[n2e_ptr[n]] / 3;

    // ---- sparsity pattern in spv_load insertion order (2D elements, then 1D)
    std::vector<std::vector<int>> rows(N);
    auto insert = [&](int a, int b) {
        if (a == b) return;
        std::vector<int> &r = rows[a];
        for (int v : r) if (v == b) return;
        r.push_back(b);
    };
    for (int e = 0; e < E; e++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) insert(m->e2d[3 * e + i], m->e2d[3 * e + j]);
    for (int e = 0; e < B; e++)
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) insert(m->e1d[2 * e + i], m->e1d[2 * e + j]);
    c.h_row_ptr.assign(N + 1, 0);
    for (int n = 0; n < N; n++) c.h_row_ptr[n + 1] = c.h_row_ptr[n] + (int)rows[n].size();
    c.nnz = c.h_row_ptr[N];
    c.h_col.resize(c.nnz);
    for (int n = 0; n < N; n++)
        for (size_t k = 0; k < rows[n].size(); k++) c.h_col[c.h_row_ptr[n] + k] = rows[n][k];
    // slot of (a, b): diagonal a, or N + position of b in row a
    auto slot_of = [&](int a, int b) -> int {
        if (a == b) return a;
        const std::vector<int> &r = rows[a];
        for (size_t k = 0; k < r.size(); k++) if (r[k] == b) return N + c.h_row_ptr[a] + (int)k;
        return -1;
    };
    // SELL-64 slots for the off-diagonal blocks (see sw2_ctx.h)
    c.nslice = (N + 63) / 64;
    std::vector<int> sell_ptr(c.nslice + 1, 0);
    for (int s = 0; s < c.nslice; s++) {
        int kmax = 0;
        for (int n = 64 * s; n < N && n < 64 * s + 64; n++) {
            int len = c.h_row_ptr[n + 1] - c.h_row_ptr[n];
            if (len > kmax) kmax = len;
        }
        sell_ptr[s + 1] = sell_ptr[s] + kmax;
    }
    {   // pad every slice to the longest row when that costs little: the
        // SpMV then runs a fully unrolled, branch-free slot loop
        int kglob = 0;
        for (int s = 0; s < c.nslice; s++) kglob = std::max(kglob, sell_ptr[s + 1] - sell_ptr[s]);
        c.kmax_row = kglob;
        c.kuni = 0;
        if (kglob >= 1 && kglob <= KUNI_MAX &&
            (double)kglob * c.nslice <= 1.25 * (double)sell_ptr[c.nslice]) {
            c.kuni = kglob;
            for (int s = 0; s <= c.nslice; s++) sell_ptr[s] = s * kglob;
        }
    }
    c.nsell = 64 * (size_t)sell_ptr[c.nslice];
    std::vector<int> sell_col(c.nsell > 0 ? c.nsell : 1);
    c.h_offd_pos.assign(c.nnz > 0 ? c.nnz : 1, 0);
    for (int s = 0; s < c.nslice; s++)
        for (int j = 0; j < sell_ptr[s + 1] - sell_ptr[s]; j++)
            for (int lane = 0; lane < 64; lane++) {
                int n = 64 * s + lane;
                size_t q = 64 * (size_t)(sell_ptr[s] + j) + lane;
                if (n < N && j < c.h_row_ptr[n + 1] - c.h_row_ptr[n]) {
                    int p = c.h_row_ptr[n] + j;
                    sell_col[q] = c.h_col[p];
                    c.h_offd_pos[p] = (int)q;
                } else {
                    sell_col[q] = n < N ? n : N - 1;
                }
            }
    const int nslots = N + c.nnz;
    std::vector<int> slot_ptr(nslots + 1, 0);
    for (int e = 0; e < E; e++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) slot_ptr[slot_of(m->e2d[3 * e + i], m->e2d[3 * e + j]) + 1]++;
    for (int e = 0; e < B; e++)
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) slot_ptr[slot_of(m->e1d[2 * e + i], m->e1d[2 * e + j]) + 1]++;
    for (int s = 0; s < nslots; s++) slot_ptr[s + 1] += slot_ptr[s];
    std::vector<int> slot_src(slot_ptr[nslots]), sfill(slot_ptr.begin(), slot_ptr.end() - 1);
    for (int e = 0; e < E; e++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                slot_src[sfill[slot_of(m->e2d[3 * e + i], m->e2d[3 * e + j])]++] = 9 * e + 3 * i + j;
    for (int e = 0; e < B; e++)
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                slot_src[sfill[slot_of(m->e1d[2 * e + i], m->e1d[2 * e + j])]++] = 9 * E + 4 * e + 2 * i + j;
    rows.clear();
    rows.shrink_to_fit();
    {   // device slot order: the N diagonal slots by node, then one slot per SELL
        // position (empty for padding), so the assembly's thread for node n finds
        // its row's blocks at N + q(n, j) and every run of 64 lanes reads and writes
        // contiguously (a slot's contributors keep their order: sums unchanged)
        std::vector<int> qinv(c.nsell > 0 ? c.nsell : 1, -1);
        for (int p = 0; p < c.nnz; p++) qinv[c.h_offd_pos[p]] = p;
        const size_t ndslots = (size_t)N + c.nsell;
        std::vector<int> dptr(ndslots + 1, 0), dsrc(slot_src.size());
        size_t w = 0;
        for (size_t ds = 0; ds < ndslots; ds++) {
            const int ps = ds < (size_t)N ? (int)ds : (qinv[ds - N] >= 0 ? N + qinv[ds - N] : -1);
            if (ps >= 0)
                for (int k = slot_ptr[ps]; k < slot_ptr[ps + 1]; k++) dsrc[w++] = slot_src[k];
            dptr[ds + 1] = (int)w;
        }
        slot_ptr.swap(dptr);
        slot_src.swap(dsrc);
        c.ndslots = ndslots;
    }

    for (int k = 0; k < 3; k++) c.loc_bc[k] = (m->node_bc && k < N) ? m->node_bc[k] : 0;

    // ---- device arrays
    hipError_t he = hipStreamCreate(&c.stream);
    if (he != hipSuccess) { delete h; return fail(ADH_SW2_EHIP, "hipStreamCreate failed"); }
    for (int i = 0; i < 8; i++) (void)hipEventCreate(&c.ev[i]);
    std::vector<int> zeros_i(N, 0);
    std::vector<double> zeros_d;
#define UP(dst, src, n)                                    \
    if ((rc = dupload(h, &c.dst, src, (size_t)(n)))) {     \
        adh_sw2_hip_destroy(h);                            \
        return rc;                                         \
    }
    UP(x, m->x, N); UP(y, m->y, N); UP(z, m->z, N);
    UP(head, (const double *)nullptr, N); UP(u, (const double *)nullptr, N); UP(v, (const double *)nullptr, N);
    UP(old_head, (const double *)nullptr, N); UP(old_u, (const double *)nullptr, N); UP(old_v, (const double *)nullptr, N);
    UP(older_head, (const double *)nullptr, N); UP(older_u, (const double *)nullptr, N);
    UP(older_v, (const double *)nullptr, N);
    UP(node_bc, m->node_bc ? m->node_bc : zeros_i.data(), N);
    UP(node_bc_val, m->node_bc_val, 3 * (size_t)N);
    UP(first_elem, first_elem.data(), N);
    UP(wd_tmp, (const int *)nullptr, N);
    UP(e2d, m->e2d, 3 * (size_t)E);
    UP(mat, m->e2d_mat, E);
    UP(ks, m->e2d_ks, E);
    std::vector<int> srcf(E, 0);
    if (m->e2d_src_flag) memcpy(srcf.data(), m->e2d_src_flag, sizeof(int) * E);
    UP(src_flag, srcf.data(), E);
    UP(src, m->e2d_src, E);
    UP(wd_flag, (const int *)nullptr, E);
    UP(djac, djac.data(), E);
    UP(grad, grad.data(), 6 * (size_t)E);
    UP(e1d, m->e1d, 2 * (size_t)B);
    UP(bc1d, m->e1d_bc, B);
    UP(val1d, m->e1d_val, B);
    UP(djac1d, djac1d.data(), B);
    UP(nrml1d, nrml1d.data(), 2 * (size_t)B);
    UP(mat_turb, m->mat_turb, c.M);
    std::vector<int> eevm(c.M, 0);
    std::vector<double> eevc(c.M, 0.0);
    if (m->mat_eev_mode) memcpy(eevm.data(), m->mat_eev_mode, sizeof(int) * c.M);
    if (m->mat_eev_coef) memcpy(eevc.data(), m->mat_eev_coef, sizeof(double) * c.M);
    UP(mat_eev_mode, eevm.data(), c.M);
    UP(mat_eev_coef, eevc.data(), c.M);
    UP(mat_ev, m->mat_ev, 3 * (size_t)c.M);
    UP(mat_cor, m->mat_cor_speed, c.M);
    if (m->wind) {
        std::vector<double> watt(c.M, 1.0);
        if (m->mat_windatt)
            for (int k = 0; k < c.M; k++) watt[k] = m->mat_windatt[k];
        UP(mat_windatt, watt.data(), c.M);
        UP(wind, (const double *)nullptr, 2 * (size_t)c.N);
    }
    {
        crlog::tables T;
        crlog::make_tables(&T);
        UP(logt, &T, 1);
    }
    UP(n2e_ptr, n2e_ptr.data(), N + 1);
    UP(n2e, n2e.data(), n2e.size());
    UP(n2b_ptr, n2b_ptr.data(), N + 1);
    UP(n2b, n2b.data(), n2b.size());
    UP(row_ptr, c.h_row_ptr.data(), N + 1);
    UP(col, c.h_col.data(), c.nnz);
    UP(sell_ptr, sell_ptr.data(), c.nslice + 1);
    UP(sell_col, sell_col.data(), c.nsell);
    UP(slot_ptr, slot_ptr.data(), c.ndslots + 1);
    UP(slot_src, slot_src.data(), slot_src.size());
    {   // element block -> its position in the slot-ordered contributor list
        c.h_blkpos.assign(slot_src.size(), 0);
        for (size_t p = 0; p < slot_src.size(); p++) c.h_blkpos[slot_src[p]] = (int)p;
        UP(blkpos, c.h_blkpos.data(), c.h_blkpos.size());
    }
    UP(pre, (const double *)nullptr, PRE_STRIDE * (size_t)E);
    UP(elist_wet, (const int *)nullptr, E);
    UP(elist_gen, (const int *)nullptr, E);
    UP(ecount, (const int *)nullptr, 2);
    UP(esplit, (const int *)nullptr, 2 * (size_t)nblocks_for(E, 256) + 2);
    c.pre_valid = 0;
    UP(ER, (const double *)nullptr, 9 * (size_t)E + 6 * (size_t)B);
    UP(EM, (const double *)nullptr, 81 * (size_t)E + 36 * (size_t)B);
    UP(R, (const double *)nullptr, 3 * (size_t)c.Np);
    UP(sol, (const double *)nullptr, 3 * (size_t)c.Np);
    UP(diag, (const double *)nullptr, 9 * (size_t)c.Np);
    UP(offd, (const double *)nullptr, 9 * c.nsell);
    UP(scale, (const double *)nullptr, 3 * (size_t)c.Np);
    UP(dinv, (const double *)nullptr, 9 * (size_t)c.Np);
    const size_t nmod = ninst > 0 ? (size_t)ninst : 1;
    UP(ksc, (const KScal *)nullptr, nmod);
    UP(norms, (const Norms *)nullptr, nmod);
    c.red.maxblocks = nblocks_for(N, 256) + 8;
    UP(red.part, (const double *)nullptr, 4 * (size_t)c.red.maxblocks);
    UP(red.count, (const unsigned *)nullptr, 8);
#undef UP
    if ((rc = alloc_solver(h))) { adh_sw2_hip_destroy(h); return rc; }
    if (hipHostMalloc((void **)&c.h_ksc, sizeof(KScal) * nmod) != hipSuccess ||
        hipHostMalloc((void **)&c.h_norms, sizeof(Norms) * nmod) != hipSuccess ||
        hipHostMalloc((void **)&c.h_flag, 64) != hipSuccess) {
        adh_sw2_hip_destroy(h);
        return fail(ADH_SW2_ENOMEM, "hipHostMalloc failed");
    }
    {   // scale_vect starts at 1.0 (fe_main.c:263)
        std::vector<double> ones(3 * (size_t)c.Np, 1.0);
        HIPCHK(hipMemcpy(c.scale, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice));
    }
    ElemData &D = c.ed;
    D.x = c.x; D.y = c.y; D.z = c.z;
    D.head = c.head; D.u = c.u; D.v = c.v;
    D.old_head = c.old_head; D.old_u = c.old_u; D.old_v = c.old_v;
    D.older_head = c.older_head; D.older_u = c.older_u; D.older_v = c.older_v;
    D.node_bc = c.node_bc; D.first_elem = c.first_elem;
    D.e2d = c.e2d; D.djac = c.djac; D.grad = c.grad; D.mat = c.mat; D.ks = c.ks; D.src_flag = c.src_flag;
    D.src = c.src; D.wd_flag = c.wd_flag; D.pre = c.pre;
    D.e1d = c.e1d; D.djac1d = c.djac1d; D.nrml1d = c.nrml1d; D.bc1d = c.bc1d; D.val1d = c.val1d;
    D.mat_turb = c.mat_turb; D.mat_eev_mode = c.mat_eev_mode; D.mat_eev_coef = c.mat_eev_coef;
    D.mat_ev = c.mat_ev; D.mat_cor = c.mat_cor;
    D.g = m->gravity; D.dt = m->dt; D.alpha = m->tau_temporal; D.tau_pg = m->tau_pg;
    D.dll = m->drying_lower_limit; D.dul = m->drying_upper_limit; D.visc = m->viscosity;
    D.logt = c.logt;
    D.wind = m->wind ? c.wind : nullptr;
    D.mat_windatt = c.mat_windatt;
    D.density = m->density;
    HIPCHK(hipDeviceSynchronize());
    if ((rc = bind_scratch(h))) {
        adh_sw2_hip_destroy(h);
        return rc;
    }
    *out = h;
    return ADH_SW2_OK;
}

extern "C" int adh_sw2_hip_create(const adh_sw2_mesh *m, const adh_sw2_opts *opts, adh_sw2_hip **out) {
    return create_impl(m, opts, out, 0);
}

extern "C" void adh_sw2_hip_destroy(adh_sw2_hip *h) {
    if (!h) return;
    Ctx &c = h->c;
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    drop_graphs(c);
    band_free(&c);
    ilu_free(&c);
    gmres_free(&c);
    for (void *p : h->allocs) (void)hipFree(p);
    if (c.h_ksc) (void)hipHostFree(c.h_ksc);
    if (c.h_norms) (void)hipHostFree(c.h_norms);
    if (c.h_flag) (void)hipHostFree(c.h_flag);
    for (int i = 0; i < 8; i++) if (c.ev[i]) (void)hipEventDestroy(c.ev[i]);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete h;
}

// what an ensemble handle runs: BiCGSTAB with 3x3 block Jacobi or the
// profile LU of each model's node blocks (tiled band LU)
static int check_batch_opts(const adh_sw2_opts *o) {
    if (o->preconditioner != 1 && o->preconditioner != 2)
        return fail(ADH_SW2_EUNSUPPORTED, "ensembles use the 3x3 block-Jacobi preconditioner or the profile LU");
    if (o->krylov != 0) return fail(ADH_SW2_EUNSUPPORTED, "ensembles use BiCGSTAB");
    return 0;
}

extern "C" int adh_sw2_hip_set_opts(adh_sw2_hip *h, const adh_sw2_opts *o) {
    if (!h || !o) return fail(ADH_SW2_EINVAL, "null argument");
    int rc = check_opts(o);
    if (rc) return rc;
    if (h->c.ninst && (rc = check_batch_opts(o))) return rc;
    h->c.opts = *o;
    if (h->c.opts.check_every < 1) h->c.opts.check_every = 1;
    return 0;
}

extern "C" int adh_sw2_hip_set_dt(adh_sw2_hip *h, double dt) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !(dt > 0)) return fail(ADH_SW2_EINVAL, "dt must be positive");
    h->c.dt = dt;
    h->c.ed.dt = dt;
    return 0;
}

extern "C" int adh_sw2_hip_set_wind(adh_sw2_hip *h, const double *stress_xy) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !stress_xy) return fail(ADH_SW2_EINVAL, "null argument");
    if (!h->c.ed.wind) return fail(ADH_SW2_EINVAL, "the model was created without winds");
    Ctx &c = h->c;
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipMemcpy(c.wind, stress_xy, sizeof(double) * 2 * (size_t)c.N, hipMemcpyHostToDevice));
    return 0;
}

// Series values as the reference averages them over [t, t + dt] before each
// attempt (sseries_setall_ts_valuesAVG, src/main/adh_func.c:542) and reads them
// in the residuals (sseries_get_value: fe_sw2_boundary_resid.c:212-214, 372;
// fe_sw2_body_resid.c:728-754).  Both kernels read them on every call, so the
// upload is all that changes.
static int upload_series(Ctx &c, int e0, int ne, int b0, int nb, const double *e1d_val, const double *e2d_src) {
    HIPCHK(hipStreamSynchronize(c.stream));
    if (e1d_val && nb) HIPCHK(hipMemcpy(c.val1d + b0, e1d_val, sizeof(double) * nb, hipMemcpyHostToDevice));
    if (e2d_src && ne) HIPCHK(hipMemcpy(c.src + e0, e2d_src, sizeof(double) * ne, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int adh_sw2_hip_set_series(adh_sw2_hip *h, const double *e1d_val, const double *e2d_src) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    return upload_series(h->c, 0, h->c.E, 0, h->c.B, e1d_val, e2d_src);
}

extern "C" int adh_sw2_hip_geometry(adh_sw2_hip *h, double *djac, double *grad, double *djac1d, double *nrml1d) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    if (djac) HIPCHK(hipMemcpy(djac, c.djac, sizeof(double) * c.E, hipMemcpyDeviceToHost));
    if (grad) HIPCHK(hipMemcpy(grad, c.grad, sizeof(double) * 6 * c.E, hipMemcpyDeviceToHost));
    if (djac1d && c.B) HIPCHK(hipMemcpy(djac1d, c.djac1d, sizeof(double) * c.B, hipMemcpyDeviceToHost));
    if (nrml1d && c.B) HIPCHK(hipMemcpy(nrml1d, c.nrml1d, sizeof(double) * 2 * c.B, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int adh_sw2_hip_pattern_nnz(adh_sw2_hip *h) { return h ? h->c.nnz : ADH_SW2_EINVAL; }

extern "C" int adh_sw2_hip_pattern(adh_sw2_hip *h, int *row_ptr, int *col) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    if (row_ptr) memcpy(row_ptr, h->c.h_row_ptr.data(), sizeof(int) * (h->c.N + 1));
    if (col && h->c.nnz) memcpy(col, h->c.h_col.data(), sizeof(int) * h->c.nnz);
    return 0;
}

static int set_pair(Ctx &c, double *dh, double *du, double *dv, const double *hh, const double *vel) {
    std::vector<double> u(c.N), v(c.N);
    for (int i = 0; i < c.N; i++) { u[i] = vel[2 * i]; v[i] = vel[2 * i + 1]; }
    HIPCHK(hipMemcpy(dh, hh, sizeof(double) * c.N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(du, u.data(), sizeof(double) * c.N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dv, v.data(), sizeof(double) * c.N, hipMemcpyHostToDevice));
    return 0;
}

static int get_pair(Ctx &c, const double *dh, const double *du, const double *dv, double *hh, double *vel) {
    if (hh) HIPCHK(hipMemcpy(hh, dh, sizeof(double) * c.N, hipMemcpyDeviceToHost));
    if (vel) {
        std::vector<double> u(c.N), v(c.N);
        HIPCHK(hipMemcpy(u.data(), du, sizeof(double) * c.N, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(v.data(), dv, sizeof(double) * c.N, hipMemcpyDeviceToHost));
        for (int i = 0; i < c.N; i++) { vel[2 * i] = u[i]; vel[2 * i + 1] = v[i]; }
    }
    return 0;
}

extern "C" int adh_sw2_hip_set_state(adh_sw2_hip *h, const double *head, const double *vel, const double *old_head,
                                     const double *old_vel, const double *older_head, const double *older_vel) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !head || !vel || !old_head || !old_vel || !older_head || !older_vel)
        return fail(ADH_SW2_EINVAL, "null argument");
    Ctx &c = h->c;
    HIPCHK(hipStreamSynchronize(c.stream));
    int rc;
    if ((rc = set_pair(c, c.head, c.u, c.v, head, vel))) return rc;
    if ((rc = set_pair(c, c.old_head, c.old_u, c.old_v, old_head, old_vel))) return rc;
    if ((rc = set_pair(c, c.older_head, c.older_u, c.older_v, older_head, older_vel))) return rc;
    c.pre_valid = 0;
    c.split_valid = 0;
    return 0;
}

extern "C" int adh_sw2_hip_get_state(adh_sw2_hip *h, double *head, double *vel, double *old_head, double *old_vel,
                                     double *older_head, double *older_vel) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    HIPCHK(hipStreamSynchronize(c.stream));
    int rc;
    if ((rc = get_pair(c, c.head, c.u, c.v, head, vel))) return rc;
    if ((rc = get_pair(c, c.old_head, c.old_u, c.old_v, old_head, old_vel))) return rc;
    if ((rc = get_pair(c, c.older_head, c.older_u, c.older_v, older_head, older_vel))) return rc;
    return 0;
}

static int sync_check(Ctx &c) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

extern "C" int adh_sw2_hip_shift(adh_sw2_hip *h) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    launch_shift(&h->c);
    return sync_check(h->c);
}

extern "C" int adh_sw2_hip_init(adh_sw2_hip *h, int *bc_mask, int *wd_flag) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    int *dmask = nullptr;
    if (bc_mask) HIPCHK(hipMalloc(&dmask, sizeof(int) * 3 * c.N));
    launch_init(&c, dmask);
    int rc = sync_check(c);
    if (!rc && bc_mask) rc = hipMemcpy(bc_mask, dmask, sizeof(int) * 3 * c.N, hipMemcpyDeviceToHost) == hipSuccess ? 0 : fail(ADH_SW2_EHIP, "copy bc_mask");
    if (dmask) (void)hipFree(dmask);
    if (!rc && wd_flag) HIPCHK(hipMemcpy(wd_flag, c.wd_flag, sizeof(int) * c.E, hipMemcpyDeviceToHost));
    return rc;
}

// adh_sw2_hip_init with the boundary mask staged in the device's `scratch`
// array: no device allocation and no free.  The kernel writes all 3N mask
// entries, so nothing left by an earlier call is read back.
extern "C" int adh_sw2_hip_init_noalloc(adh_sw2_hip *h, int *bc_mask, int *wd_flag) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    if ((size_t)c.N > NOALLOC_MAX_NODES)
        return fail(ADH_SW2_EINVAL, "init_noalloc: the model has " + std::to_string(c.N) +
                                        " nodes, more than the 2097152 its staging area holds (use adh_sw2_hip_init)");
    if (!h->scratch) return fail(ADH_SW2_EHIP, "init_noalloc: no staging area on this handle's device");
    launch_init(&c, bc_mask ? h->scratch : nullptr);
    int rc = sync_check(c);
    if (rc) return rc;
    if (bc_mask) HIPCHK(hipMemcpy(bc_mask, h->scratch, sizeof(int) * 3 * c.N, hipMemcpyDeviceToHost));
    if (wd_flag) HIPCHK(hipMemcpy(wd_flag, c.wd_flag, sizeof(int) * c.E, hipMemcpyDeviceToHost));
    return 0;
}

// node vectors [3N] / diagonal blocks [9N] between node-major (host) and the
// device's sliced layout
static int nodes_to_host(Ctx &c, const double *dev, double *out, int per) {
    std::vector<double> t((size_t)per * c.Np);
    HIPCHK(hipMemcpy(t.data(), dev, sizeof(double) * t.size(), hipMemcpyDeviceToHost));
    for (int n = 0; n < c.N; n++)
        for (int k = 0; k < per; k++) out[(size_t)per * n + k] = t[per == 3 ? vidx3(n, k) : didx9(n, k)];
    return 0;
}

static int nodes_from_host(Ctx &c, double *dev, const double *in, int per) {
    std::vector<double> t((size_t)per * c.Np, 0.0);
    for (int n = 0; n < c.N; n++)
        for (int k = 0; k < per; k++) t[per == 3 ? vidx3(n, k) : didx9(n, k)] = in[(size_t)per * n + k];
    HIPCHK(hipMemcpy(dev, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int adh_sw2_hip_resid(adh_sw2_hip *h, double *residual) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    launch_resid(&c);
    int rc = sync_check(c);
    if (rc) return rc;
    if (residual && (rc = nodes_to_host(c, c.R, residual, 3))) return rc;
    return 0;
}

// off-diagonal blocks between the reference's row order (host) and SELL-64 (device)
static int offd_to_host(Ctx &c, double *out) {
    std::vector<double> sv(9 * c.nsell);
    if (c.mat32) {   // refined forms: the FP32 matrix (unscaled until the set-up scales it in place)
        std::vector<float> f(sv.size());
        HIPCHK(hipMemcpy(f.data(), c.offd32, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); i++) sv[i] = (double)f[i];
    } else {
        HIPCHK(hipMemcpy(sv.data(), c.offd, sizeof(double) * sv.size(), hipMemcpyDeviceToHost));
    }
    for (int p = 0; p < c.nnz; p++)
        for (int e = 0; e < 9; e++) out[9 * (size_t)p + e] = sv[sell_vidx((size_t)c.h_offd_pos[p], e)];
    return 0;
}

static int offd_from_host(Ctx &c, const double *in) {
    std::vector<double> sv(9 * c.nsell, 0.0);
    for (int p = 0; p < c.nnz; p++)
        for (int e = 0; e < 9; e++) sv[sell_vidx((size_t)c.h_offd_pos[p], e)] = in[9 * (size_t)p + e];
    HIPCHK(hipMemcpy(c.offd, sv.data(), sizeof(double) * sv.size(), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int adh_sw2_hip_load(adh_sw2_hip *h, double *diagonal, double *offdiag, double *residual) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    launch_jacobian(&c);
    launch_assemble(&c);
    int rc = sync_check(c);
    if (rc) return rc;
    if (diagonal && (rc = nodes_to_host(c, c.diag, diagonal, 9))) return rc;
    if (offdiag && c.nnz && (rc = offd_to_host(c, offdiag))) return rc;
    if (residual && (rc = nodes_to_host(c, c.R, residual, 3))) return rc;
    return 0;
}

extern "C" int adh_sw2_hip_elem_jacobians(adh_sw2_hip *h, double *em2d, double *em1d) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    HIPCHK(hipStreamSynchronize(c.stream));
    // the device keeps element blocks in slot order (k_assemble reads them
    // contiguously), in FP32 after a refined-form Jacobian
    std::vector<double> ems(9 * c.h_blkpos.size());
    if (c.em32) {
        std::vector<float> f(ems.size());
        HIPCHK(hipMemcpy(f.data(), c.EM, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); i++) ems[i] = (double)f[i];
    } else {
        HIPCHK(hipMemcpy(ems.data(), c.EM, sizeof(double) * ems.size(), hipMemcpyDeviceToHost));
    }
    const size_t nb2 = 9 * (size_t)c.E;
    for (size_t blk = 0; blk < c.h_blkpos.size(); blk++) {
        double *dst = blk < nb2 ? (em2d ? em2d + 9 * blk : nullptr) : (em1d ? em1d + 9 * (blk - nb2) : nullptr);
        if (!dst) continue;
        for (int k = 0; k < 9; k++) dst[k] = ems[9 * (size_t)c.h_blkpos[blk] + k];
    }
    return 0;
}

extern "C" int adh_sw2_hip_inc(adh_sw2_hip *h, const double *sol) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    int rc;
    if (sol && (rc = nodes_from_host(c, c.sol, sol, 3))) return rc;
    launch_inc(&c, 1.0);
    return sync_check(c);
}

// builds the profile-LU block structure when the options ask for it
static int ensure_band(Ctx &c) {
    if (c.opts.preconditioner == 3) {   // block ILU(0): the structure once per pattern and factor precision
        if (c.ilu_ready && c.ilu_f32 == (c.opts.precision >= 1)) return 0;
        std::string err = ilu_structure(&c);
        if (!err.empty()) return fail(ADH_SW2_ENOMEM, err);
        return 0;
    }
    if (c.opts.preconditioner != 1) return 0;
    if (c.band_nblock == std::min(c.opts.nblock, c.N) && c.band_lu_tiled_opt == c.opts.lu_tiled) return 0;
    c.band_lu_tiled_opt = c.opts.lu_tiled;
    std::string err = band_structure(&c, c.opts.nblock);
    if (!err.empty()) return fail(ADH_SW2_EUNSUPPORTED, err);
    return 0;
}

extern "C" int adh_sw2_hip_linear_sys_setup(adh_sw2_hip *h, const double *x, double *scale_vect) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    // The set-up scales A, b and x in place (as spv_jacobi does).  The
    // refined forms keep the scaled off-diagonal blocks only in FP32, so a
    // second scaling of the same system would start from stale FP64 blocks:
    // refuse it; load or set_system gives the set-up a fresh system.
    if (c.sys_scaled)
        return fail(ADH_SW2_EINVAL, "linear_sys_setup: the system is already scaled (call load or set_system first)");
    int rc0 = ensure_band(c);
    if (rc0) return rc0;
    // x is the solve's initial shift (solv_bcgstab.c:96 keeps it as u0 and
    // solv_return_old.c:23-25 adds it back); fe_newton.c:470 zeroes it before
    // every solve, which a null x stands for.
    if (x) {
        if ((rc0 = nodes_from_host(c, c.sol, x, 3))) return rc0;
    } else {
        launch_zero(&c, c.sol, 3 * (size_t)c.Np);
    }
    launch_setup(&c);
    int rc = sync_check(c);
    if (rc) return rc;
    if (scale_vect && (rc = nodes_to_host(c, c.scale, scale_vect, 3))) return rc;
    return 0;
}

// BiCGSTAB on the scaled device system; returns the reference's YES/NO.
// Iteration cap of the all-FP32 Krylov solve: FP32 vectors can stagnate
// above the 1e-5 relative stop on large meshes, and the FP64 Newton loop
// around it is the refinement that finishes the job.
constexpr int FP32_MAX_LIN_IT = 200;

static int lin_it_cap(const Ctx &c) {
    return c.opts.precision == 2 ? std::min(c.opts.max_lin_it, FP32_MAX_LIN_IT) : c.opts.max_lin_it;
}

static void drop_graphs(Ctx &c) {
    for (int k = 0; k <= KGRAPH_MAX; k++)
        if (c.kgraph[k]) { (void)hipGraphExecDestroy(c.kgraph[k]); c.kgraph[k] = nullptr; }
}

// k BiCGSTAB iterations: a captured graph of the 3k kernels and their
// finishers, replayed (captured on first use for this k)
static int run_iters(Ctx &c, int k) {
    if (k > KGRAPH_MAX || !c.opts.graphs || c.band_tiled) {   // the tiled LU's sweeps: a launch per tile, eager
        launch_krylov_iters(&c, k);
        HIPCHK(hipGetLastError());
        return 0;
    }
    // the captured kernels' arguments: the precision, the preconditioner and its storage
    const double *pstore = c.opts.preconditioner == 3 ? (const double *)c.ilu_L : c.band;
    if (c.kgraph_prec != c.opts.precision || c.kgraph_pc != c.opts.preconditioner || c.kgraph_band != pstore) {
        drop_graphs(c);
        c.kgraph_prec = c.opts.precision;
        c.kgraph_pc = c.opts.preconditioner;
        c.kgraph_band = pstore;
    }
    if (!c.kgraph[k]) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        launch_krylov_iters(&c, k);
        HIPCHK(hipStreamEndCapture(c.stream, &g));
        hipError_t e = hipGraphInstantiate(&c.kgraph[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c.kgraph[k] = nullptr;
            return fail(ADH_SW2_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    HIPCHK(hipGraphLaunch(c.kgraph[k], c.stream));
    return 0;
}

// Krylov chunk lengths between host convergence checks: the first from the
// previous solve's count (opts.check_every for a model's first solve), then
// `more` at a time; with graph replay both are multiples of 4, so few graphs
// are captured.
static void chunk_plan(const Ctx &c, int &first, int &more) {
    const bool graphs = c.opts.graphs != 0;
    if (c.last_lin_it <= 0) first = c.opts.check_every;
    else first = graphs ? ((c.last_lin_it + 3) / 4) * 4 : c.last_lin_it;
    more = graphs ? 4 : 2;
}

// zero_guess: sol holds a zero initial guess (the Newton step zeroes it), so
// the epilogue is idempotent and is queued behind every chunk before the
// convergence flag is read; otherwise it runs once, after the last chunk.
static int krylov(Ctx &c, int *lin_it, int *verdict, bool zero_guess) {
    if (c.opts.krylov == 1) {   // restarted GMRES: host-driven cycles
        std::string err = gmres_alloc(&c);
        if (!err.empty()) return fail(ADH_SW2_ENOMEM, err);
        bool conv = false;
        const int it = launch_gmres(&c, lin_it_cap(c), &conv);
        HIPCHK(hipGetLastError());
        launch_krylov_epilogue(&c, zero_guess);
        c.last_lin_it = it;
        if (lin_it) *lin_it = it;
        *verdict = conv ? ADH_SW2_YES : ADH_SW2_NO;
        return 0;
    }
    KScal init{};
    init.atol = c.opts.lin_atol;
    init.rtol = c.opts.lin_rtol;
    init.max_it = lin_it_cap(c);
    HIPCHK(hipMemcpyAsync(c.ksc, &init, sizeof(KScal), hipMemcpyHostToDevice, c.stream));
    launch_krylov_prologue(&c);
    // first chunk: the previous solve's iteration count (consecutive Newton
    // systems converge alike), then chunks of 2 (chunk_plan).  The device flag
    // makes iterations launched past convergence return at entry, so the
    // first chunk follows the prologue without a host check.
    int first, more;
    chunk_plan(c, first, more);
    int left = lin_it_cap(c);
    // with the tiled band LU an iteration is two sweeps of a launch per tile:
    // the flag is read after the prologue and after every iteration instead
    bool done_early = false;
    if (c.opts.preconditioner == 1 && c.band_tiled) {
        first = more = 1;
        HIPCHK(hipMemcpyAsync(c.h_ksc, c.ksc, sizeof(KScal), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        done_early = c.h_ksc->done;
        if (done_early && zero_guess) launch_krylov_epilogue(&c, true);
    }
    for (int chunk = 0; !done_early; chunk++) {
        int want = chunk == 0 ? first : more;
        int k = std::max(1, std::min(want, left));
        int rc = run_iters(c, k);
        if (rc) return rc;
        if (zero_guess) launch_krylov_epilogue(&c, true);
        HIPCHK(hipMemcpyAsync(c.h_ksc, c.ksc, sizeof(KScal), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        if (c.h_ksc->done) break;
        left = lin_it_cap(c) - c.h_ksc->it;
    }
    c.last_lin_it = c.h_ksc->it;
    if (!zero_guess) launch_krylov_epilogue(&c, false);
    HIPCHK(hipGetLastError());
    if (lin_it) *lin_it = c.h_ksc->it;
    *verdict = (c.h_ksc->it == lin_it_cap(c)) ? ADH_SW2_NO : ADH_SW2_YES;
    return 0;
}

// y = M^-1 v with the preconditioner the last set-up built (profile LU or
// ILU(0)); for tests of the preconditioner on its own.
extern "C" int adh_sw2_hip_precond_apply(adh_sw2_hip *h, const double *v, double *y) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !v || !y) return fail(ADH_SW2_EINVAL, "null argument");
    Ctx &c = h->c;
    if (!c.sys_scaled) return fail(ADH_SW2_EINVAL, "precond_apply: no preconditioner set up (call linear_sys_setup)");
    if (c.opts.preconditioner == 2) return fail(ADH_SW2_EUNSUPPORTED, "precond_apply: profile LU or ILU(0) only");
    int rc;
    if ((rc = nodes_from_host(c, c.kAs, v, 3))) return rc;
    if (c.opts.preconditioner == 3) launch_ilu_solve(&c, c.kAs, nullptr);
    else launch_band_solve(&c, c.kAs, nullptr);
    if ((rc = sync_check(c))) return rc;
    return nodes_to_host(c, c.kAs, y, 3);
}

// The ILU(0) node colouring: color[N], *ncolor.
extern "C" int adh_sw2_hip_ilu_colors(adh_sw2_hip *h, int *color, int *ncolor) {
    if (!h || !color || !ncolor) return fail(ADH_SW2_EINVAL, "null argument");
    Ctx &c = h->c;
    if (!c.ilu_ready) return fail(ADH_SW2_EINVAL, "ilu_colors: no ILU(0) structure (preconditioner 3, after a set-up)");
    std::copy(c.h_ilu_color.begin(), c.h_ilu_color.end(), color);
    *ncolor = c.ilu_ncolor;
    return 0;
}

extern "C" int adh_sw2_hip_linear_sys_solve(adh_sw2_hip *h, double *sol, int *lin_it, int *verdict) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    int vd = ADH_SW2_NO;
    int rc = krylov(c, lin_it, &vd, false);
    if (rc) return rc;
    if ((rc = sync_check(c))) return rc;
    if (verdict) *verdict = vd;
    if (sol && (rc = nodes_to_host(c, c.sol, sol, 3))) return rc;
    return 0;
}

extern "C" int adh_sw2_hip_set_system(adh_sw2_hip *h, const double *diagonal, const double *offdiag, const double *rhs) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !diagonal || !rhs || (!offdiag && h->c.nnz)) return fail(ADH_SW2_EINVAL, "null argument");
    Ctx &c = h->c;
    int rc;
    if ((rc = nodes_from_host(c, c.diag, diagonal, 9))) return rc;
    if (c.nnz && (rc = offd_from_host(c, offdiag))) return rc;
    c.scale_valid = 0;
    c.sys_scaled = 0;
    c.mat32 = 0;   // the system is now the FP64 one given here
    if ((rc = nodes_from_host(c, c.R, rhs, 3))) return rc;
    HIPCHK(hipMemset(c.sol, 0, sizeof(double) * 3 * c.Np));
    return 0;
}

extern "C" int adh_sw2_hip_get_system(adh_sw2_hip *h, double *diagonal, double *offdiag, double *rhs) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h) return fail(ADH_SW2_EINVAL, "null handle");
    Ctx &c = h->c;
    HIPCHK(hipStreamSynchronize(c.stream));
    int rc;
    if (diagonal && (rc = nodes_to_host(c, c.diag, diagonal, 9))) return rc;
    if (offdiag && c.nnz && (rc = offd_to_host(c, offdiag))) return rc;
    if (rhs && (rc = nodes_to_host(c, c.R, rhs, 3))) return rc;
    return 0;
}

extern "C" int adh_sw2_hip_probe_math(const double *x, int n, double *log_out, double *sqrt_out, double *div_out) {
    if (!x || n <= 0 || !log_out || !sqrt_out || !div_out) return fail(ADH_SW2_EINVAL, "bad probe arguments");
    double *dx = nullptr, *dl = nullptr, *ds = nullptr, *dd = nullptr;
    crlog::tables *dt = nullptr;
    crlog::tables T;
    crlog::make_tables(&T);
    int rc = 0;
    if (hipMalloc(&dx, sizeof(double) * n) != hipSuccess || hipMalloc(&dl, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&ds, sizeof(double) * n) != hipSuccess || hipMalloc(&dd, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&dt, sizeof(T)) != hipSuccess) {
        rc = fail(ADH_SW2_ENOMEM, "probe alloc");
    } else if (hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dt, &T, sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(ADH_SW2_EHIP, "probe upload");
    } else {
        launch_probe(dx, n, dt, dl, ds, dd);
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(log_out, dl, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(sqrt_out, ds, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(div_out, dd, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(ADH_SW2_EHIP, "probe run");
    }
    if (dx) (void)hipFree(dx);
    if (dl) (void)hipFree(dl);
    if (ds) (void)hipFree(ds);
    if (dd) (void)hipFree(dd);
    if (dt) (void)hipFree(dt);
    return rc;
}

extern "C" int adh_sw2_hip_time_kernel(adh_sw2_hip *h, int which, int reps, double *avg_ms) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !avg_ms || reps < 1 || which < 0 || which > 4) return fail(ADH_SW2_EINVAL, "bad arguments");
    Ctx &c = h->c;
    double *copy_dst = nullptr;
    const size_t copy_n = 81 * (size_t)c.E;  // the element-matrix slab: 648 MB at 1M triangles
    if (which == 4) HIPCHK(hipMalloc(&copy_dst, sizeof(double) * (copy_n ? copy_n : 1)));
    if (which == 0) {
        KScal k{};
        k.rho = 1.0; k.rhop = 1.0; k.alpha = 1.0; k.omega = 1.0; k.max_it = 1 << 30; k.done = 0;
        HIPCHK(hipMemcpy(c.ksc, &k, sizeof(KScal), hipMemcpyHostToDevice));
    }
    auto launch = [&]() {
        if (which == 0) launch_k1_only(&c);
        else if (which == 1) launch_jac2d_only(&c);
        else if (which == 2) launch_resid2d_only(&c);
        else if (which == 3) launch_assemble_only(&c);
        else launch_copy(&c, c.EM, copy_dst, copy_n);
    };
    launch();  // warm-up
    HIPCHK(hipStreamSynchronize(c.stream));
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, c.stream));
    for (int i = 0; i < reps; i++) launch();
    HIPCHK(hipEventRecord(b, c.stream));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIPCHK(hipGetLastError());
    if (copy_dst) HIPCHK(hipFree(copy_dst));
    *avg_ms = ms / reps;
    return 0;
}

static float elapsed(Ctx &c, int a, int b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c.ev[a], c.ev[b]);
    return ms;
}

static int read_norms(Ctx &c, Norms *out) {
    HIPCHK(hipMemcpyAsync(c.h_norms, c.norms, sizeof(Norms), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    *out = *c.h_norms;
    return 0;
}

static void hist(adh_sw2_report *r, const Norms &n, int lin_it) {
    if (!r) return;
    if (r->nhist < r->hist_cap) {
        if (r->hist_max) r->hist_max[r->nhist] = n.maxn;
        if (r->hist_l2) r->hist_l2[r->nhist] = n.l2;
        if (r->hist_inc) r->hist_inc[r->nhist] = n.incn;
        if (r->hist_lin_it) r->hist_lin_it[r->nhist] = lin_it;
    }
    r->nhist++;
}

// fe_newton (src/fe/fe_newton.c:214-912), serial non-PETSc branch with the
// SW2 callbacks.
static int newton(Ctx &c, adh_sw2_report *rep, int *verdict) {
    const adh_sw2_opts &o = c.opts;
    int rc;
    if ((rc = ensure_band(c))) return rc;
    Norms nm;
    (void)hipEventRecord(c.ev[0], c.stream);
    launch_init(&c, nullptr);
    (void)hipEventRecord(c.ev[1], c.stream);
    launch_resid(&c);
    launch_norms(&c);
    (void)hipEventRecord(c.ev[2], c.stream);
    if ((rc = read_norms(c, &nm))) return rc;
    if (rep) rep->ms_resid += elapsed(c, 1, 2);
    hist(rep, nm, 0);
    int it = 0, solv_flag = ADH_SW2_YES, keep = 0;
    do {
        int lin_it = 0, vd = ADH_SW2_YES;
        (void)hipEventRecord(c.ev[3], c.stream);
        launch_jacobian(&c);
        launch_assemble(&c);
        launch_zero(&c, c.sol, 3 * (size_t)c.Np);
        (void)hipEventRecord(c.ev[4], c.stream);
        launch_setup(&c);
        (void)hipEventRecord(c.ev[5], c.stream);
        if ((rc = krylov(c, &lin_it, &vd, true))) return rc;   // sol zeroed above
        (void)hipEventRecord(c.ev[6], c.stream);
        // FP64: the reference's rule, a Krylov failure fails the Newton step.
        // FP32-refined forms: an unconverged inner solve is still a
        // correction; the FP64 residual tests and the line search decide.
        solv_flag = (o.precision >= 1) ? ADH_SW2_YES : vd;
        launch_inc_flip(&c);   // state += sol; sol = -sol for the line search
        double old_l2 = nm.l2;
        launch_resid(&c);
        launch_norms(&c);
        (void)hipEventRecord(c.ev[7], c.stream);
        if ((rc = read_norms(c, &nm))) return rc;
        if (rep) {
            rep->ms_load += elapsed(c, 3, 4);
            rep->ms_setup += elapsed(c, 4, 5);
            rep->ms_solve += elapsed(c, 5, 6);
            rep->ms_resid += elapsed(c, 6, 7);
            rep->lin_iterations += lin_it;
        }
        if (isnan(nm.l2) || isinf(nm.l2)) solv_flag = ADH_SW2_NO;
        int cuts = 0;
        while (nm.maxn > o.tol_nonlin && nm.l2 > old_l2 && cuts < o.max_cuts) {
            launch_scale_sol(&c, 0.5);
            launch_inc(&c, 1.0);
            launch_resid(&c);
            launch_norms(&c);
            if ((rc = read_norms(c, &nm))) return rc;
            cuts++;
        }
        it++;
        hist(rep, nm, lin_it);
        keep = (solv_flag == ADH_SW2_YES) && (it < o.max_nonlin_it) && (nm.maxn < 1.0E+20) &&
               (nm.maxn > o.tol_nonlin && nm.incn > o.inc_nonlin);
    } while (keep);
    bool failed = (solv_flag != ADH_SW2_YES) || ((nm.maxn > o.tol_nonlin) && (nm.incn > o.inc_nonlin)) ||
                  (nm.maxn >= 1.0E+20);
    if (rep) {
        rep->iterations += it;
        rep->converged = !failed;
        rep->solv_fail = solv_flag != ADH_SW2_YES;
        rep->ms_total += elapsed(c, 0, 7);
    }
    if (failed) {
        launch_init(&c, nullptr);  // fe_newton.c:883
        if ((rc = sync_check(c))) return rc;
        *verdict = ADH_SW2_NO;
        return 0;
    }
    *verdict = ADH_SW2_YES;
    return sync_check(c);
}

extern "C" int adh_sw2_hip_newton(adh_sw2_hip *h, adh_sw2_report *rep, int *verdict) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !verdict) return fail(ADH_SW2_EINVAL, "null argument");
    if (rep) {
        rep->iterations = rep->lin_iterations = rep->nhist = 0;
        rep->ms_load = rep->ms_setup = rep->ms_solve = rep->ms_resid = rep->ms_total = 0;
    }
    return newton(h->c, rep, verdict);
}

extern "C" int adh_sw2_hip_timestep(adh_sw2_hip *h, adh_sw2_report *rep, int *verdict, int max_retries) {
    if (h && h->c.ninst) return fail(ADH_SW2_EINVAL, "an ensemble handle takes the adh_sw2_hip_batch_* calls");
    if (!h || !verdict) return fail(ADH_SW2_EINVAL, "null argument");
    if (rep) {
        rep->iterations = rep->lin_iterations = rep->nhist = 0;
        rep->ms_load = rep->ms_setup = rep->ms_solve = rep->ms_resid = rep->ms_total = 0;
    }
    for (int attempt = 0; attempt <= max_retries; attempt++) {
        launch_shift(&h->c);   // fe_main calls fe_sw2 on every attempt (fe_main.c:348-353)
        int rc = newton(h->c, rep, verdict);
        if (rc) return rc;
        if (*verdict == ADH_SW2_YES) return 0;
        double dt = h->c.dt * 0.25;   // tc_scale, DT_REDUCE_FACTOR (define.h:123)
        if (dt <= SMALL) return fail(ADH_SW2_EINVAL, "time step reduced below DBL_EPSILON");
        adh_sw2_hip_set_dt(h, dt);
    }
    return 0;
}

// ===========================================================================
// Ensembles: many independent SW2 models advanced together.  The models are
// laid side by side in one mesh (each model's nodes start on a multiple of
// 256; the gap nodes are isolated and stay zero), so every phase -- element
// residuals and Jacobians, assembly, scaling, each BiCGSTAB kernel -- is one
// launch over all models; reductions and Krylov scalars are per model
// (256-node blocks never straddle two models), and the Newton/time-step
// decisions are taken per model as the single-model path takes them.  A
// model's arithmetic is the single-model arithmetic: same block partitions,
// same reduction order, same element code.

static int set_mask(Ctx &c, const std::vector<int> &m) {
    c.split_valid = 0;  // the element lists hold the masked models only
    HIPCHK(hipMemcpyAsync(c.inst_mask, m.data(), sizeof(int) * c.ninst, hipMemcpyHostToDevice, c.stream));
    c.h_inst_mask = m;
    // the host vector may change before the copy runs
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

static int read_norms_all(Ctx &c, std::vector<Norms> &out) {
    HIPCHK(hipMemcpyAsync(c.h_norms, c.norms, sizeof(Norms) * c.ninst, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    out.assign(c.h_norms, c.h_norms + c.ninst);
    return 0;
}

// BiCGSTAB for the active models (inactive ones start converged)
static int krylov_batch(Ctx &c, const std::vector<int> &active, std::vector<int> &lin_it, std::vector<int> &vd) {
    const int n = c.ninst, cap = lin_it_cap(c);
    std::vector<KScal> init(n);
    for (int i = 0; i < n; i++) {
        init[i] = KScal{};
        init[i].atol = c.opts.lin_atol;
        init[i].rtol = c.opts.lin_rtol;
        init[i].max_it = active[i] ? cap : 0;
    }
    HIPCHK(hipMemcpyAsync(c.ksc, init.data(), sizeof(KScal) * n, hipMemcpyHostToDevice, c.stream));
    launch_krylov_prologue(&c);
    // chunks as in krylov() with a zero guess (newton_batch zeroes sol): no
    // host check between the prologue and the first chunk, the idempotent
    // epilogue queued behind every chunk before the flags are read
    int first, more;
    chunk_plan(c, first, more);
    int maxit = 0;
    int left = cap;
    // the tiled band LU: flags read after the prologue and every iteration,
    // as krylov() does (a direct solve usually ends at the prologue)
    bool all_done0 = false;
    if (c.opts.preconditioner == 1 && c.band_tiled) {
        first = more = 1;
        HIPCHK(hipMemcpyAsync(c.h_ksc, c.ksc, sizeof(KScal) * n, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        all_done0 = true;
        for (int i = 0; i < n; i++) all_done0 = all_done0 && c.h_ksc[i].done;
        if (all_done0) launch_krylov_epilogue(&c, true);
    }
    for (int chunk = 0; !all_done0; chunk++) {
        int want = chunk == 0 ? first : more;
        int rc = run_iters(c, std::max(1, std::min(want, left)));
        if (rc) return rc;
        launch_krylov_epilogue(&c, true);
        HIPCHK(hipMemcpyAsync(c.h_ksc, c.ksc, sizeof(KScal) * n, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        left = 0;
        bool all_done = true;
        for (int i = 0; i < n; i++)
            if (!c.h_ksc[i].done) {
                all_done = false;
                left = std::max(left, cap - c.h_ksc[i].it);
            }
        if (all_done) break;
    }
    HIPCHK(hipGetLastError());
    lin_it.assign(n, 0);
    vd.assign(n, ADH_SW2_YES);
    for (int i = 0; i < n; i++) {
        if (!active[i]) continue;
        lin_it[i] = c.h_ksc[i].it;
        vd[i] = c.h_ksc[i].it == cap ? ADH_SW2_NO : ADH_SW2_YES;
        maxit = std::max(maxit, lin_it[i]);
    }
    c.last_lin_it = maxit;
    return 0;
}

// fe_newton for every model with run[i] set (src/fe/fe_newton.c:214-912)
static int newton_batch(Ctx &c, const std::vector<int> &run, adh_sw2_report *reps, std::vector<int> &verdict) {
    const adh_sw2_opts &o = c.opts;
    const int n = c.ninst;
    int rc;
    std::vector<double> tol(n);
    for (int i = 0; i < n; i++) tol[i] = std::isnan(c.h_inst_tol[i]) ? o.tol_nonlin : c.h_inst_tol[i];
    std::vector<Norms> nm;
    if ((rc = ensure_band(c))) return rc;   // the profile LU's bands, once per partition
    if ((rc = set_mask(c, run))) return rc;
    launch_init(&c, nullptr);   // masked: only the models of this attempt
    launch_resid(&c);
    launch_norms(&c);
    if ((rc = read_norms_all(c, nm))) return rc;
    std::vector<int> active(run), it(n, 0), solv(n, ADH_SW2_YES), lin_it, vd;
    for (int i = 0; i < n; i++)
        if (run[i] && reps) hist(reps + i, nm[i], 0);
    auto any = [](const std::vector<int> &v) { for (int x : v) if (x) return true; return false; };
    while (any(active)) {
        if ((rc = set_mask(c, active))) return rc;
        launch_jacobian(&c);
        launch_assemble(&c);
        launch_zero(&c, c.sol, 3 * (size_t)c.Np);
        launch_setup(&c);
        if ((rc = krylov_batch(c, active, lin_it, vd))) return rc;
        for (int i = 0; i < n; i++)
            if (active[i]) solv[i] = (o.precision >= 1) ? ADH_SW2_YES : vd[i];
        if ((rc = set_mask(c, active))) return rc;
        launch_inc_flip(&c);   // state += sol; sol = -sol for the line search
        std::vector<double> old_l2(n);
        for (int i = 0; i < n; i++) old_l2[i] = nm[i].l2;
        launch_resid(&c);
        launch_norms(&c);
        {
            // finished models keep their last norms (their correction
            // vector no longer belongs to them)
            std::vector<Norms> nm2;
            if ((rc = read_norms_all(c, nm2))) return rc;
            for (int i = 0; i < n; i++)
                if (active[i]) nm[i] = nm2[i];
        }
        std::vector<int> cuts(n, 0);
        for (int i = 0; i < n; i++)
            if (active[i] && (isnan(nm[i].l2) || isinf(nm[i].l2))) solv[i] = ADH_SW2_NO;
        for (;;) {   // line search, for the models that need it
            std::vector<int> cut(n, 0);
            for (int i = 0; i < n; i++)
                cut[i] = active[i] && nm[i].maxn > tol[i] && nm[i].l2 > old_l2[i] && cuts[i] < o.max_cuts;
            if (!any(cut)) break;
            if ((rc = set_mask(c, cut))) return rc;
            launch_scale_sol(&c, 0.5);
            launch_inc(&c, 1.0);
            launch_resid(&c);
            launch_norms(&c);
            std::vector<Norms> nm2;
            if ((rc = read_norms_all(c, nm2))) return rc;
            for (int i = 0; i < n; i++)
                if (cut[i]) { nm[i] = nm2[i]; cuts[i]++; }
        }
        for (int i = 0; i < n; i++) {
            if (!active[i]) continue;
            it[i]++;
            if (reps) {
                hist(reps + i, nm[i], lin_it[i]);
                reps[i].lin_iterations += lin_it[i];
            }
            const bool keep = (solv[i] == ADH_SW2_YES) && (it[i] < o.max_nonlin_it) && (nm[i].maxn < 1.0E+20) &&
                              (nm[i].maxn > tol[i] && nm[i].incn > o.inc_nonlin);
            if (!keep) active[i] = 0;
        }
    }
    std::vector<int> failed(n, 0);
    verdict.assign(n, ADH_SW2_YES);
    for (int i = 0; i < n; i++) {
        if (!run[i]) continue;
        const bool f = (solv[i] != ADH_SW2_YES) || ((nm[i].maxn > tol[i]) && (nm[i].incn > o.inc_nonlin)) ||
                       (nm[i].maxn >= 1.0E+20);
        failed[i] = f;
        verdict[i] = f ? ADH_SW2_NO : ADH_SW2_YES;
        if (reps) {
            reps[i].iterations += it[i];
            reps[i].converged = !f;
            reps[i].solv_fail = solv[i] != ADH_SW2_YES;
        }
    }
    if (any(failed)) {
        if ((rc = set_mask(c, failed))) return rc;
        launch_init(&c, nullptr);  // fe_newton.c:883, for the failed models only
    }
    return sync_check(c);
}

extern "C" int adh_sw2_hip_batch_create(const adh_sw2_mesh *meshes, int n, const adh_sw2_opts *opts,
                                        adh_sw2_hip **out) {
    if (!meshes || n < 1 || !out) return fail(ADH_SW2_EINVAL, "null argument or empty ensemble");
    int rc;
    if (opts && (rc = check_batch_opts(opts))) return rc;
    for (int i = 0; i < n; i++)
        if ((rc = validate(meshes + i))) return rc;
    const adh_sw2_mesh &m0 = meshes[0];
    for (int i = 1; i < n; i++) {
        const adh_sw2_mesh &m = meshes[i];
        if (m.gravity != m0.gravity || m.tau_temporal != m0.tau_temporal || m.tau_pg != m0.tau_pg ||
            m.drying_lower_limit != m0.drying_lower_limit || m.drying_upper_limit != m0.drying_upper_limit ||
            m.viscosity != m0.viscosity || m.density != m0.density)
            return fail(ADH_SW2_EINVAL,
                        "ensemble models must share gravity, tau, drying limits, viscosity and density");
    }
    // union layout
    std::vector<int> node0(n), E0(n), B0(n), M0(n);
    long long nn = 0, ne = 0, nb = 0, nm = 0;
    for (int i = 0; i < n; i++) {
        node0[i] = (int)nn; E0[i] = (int)ne; B0[i] = (int)nb; M0[i] = (int)nm;
        nn += meshes[i].nnodes;
        if (i + 1 < n) nn = (nn + 255) / 256 * 256;
        ne += meshes[i].nelems2d; nb += meshes[i].nelems1d; nm += meshes[i].nmat;
    }
    if (nn > 2000000000LL || ne > MAX_TRIANGLES) return fail(ADH_SW2_EINVAL, "ensemble too large");
    const int N = (int)nn, E = (int)ne, B = (int)nb, M = (int)nm;
    std::vector<double> x(N, 0.0), y(N, 0.0), z(N, 0.0), bcval(3 * (size_t)N, 0.0);
    std::vector<int> node_bc(N, 0), e2d(3 * (size_t)E), emat(E), esf(E), e1d(2 * (size_t)B), eadj(B), ebc(B);
    std::vector<double> eks(E), esrc(E), eval(B);
    std::vector<int> mturb(M), mmode(M);
    std::vector<double> mcoef(M), mev(3 * (size_t)M), mcor(M);
    std::vector<double> gdj(E), ggr(6 * (size_t)E), gdj1(B), gnr1(2 * (size_t)B);
    std::vector<double> mwatt(M, 1.0);
    std::vector<int> mwflag(M, 0);
    int any_wind = 0;
    for (int i = 0; i < n; i++) {
        const adh_sw2_mesh &m = meshes[i];
        std::vector<double> dj, gr, dj1, nr1;
        mesh_geometry(&m, dj, gr, dj1, nr1);
        std::copy(dj.begin(), dj.end(), gdj.begin() + E0[i]);
        std::copy(gr.begin(), gr.end(), ggr.begin() + 6 * (size_t)E0[i]);
        std::copy(dj1.begin(), dj1.end(), gdj1.begin() + B0[i]);
        std::copy(nr1.begin(), nr1.end(), gnr1.begin() + 2 * (size_t)B0[i]);
        for (int k = 0; k < m.nnodes; k++) {
            x[node0[i] + k] = m.x[k]; y[node0[i] + k] = m.y[k]; z[node0[i] + k] = m.z[k];
            if (m.node_bc) node_bc[node0[i] + k] = m.node_bc[k];
            if (m.node_bc_val)
                for (int j = 0; j < 3; j++) bcval[3 * (size_t)(node0[i] + k) + j] = m.node_bc_val[3 * (size_t)k + j];
        }
        for (int e = 0; e < m.nelems2d; e++) {
            for (int j = 0; j < 3; j++) e2d[3 * (size_t)(E0[i] + e) + j] = m.e2d[3 * (size_t)e + j] + node0[i];
            emat[E0[i] + e] = m.e2d_mat[e] + M0[i];
            eks[E0[i] + e] = m.e2d_ks[e];
            esf[E0[i] + e] = m.e2d_src_flag ? m.e2d_src_flag[e] : 0;
            esrc[E0[i] + e] = m.e2d_src ? m.e2d_src[e] : 0.0;
        }
        for (int e = 0; e < m.nelems1d; e++) {
            for (int j = 0; j < 2; j++) e1d[2 * (size_t)(B0[i] + e) + j] = m.e1d[2 * (size_t)e + j] + node0[i];
            eadj[B0[i] + e] = m.e1d_adj ? m.e1d_adj[e] + E0[i] : E0[i];
            ebc[B0[i] + e] = m.e1d_bc[e];
            eval[B0[i] + e] = m.e1d_val[e];
        }
        for (int k = 0; k < m.nmat; k++) {
            mturb[M0[i] + k] = m.mat_turb[k];
            mmode[M0[i] + k] = m.mat_eev_mode ? m.mat_eev_mode[k] : 0;
            mcoef[M0[i] + k] = m.mat_eev_coef ? m.mat_eev_coef[k] : 0.0;
            for (int j = 0; j < 3; j++) mev[3 * (size_t)(M0[i] + k) + j] = m.mat_ev[3 * (size_t)k + j];
            mcor[M0[i] + k] = m.mat_cor_speed[k];
            if (m.wind && m.mat_windatt) mwatt[M0[i] + k] = m.mat_windatt[k];
            if (m.wind && m.mat_wind_flag) mwflag[M0[i] + k] = m.mat_wind_flag[k];
        }
        any_wind |= m.wind != 0;
    }
    adh_sw2_mesh u = {};
    u.nnodes = N; u.nelems2d = E; u.nelems1d = B; u.nmat = M;
    u.x = x.data(); u.y = y.data(); u.z = z.data(); u.node_bc = node_bc.data(); u.node_bc_val = bcval.data();
    u.e2d = e2d.data(); u.e2d_mat = emat.data(); u.e2d_ks = eks.data(); u.e2d_src_flag = esf.data();
    u.e2d_src = esrc.data();
    u.e1d = B ? e1d.data() : nullptr; u.e1d_adj = B ? eadj.data() : nullptr; u.e1d_bc = B ? ebc.data() : nullptr;
    u.e1d_val = B ? eval.data() : nullptr;
    u.mat_turb = mturb.data(); u.mat_eev_mode = mmode.data(); u.mat_eev_coef = mcoef.data(); u.mat_ev = mev.data();
    u.mat_cor_speed = mcor.data();
    u.wind = any_wind;
    u.mat_windatt = mwatt.data();
    u.mat_wind_flag = mwflag.data();
    u.density = m0.density;
    u.e2d_djac = gdj.data(); u.e2d_grad = ggr.data();
    u.e1d_djac = B ? gdj1.data() : nullptr; u.e1d_nrml = B ? gnr1.data() : nullptr;
    u.gravity = m0.gravity; u.dt = m0.dt; u.tau_temporal = m0.tau_temporal; u.tau_pg = m0.tau_pg;
    u.drying_lower_limit = m0.drying_lower_limit; u.drying_upper_limit = m0.drying_upper_limit;
    u.viscosity = m0.viscosity;
    adh_sw2_hip *h = nullptr;
    if ((rc = create_impl(&u, opts, &h, n))) return rc;
    Ctx &c = h->c;
    // model tables
    // per 256-node block: its model and that model's end (gap nodes are never
    // computed); per node: its model, -1 for gap and padding nodes
    const int nblk = c.Np / 256 + 1;
    std::vector<int> blk_inst(2 * (size_t)nblk, 0), inst_blk0(n + 1), inst_nodes(n), node_inst(c.Np, -1), elem_inst(E),
        edge_inst(B > 0 ? B : 1, 0), loc_bc(3 * (size_t)n, 0), mask(n, 1);
    std::vector<double> dts(n);
    for (int i = 0; i < n; i++) {
        const int a = node0[i], b = (i + 1 < n) ? node0[i + 1] : c.Np;
        inst_blk0[i] = a / 256;
        inst_nodes[i] = meshes[i].nnodes;
        for (int k = a; k < a + meshes[i].nnodes; k++) node_inst[k] = i;
        for (int k = a / 256; k < (b + 255) / 256 && k < nblk; k++) {
            blk_inst[2 * k] = i;
            blk_inst[2 * k + 1] = a + meshes[i].nnodes;
        }
        for (int e = 0; e < meshes[i].nelems2d; e++) elem_inst[E0[i] + e] = i;
        for (int e = 0; e < meshes[i].nelems1d; e++) edge_inst[B0[i] + e] = i;
        for (int k = 0; k < 3; k++)
            loc_bc[3 * i + k] = (meshes[i].node_bc && k < meshes[i].nnodes) ? meshes[i].node_bc[k] : 0;
        dts[i] = meshes[i].dt;
    }
    inst_blk0[n] = (c.N + 255) / 256;
    auto up = [&](auto **dst, const auto *src, size_t cnt) -> int { return dupload(h, dst, src, cnt); };
    if ((rc = up(&c.blk_inst, blk_inst.data(), blk_inst.size())) || (rc = up(&c.inst_blk0, inst_blk0.data(), n + 1)) ||
        (rc = up(&c.inst_nodes, inst_nodes.data(), n)) || (rc = up(&c.node_inst, node_inst.data(), node_inst.size())) ||
        (rc = up(&c.inst_mask, mask.data(), n)) || (rc = up(&c.elem_inst, elem_inst.data(), elem_inst.size())) ||
        (rc = up(&c.edge_inst, edge_inst.data(), edge_inst.size())) || (rc = up(&c.dt_inst, dts.data(), n)) ||
        (rc = up(&c.loc_bc_inst, loc_bc.data(), loc_bc.size()))) {
        adh_sw2_hip_destroy(h);
        return rc;
    }
    c.h_inst_node0 = node0;
    c.h_inst_N = inst_nodes;
    c.h_inst_mask = mask;
    c.h_inst_e0 = E0;
    c.h_inst_b0 = B0;
    c.h_inst_E.resize(n);
    c.h_inst_B.resize(n);
    for (int i = 0; i < n; i++) {
        c.h_inst_E[i] = meshes[i].nelems2d;
        c.h_inst_B[i] = meshes[i].nelems1d;
    }
    c.h_inst_dt = dts;
    c.h_inst_tol.assign(n, std::nan(""));
    c.ninst = n;
    c.ed.dt_inst = c.dt_inst;
    c.ed.elem_inst = c.elem_inst;
    c.ed.edge_inst = c.edge_inst;
    c.ed.loc_bc_inst = c.loc_bc_inst;
    HIPCHK(hipDeviceSynchronize());
    *out = h;
    return ADH_SW2_OK;
}

extern "C" int adh_sw2_hip_batch_size(adh_sw2_hip *h) { return h ? h->c.ninst : ADH_SW2_EINVAL; }

static int model_range(adh_sw2_hip *h, int i, int *a, int *nn) {
    if (!h || h->c.ninst < 1) return fail(ADH_SW2_EINVAL, "not an ensemble handle");
    if (i < 0 || i >= h->c.ninst) return fail(ADH_SW2_EINVAL, "model index out of range");
    *a = h->c.h_inst_node0[i];
    *nn = h->c.h_inst_N[i];
    return 0;
}

extern "C" int adh_sw2_hip_batch_set_state(adh_sw2_hip *h, int i, const double *head, const double *vel,
                                           const double *old_head, const double *old_vel, const double *older_head,
                                           const double *older_vel) {
    int a, nn, rc;
    if ((rc = model_range(h, i, &a, &nn))) return rc;
    if (!head || !vel || !old_head || !old_vel || !older_head || !older_vel) return fail(ADH_SW2_EINVAL, "null argument");
    Ctx &c = h->c;
    HIPCHK(hipStreamSynchronize(c.stream));
    const double *src[6] = {head, vel, old_head, old_vel, older_head, older_vel};
    double *hd[3] = {c.head, c.old_head, c.older_head}, *ud[3] = {c.u, c.old_u, c.older_u},
           *vd[3] = {c.v, c.old_v, c.older_v};
    std::vector<double> u(nn), v(nn);
    for (int k = 0; k < 3; k++) {
        HIPCHK(hipMemcpy(hd[k] + a, src[2 * k], sizeof(double) * nn, hipMemcpyHostToDevice));
        for (int j = 0; j < nn; j++) { u[j] = src[2 * k + 1][2 * j]; v[j] = src[2 * k + 1][2 * j + 1]; }
        HIPCHK(hipMemcpy(ud[k] + a, u.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(vd[k] + a, v.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    }
    c.pre_valid = 0;
    c.split_valid = 0;
    return 0;
}

extern "C" int adh_sw2_hip_batch_get_state(adh_sw2_hip *h, int i, double *head, double *vel) {
    int a, nn, rc;
    if ((rc = model_range(h, i, &a, &nn))) return rc;
    Ctx &c = h->c;
    HIPCHK(hipStreamSynchronize(c.stream));
    if (head) HIPCHK(hipMemcpy(head, c.head + a, sizeof(double) * nn, hipMemcpyDeviceToHost));
    if (vel) {
        std::vector<double> u(nn), v(nn);
        HIPCHK(hipMemcpy(u.data(), c.u + a, sizeof(double) * nn, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(v.data(), c.v + a, sizeof(double) * nn, hipMemcpyDeviceToHost));
        for (int j = 0; j < nn; j++) { vel[2 * j] = u[j]; vel[2 * j + 1] = v[j]; }
    }
    return 0;
}

extern "C" int adh_sw2_hip_batch_set_dt(adh_sw2_hip *h, int i, double dt) {
    int a, nn, rc;
    if ((rc = model_range(h, i, &a, &nn))) return rc;
    if (!(dt > 0)) return fail(ADH_SW2_EINVAL, "dt must be positive");
    Ctx &c = h->c;
    c.h_inst_dt[i] = dt;
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipMemcpy(c.dt_inst + i, &dt, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int adh_sw2_hip_batch_get_dt(adh_sw2_hip *h, int i, double *dt) {
    int a, nn, rc;
    if ((rc = model_range(h, i, &a, &nn))) return rc;
    if (!dt) return fail(ADH_SW2_EINVAL, "null argument");
    *dt = h->c.h_inst_dt[i];
    return 0;
}

extern "C" int adh_sw2_hip_batch_set_wind(adh_sw2_hip *h, int i, const double *stress_xy) {
    int a, nn, rc;
    if ((rc = model_range(h, i, &a, &nn))) return rc;
    if (!stress_xy) return fail(ADH_SW2_EINVAL, "null argument");
    Ctx &c = h->c;
    if (!c.ed.wind) return fail(ADH_SW2_EINVAL, "the ensemble was created without winds");
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipMemcpy(c.wind + 2 * (size_t)a, stress_xy, sizeof(double) * 2 * (size_t)nn, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int adh_sw2_hip_batch_set_series(adh_sw2_hip *h, int i, const double *e1d_val, const double *e2d_src) {
    int a, nn, rc;
    if ((rc = model_range(h, i, &a, &nn))) return rc;
    Ctx &c = h->c;
    return upload_series(c, c.h_inst_e0[i], c.h_inst_E[i], c.h_inst_b0[i], c.h_inst_B[i], e1d_val, e2d_src);
}

extern "C" int adh_sw2_hip_batch_set_tol(adh_sw2_hip *h, int i, double tol_nonlin) {
    int a, nn, rc;
    if ((rc = model_range(h, i, &a, &nn))) return rc;
    if (!(tol_nonlin > 0)) return fail(ADH_SW2_EINVAL, "tol_nonlin must be positive");
    h->c.h_inst_tol[i] = tol_nonlin;
    return 0;
}

// fe_newton's opening residual (init, residual, norms) for every model
extern "C" int adh_sw2_hip_batch_residual_norm(adh_sw2_hip *h, double *maxn) {
    if (!h || h->c.ninst < 1 || !maxn) return fail(ADH_SW2_EINVAL, "null argument or not an ensemble");
    Ctx &c = h->c;
    int rc;
    std::vector<Norms> nm;
    if ((rc = set_mask(c, std::vector<int>(c.ninst, 1)))) return rc;
    launch_init(&c, nullptr);
    launch_resid(&c);
    launch_norms(&c);
    if ((rc = read_norms_all(c, nm))) return rc;
    for (int i = 0; i < c.ninst; i++) maxn[i] = nm[i].maxn;
    return sync_check(c);
}

// One time step of every model (fe_main's attempts per model: shift, Newton,
// dt * 0.25 after a failure, src/main/adh_func.c:535-549).  reps and
// verdicts hold one entry per model.
extern "C" int adh_sw2_hip_batch_timestep(adh_sw2_hip *h, const int *which, adh_sw2_report *reps, int *verdicts,
                                          int max_retries) {
    if (!h || h->c.ninst < 1 || !verdicts) return fail(ADH_SW2_EINVAL, "null argument or not an ensemble");
    Ctx &c = h->c;
    const int n = c.ninst;
    if (reps)
        for (int i = 0; i < n; i++) {
            reps[i].iterations = reps[i].lin_iterations = reps[i].nhist = 0;
            reps[i].ms_load = reps[i].ms_setup = reps[i].ms_solve = reps[i].ms_resid = reps[i].ms_total = 0;
        }
    std::vector<int> run(n, 1), vd;
    for (int i = 0; i < n; i++) {
        if (which) run[i] = which[i] != 0;
        verdicts[i] = run[i] ? ADH_SW2_NO : ADH_SW2_YES;
    }
    for (int attempt = 0; attempt <= max_retries; attempt++) {
        int rc;
        if ((rc = set_mask(c, run))) return rc;
        launch_shift(&c);   // masked: the models of this attempt
        if ((rc = newton_batch(c, run, reps, vd))) return rc;
        bool again = false;
        for (int i = 0; i < n; i++) {
            if (!run[i]) continue;
            verdicts[i] = vd[i];
            if (vd[i] == ADH_SW2_YES) { run[i] = 0; continue; }
            double dt = c.h_inst_dt[i] * 0.25;   // tc_scale, DT_REDUCE_FACTOR (define.h:123)
            if (dt <= SMALL) { run[i] = 0; continue; }
            if ((rc = adh_sw2_hip_batch_set_dt(h, i, dt))) return rc;
            again = true;
        }
        if (!again) break;
    }
    return 0;
}

// ===========================================================================
// Launch-shape autotuning of the element list kernels

extern "C" int adh_sw2_hip_get_tuning(adh_sw2_hip *h, adh_sw2_tuning *out) {
    if (!h || !out) return fail(ADH_SW2_EINVAL, "null argument");
    for (int k = 0; k < 4; k++) {
        out->block[k] = h->c.tune.bs[k];
        out->grid[k] = h->c.tune.grid[k];
        out->layout[k] = h->c.tune.layout[k];
        out->ms[k] = 0.0;
    }
    return 0;
}

extern "C" int adh_sw2_hip_set_tuning(adh_sw2_hip *h, const adh_sw2_tuning *t) {
    if (!h || !t) return fail(ADH_SW2_EINVAL, "null argument");
    for (int k = 0; k < 4; k++) {
        if (t->block[k] != 64 && t->block[k] != 128 && t->block[k] != 256)
            return fail(ADH_SW2_EINVAL, "list kernel block size must be 64, 128 or 256");
        if (t->grid[k] < 1 || t->grid[k] > (1 << 20)) return fail(ADH_SW2_EINVAL, "list kernel grid out of range");
        if (t->layout[k] < 0 || t->layout[k] > (k < 2 ? 3 : 0))
            return fail(ADH_SW2_EINVAL, "layout must be 0 to 3 for the Jacobian lists, 0 for the residual lists");
    }
    for (int k = 0; k < 4; k++) {
        h->c.tune.bs[k] = t->block[k];
        h->c.tune.grid[k] = t->grid[k];
        h->c.tune.layout[k] = t->layout[k];
    }
    return 0;
}

extern "C" int adh_sw2_hip_autotune(adh_sw2_hip *h, int reps, adh_sw2_tuning *out) {
    if (!h || reps < 1) return fail(ADH_SW2_EINVAL, "null handle or reps < 1");
    Ctx &c = h->c;
    int rc;
    if (c.ninst && (rc = set_mask(c, std::vector<int>(c.ninst, 1)))) return rc;
    if (!c.E) {
        if (out) adh_sw2_hip_get_tuning(h, out);
        return 0;
    }
    ensure_lists(&c);
    int cnt[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(cnt, c.ecount, sizeof(cnt), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    static const int BS[3] = {64, 128, 256};
    static const int GRID[8] = {256, 512, 1024, 2048, 4096, 8192, 16384, 32768};
    adh_sw2_tuning best;
    adh_sw2_hip_get_tuning(h, &best);
    for (int phase = 0; phase < 4; phase++) {
        const int items = (phase < 2 ? 18 : 1) * cnt[phase & 1];
        if (items == 0) continue;
        double tbest = 1e30;
        for (int layout : {0, 1, 2, 3}) {
            if (layout && phase >= 2) continue;
            for (int b : BS) {
                const int full = nblocks_for(items, b);
                for (int gi = 0; gi < 8; gi++) {
                    const int g = std::min(GRID[gi], full);
                    if (gi > 0 && std::min(GRID[gi - 1], full) == g) continue;   // same shape as the previous one
                    launch_list_phase(&c, phase, b, g, layout);
                    (void)hipEventRecord(c.ev[0], c.stream);
                    for (int r = 0; r < reps; r++) launch_list_phase(&c, phase, b, g, layout);
                    (void)hipEventRecord(c.ev[1], c.stream);
                    HIPCHK(hipEventSynchronize(c.ev[1]));
                    HIPCHK(hipGetLastError());
                    const double ms = (double)elapsed(c, 0, 1) / reps;
                    if (ms < tbest) {
                        tbest = ms;
                        best.block[phase] = b;
                        best.grid[phase] = g;
                        best.layout[phase] = layout;
                        best.ms[phase] = ms;
                    }
                }
            }
        }
    }
    for (int k = 0; k < 4; k++) {
        c.tune.bs[k] = best.block[k];
        c.tune.grid[k] = best.grid[k];
        c.tune.layout[k] = best.layout[k];
    }
    if (out) *out = best;
    return sync_check(c);
}


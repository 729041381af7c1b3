// Element-level kernels of the SW2 Newton step on gfx950: residual,
// finite-difference Jacobian, ordered assembly into 3x3 block rows, and the
// per-node state kernels (init / wet-dry flags / shift / increment / norms).
//
// Layout in HBM: node fields are structure-of-arrays doubles; element
// connectivity and geometry are element-indexed arrays, so one thread per
// element (or per element column) issues coalesced loads for everything but
// the node gathers.  Element results go to contiguous per-element slabs
// (ER: 9 doubles per triangle, EM: 81 per triangle) and are then gathered
// per node / per matrix block in ascending element order -- the order in
// which the reference accumulates (fe_sw2_resid.c:74-191,
// fe_assemble_matrix.c:166-224) -- so sums are bit-identical to it without
// atomics.
#include <algorithm>

#include "sw2_ctx.h"

namespace sw2 {

int nblocks_for(int n, int bs) { return n <= 0 ? 1 : (n + bs - 1) / bs; }

static void launch_resid2d_split(Ctx *c, const ElemData &D);

// The refined forms (precision >= 1) store the element Jacobian blocks in
// FP32 (the matrix they assemble is stored in FP32 anyway); the FP64 form in
// FP64.  Returns the FP32 view of the slab, or null for FP64 storage.
static float *em32_of(Ctx *c) { return c->opts.precision >= 1 ? (float *)c->EM : nullptr; }

// ------------------------------------------------------------------ residual
// The same over a triangle list (grid-stride over the device-side count):
// WET lists hold triangles with positive depths, run without clipping code.
// Software pipeline: the triangle id two iterations ahead and its node ids
// one iteration ahead are loaded while the current triangle is computed.
template <bool WET, int BS>
__global__ void __launch_bounds__(BS) k_resid2d_list(ElemData D, const int *list, const int *count, double *ER) {
    const int n = *count;
    const int stride = gridDim.x * blockDim.x;
    int li = blockIdx.x * blockDim.x + threadIdx.x;
    int ie_a = li < n ? list[li] : 0;
    int ie_b = li + stride < n ? list[li + stride] : 0;
    int nd_a[3] = {0, 0, 0};
    if (li < n)
        for (int k = 0; k < 3; k++) nd_a[k] = D.e2d[3 * ie_a + k];
    for (; li < n; li += stride) {
        const int ie = ie_a;
        const int nodes[3] = {nd_a[0], nd_a[1], nd_a[2]};
        ie_a = ie_b;
        ie_b = li + 2 * stride < n ? list[li + 2 * stride] : 0;
        if (li + stride < n)
            for (int k = 0; k < 3; k++) nd_a[k] = D.e2d[3 * ie_a + k];
        d3 r[3];
        body_resid<WET>(D, D.pre + PRE_STRIDE * (size_t)ie, ie, 0.0, 0, PERTURB_NONE, 0, true, r, nodes);
        double *o = ER + 9 * (size_t)ie;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            o[3 * i + 0] = r[i].x_eq;
            o[3 * i + 1] = r[i].y_eq;
            o[3 * i + 2] = r[i].c_eq;
        }
    }
}

__global__ void __launch_bounds__(128) k_resid1d(ElemData D, int B, double *ER1) {
    int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= B) return;
    d3 r[3];
    boundary_resid(D, ie, 0.0, 0, PERTURB_NONE, 0, true, r);
    double *o = ER1 + 6 * (size_t)ie;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        o[3 * i + 0] = r[i].x_eq;
        o[3 * i + 1] = r[i].y_eq;
        o[3 * i + 2] = r[i].c_eq;
    }
}

// residual[3n..] = -(sum of element contributions) in ascending element
// order, 2D first then 1D; dry-node momentum contributions are dropped and
// dry-node velocities zeroed (fe_sw2_resid.c:85-93, :154-162).
__global__ void k_resid_gather(int N, const int *n2e_ptr, const int *n2e, const int *n2b_ptr, const int *n2b,
                               const double *ER, const double *ER1, const double *old_head, double *u, double *v,
                               double *R) {
    int n = xcd_block() * blockDim.x + threadIdx.x;   // a node's elements are shared with its neighbours
    if (n >= N) return;
    bool dry = old_head[n] <= 0.;
    double rx = 0.0, ry = 0.0, rc = 0.0;
    const int k0 = n2e_ptr[n], k1 = n2e_ptr[n + 1];
    int k = k0;
    // batches of up to 8 contributions: the indices, then the values, are
    // loaded before the (in-order) sums, so their latencies overlap
    for (; k < k1; k += 8) {
        const int m = min(8, k1 - k);
        int sidx[8];
#pragma unroll
        for (int t = 0; t < 8; t++) sidx[t] = t < m ? n2e[k + t] : 0;
        double vx[8], vy[8], vc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const double *o = ER + 9 * (size_t)(sidx[t] / 3) + 3 * (sidx[t] % 3);
            vx[t] = t < m ? o[0] : 0.0;
            vy[t] = t < m ? o[1] : 0.0;
            vc[t] = t < m ? o[2] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (t < m) {
                if (!dry) { rx -= vx[t]; ry -= vy[t]; }
                rc -= vc[t];
            }
        }
    }
    for (int k = n2b_ptr[n]; k < n2b_ptr[n + 1]; k++) {
        int s = n2b[k];
        const double *o = ER1 + 6 * (size_t)(s / 2) + 3 * (s % 2);
        if (!dry) { rx -= o[0]; ry -= o[1]; }
        rc -= o[2];
    }
    R[vidx3(n, 0)] = rx;
    R[vidx3(n, 1)] = ry;
    R[vidx3(n, 2)] = rc;
    if (dry && (n2e_ptr[n + 1] > n2e_ptr[n] || n2b_ptr[n + 1] > n2b_ptr[n])) {
        u[n] = 0.;
        v[n] = 0.;
    }
}

void launch_resid(Ctx *c) {
    if (!c->pre_valid) launch_old_terms(c);
    ElemData D = c->ed;
    D.first_elem = c->first_elem;
    launch_resid2d_split(c, D);
    if (c->B) hipLaunchKernelGGL(k_resid1d, dim3(nblocks_for(c->B, 128)), dim3(128), 0, c->stream, D, c->B,
                                 c->ER + 9 * (size_t)c->E);
    hipLaunchKernelGGL(k_resid_gather, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N, c->n2e_ptr,
                       c->n2e, c->n2b_ptr, c->n2b, c->ER, c->ER + 9 * (size_t)c->E, c->old_head, c->u, c->v, c->R);
}

// ------------------------------------------------------------------ Jacobian
// EM[e*81 + (i*3+j)*9 + r*3 + var] = d elem_rhs[i].eq_r / d var_j, column
// q = 3*j + var (var 0 u, 1 v, 2 h).
// The same over a list of triangles (`count` entries), two lanes per column:
// the even lane evaluates the +eps residual and the odd lane the -eps one, so
// each lane carries one residual evaluation; the pair forms the central
// difference with a lane exchange.  WET lists hold triangles whose depths stay
// positive under every perturbation and run without the wet/dry clipping code.
// OCC > 0 asks the compiler for at least OCC waves per SIMD (fewer registers,
// possibly spills); 0 leaves the register budget to the compiler.
template <bool WET, int BS, int OCC>
__global__ void __launch_bounds__(BS, OCC > 0 ? OCC : 1) k_jac2d_list(ElemData D, const int *list, const int *count,
                                                                      int b0, int b1, int b2, const int *blkpos,
                                                                      double *EM, float *EM32) {
    const int nt = 18 * *count;
    const int stride = gridDim.x * blockDim.x;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    // Grid-stride over (triangle, column, sign); the stride is a multiple of
    // the wavefront, so both lanes of a pair run and leave together.
    // Software pipeline: the triangle id two iterations ahead and the node ids
    // one iteration ahead are loaded while the current triangle is computed,
    // so an iteration waits on one memory round trip (the node data) instead
    // of three dependent ones (list -> triangle -> nodes).
    int ie_a = t < nt ? list[t / 18] : 0;
    int ie_b = t + stride < nt ? list[(t + stride) / 18] : 0;
    int nd_a[3] = {0, 0, 0};
    if (t < nt)
        for (int k = 0; k < 3; k++) nd_a[k] = D.e2d[3 * ie_a + k];
    for (; t < nt; t += stride) {
    const int ie = ie_a;
    const int nodes[3] = {nd_a[0], nd_a[1], nd_a[2]};
    ie_a = ie_b;
    ie_b = t + 2 * stride < nt ? list[(t + 2 * stride) / 18] : 0;
    if (t + stride < nt)
        for (int k = 0; k < 3; k++) nd_a[k] = D.e2d[3 * ie_a + k];
    const int q = (t >> 1) % 9, j = q / 3, var = q % 3;
    const int sgn = (t & 1) ? -1 : +1;
    const int n = nodes[j];
    const double val = var == 0 ? D.u[n] : (var == 1 ? D.v[n] : D.head[n]);
    const double eps = num_diff_eps(val), eps2 = 2.0 * eps;
    const int pv = var == 0 ? PERTURB_U : (var == 1 ? PERTURB_V : PERTURB_H);
    d3 r[3];
    body_resid<WET>(D, D.pre + PRE_STRIDE * (size_t)ie, ie, eps, j, pv, sgn, false, r, nodes);
    double col[9];
    const double inv = 1.0 / eps2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // even lanes hold P (+eps), odd lanes M (-eps); both lanes form P - M
        double mx = __shfl_xor(r[i].x_eq, 1, 64), my = __shfl_xor(r[i].y_eq, 1, 64),
               mc = __shfl_xor(r[i].c_eq, 1, 64);
        col[3 * i + 0] = div_rn(r[i].x_eq - mx, eps2, inv);
        col[3 * i + 1] = div_rn(r[i].y_eq - my, eps2, inv);
        col[3 * i + 2] = div_rn(r[i].c_eq - mc, eps2, inv);
    }
    if (t & 1) continue;
    const int *lbi = D.loc_bc_inst ? D.loc_bc_inst + 3 * D.elem_inst[ie] : nullptr;
    const int lb[3] = {lbi ? lbi[0] : b0, lbi ? lbi[1] : b1, lbi ? lbi[2] : b2};
    jac_column_db<3>(D, nodes, lb, j, var, col);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const size_t o = 9 * (size_t)blkpos[9 * ie + 3 * i + j];
#pragma unroll
        for (int rr = 0; rr < 3; rr++) {
            if (EM32) EM32[o + 3 * rr + var] = (float)col[3 * i + rr];
            else EM[o + 3 * rr + var] = col[3 * i + rr];
        }
    }
    }
}

// The same Jacobian columns in a column-uniform layout: wave w takes column
// q = w % 9 (node j = q / 3, variable q % 3) of the 32 triangles of group
// w / 9, lanes 0-31 evaluating the +eps residuals and lanes 32-63 the -eps
// ones of the same triangles.  The perturbed node and variable are then the
// same across the wave (scalar registers), so the residual's branches on them
// are taken or skipped by whole waves -- the u and v columns skip the
// pressure block (fe_sw2_body_resid.c:761-862 runs for depth perturbations
// only) instead of running it under a lane mask -- and each triangle's data
// is loaded by 64 lanes of 32 different triangles rather than broadcast to 18
// lanes of one.  Every (triangle, column, sign) evaluates exactly what the
// lane-per-column kernel evaluates, so the blocks are bit-identical.
template <bool WET, int BS>
__global__ void __launch_bounds__(BS) k_jac2d_cols(ElemData D, const int *list, const int *count, int b0, int b1,
                                                   int b2, const int *blkpos, double *EM, float *EM32) {
    constexpr int WPB = BS / 64;
    const int n = *count;
    const int nunits = 9 * ((n + 31) >> 5);
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool minus = lane >= 32;
    for (int w = blockIdx.x * WPB + wib; w < nunits; w += gridDim.x * WPB) {
        const int q = w % 9, grp = w / 9;
        const int j = q / 3, var = q % 3;
        const int li = grp * 32 + (lane & 31);
        const bool act = li < n;
        // idle lanes of a group's last wave recompute its first triangle and store nothing
        const int ie = list[act ? li : grp * 32];
        const int nodes[3] = {D.e2d[3 * ie], D.e2d[3 * ie + 1], D.e2d[3 * ie + 2]};
        const int nd = pick3(nodes, j);
        const double val = var == 0 ? D.u[nd] : (var == 1 ? D.v[nd] : D.head[nd]);
        const double eps = num_diff_eps(val), eps2 = 2.0 * eps;
        const int pv = var == 0 ? PERTURB_U : (var == 1 ? PERTURB_V : PERTURB_H);
        d3 r[3];
        body_resid<WET>(D, D.pre + PRE_STRIDE * (size_t)ie, ie, eps, j, pv, minus ? -1 : +1, false, r, nodes);
        double col[9];
        const double inv = 1.0 / eps2;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            // lanes l and l ^ 32 hold the +eps and -eps residuals of one triangle
            double mx = __shfl_xor(r[i].x_eq, 32, 64), my = __shfl_xor(r[i].y_eq, 32, 64),
                   mc = __shfl_xor(r[i].c_eq, 32, 64);
            col[3 * i + 0] = div_rn(r[i].x_eq - mx, eps2, inv);
            col[3 * i + 1] = div_rn(r[i].y_eq - my, eps2, inv);
            col[3 * i + 2] = div_rn(r[i].c_eq - mc, eps2, inv);
        }
        if (minus || !act) continue;
        const int *lbi = D.loc_bc_inst ? D.loc_bc_inst + 3 * D.elem_inst[ie] : nullptr;
        const int lb[3] = {lbi ? lbi[0] : b0, lbi ? lbi[1] : b1, lbi ? lbi[2] : b2};
        jac_column_db<3>(D, nodes, lb, j, var, col);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const size_t o = 9 * (size_t)blkpos[9 * ie + 3 * i + j];
#pragma unroll
            for (int rr = 0; rr < 3; rr++) {
                if (EM32) EM32[o + 3 * rr + var] = (float)col[3 * i + rr];
                else EM[o + 3 * rr + var] = col[3 * i + rr];
            }
        }
    }
}

// The column-uniform layout with every triangle's inputs staged in LDS
// first.  Each wave copies the inputs of its 32 triangles -- node
// coordinates, depths and velocities, P1 gradients, the 24 old-state terms,
// djac, roughness, rain, material, wet/dry flag and the nodes' boundary flags
// -- with 29 loads per lane issued back to back (lanes l and l + 32 share a
// triangle and take alternate fields), waits once, and then evaluates the
// residuals reading those inputs from LDS.  In the global-memory layouts
// many of these loads sit deep in the element code, each waited for on its
// own; here they cost one round trip per triangle group.  Same arithmetic,
// same results.
constexpr int LREC = 58;   // doubles per staged triangle (integers held as exact doubles)
enum : int { LR_NODE = 0, LR_GRAD = 18, LR_PRE = 24, LR_DJAC = 48, LR_KS = 49, LR_SRC = 50, LR_MAT = 51,
             LR_WD = 52, LR_SRCF = 53, LR_NBC = 54, LR_ROUGH0 = 57 };

// R0: the record also carries the unperturbed drag coefficient (LR_ROUGH0),
// which the residual then uses for u and v perturbations
template <bool R0>
struct LdsSrcT {
    static constexpr bool has_rough0 = R0;
    const double *r;   // the triangle's staged record (LDS)
    __device__ __forceinline__ double rough0() const { return r[LR_ROUGH0]; }
    __device__ __forceinline__ double x(int i) const { return r[6 * i]; }
    __device__ __forceinline__ double y(int i) const { return r[6 * i + 1]; }
    __device__ __forceinline__ double z(int i) const { return r[6 * i + 2]; }
    __device__ __forceinline__ double head(int i) const { return r[6 * i + 3]; }
    __device__ __forceinline__ double u(int i) const { return r[6 * i + 4]; }
    __device__ __forceinline__ double v(int i) const { return r[6 * i + 5]; }
    __device__ __forceinline__ int node_bc(int i) const { return (int)r[LR_NBC + i]; }
    __device__ __forceinline__ double grad(int k) const { return r[LR_GRAD + k]; }
    __device__ __forceinline__ double djac() const { return r[LR_DJAC]; }
    __device__ __forceinline__ int mat() const { return (int)r[LR_MAT]; }
    __device__ __forceinline__ int wd_flag() const { return (int)r[LR_WD]; }
    __device__ __forceinline__ double ks() const { return r[LR_KS]; }
    __device__ __forceinline__ int src_flag() const { return (int)r[LR_SRCF]; }
    __device__ __forceinline__ double src() const { return r[LR_SRC]; }
    __device__ __forceinline__ double pre(int k) const { return r[LR_PRE + k]; }
};
using LdsSrc = LdsSrcT<false>;

// The triangle's 57 inputs into rec: lane half 0 takes one field of each
// pair, half 1 the other (both fields of a step come from one array where
// possible, so a step is one load instruction for the whole wave).
__device__ __forceinline__ void stage_triangle(const ElemData &D, int ie, const int *nd, int half, double *rec) {
    double val[29];
    int fld[29];
    const int na = half ? nd[1] : nd[0];
    // node fields of nodes 0 and 1
    const double *narr[6] = {D.x, D.y, D.z, D.head, D.u, D.v};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        val[c] = narr[c][na];
        fld[c] = (half ? 6 : 0) + c;
    }
    // node 2: (x, y), (z, h), (u, v)
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double *a = half ? narr[2 * c + 1] : narr[2 * c];
        val[6 + c] = a[nd[2]];
        fld[6 + c] = 12 + 2 * c + half;
    }
    // gradients and old-state terms
#pragma unroll
    for (int m = 0; m < 3; m++) {
        val[9 + m] = D.grad[6 * (size_t)ie + 2 * m + half];
        fld[9 + m] = LR_GRAD + 2 * m + half;
    }
    const double *pre = D.pre + PRE_STRIDE * (size_t)ie;
#pragma unroll
    for (int m = 0; m < 12; m++) {
        val[12 + m] = pre[2 * m + half];
        fld[12 + m] = LR_PRE + 2 * m + half;
    }
    val[24] = half ? D.ks[ie] : D.djac[ie];
    fld[24] = half ? LR_KS : LR_DJAC;
    val[25] = D.src[ie];
    fld[25] = LR_SRC;   // both halves write the same value
    val[26] = (double)(half ? D.wd_flag[ie] : D.mat[ie]);
    fld[26] = half ? LR_WD : LR_MAT;
    val[27] = (double)(half ? D.node_bc[nd[0]] : D.src_flag[ie]);
    fld[27] = half ? LR_NBC : LR_SRCF;
    val[28] = (double)D.node_bc[half ? nd[2] : nd[1]];
    fld[28] = LR_NBC + 1 + half;
#pragma unroll
    for (int k = 0; k < 29; k++) rec[fld[k]] = val[k];
}

// One Jacobian column (node j, variable var) of the staged triangle: the
// +eps residual on lanes 0-31, the -eps one on lanes 32-63, the difference
// quotient, the Dirichlet edits, and the stores of the +eps lanes.
template <bool WET, typename SRC>
__device__ __forceinline__ void lds_column(const ElemData &DL, const SRC &S, int ie, const int *nodes, int j,
                                           int var, bool minus, bool act, int b0, int b1, int b2,
                                           const int *blkpos, double *EM, float *EM32) {
    const double val = var == 0 ? S.u(j) : (var == 1 ? S.v(j) : S.head(j));
    const double eps = num_diff_eps(val), eps2 = 2.0 * eps;
    const int pv = var == 0 ? PERTURB_U : (var == 1 ? PERTURB_V : PERTURB_H);
    d3 r[3];
    body_resid_src<WET>(DL, S, ie, eps, j, pv, minus ? -1 : +1, false, r, nodes);
    double col[9];
    const double inv = 1.0 / eps2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double mx = __shfl_xor(r[i].x_eq, 32, 64), my = __shfl_xor(r[i].y_eq, 32, 64),
               mc = __shfl_xor(r[i].c_eq, 32, 64);
        col[3 * i + 0] = div_rn(r[i].x_eq - mx, eps2, inv);
        col[3 * i + 1] = div_rn(r[i].y_eq - my, eps2, inv);
        col[3 * i + 2] = div_rn(r[i].c_eq - mc, eps2, inv);
    }
    if (minus || !act) return;
    const int *lbi = DL.loc_bc_inst ? DL.loc_bc_inst + 3 * DL.elem_inst[ie] : nullptr;
    const int lb[3] = {lbi ? lbi[0] : b0, lbi ? lbi[1] : b1, lbi ? lbi[2] : b2};
    jac_column_db<3>(DL, nodes, lb, j, var, col);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const size_t o = 9 * (size_t)blkpos[9 * ie + 3 * i + j];
#pragma unroll
        for (int rr = 0; rr < 3; rr++) {
            if (EM32) EM32[o + 3 * rr + var] = (float)col[3 * i + rr];
            else EM[o + 3 * rr + var] = col[3 * i + rr];
        }
    }
}

// GROUP = false: wave w takes column w % 9 of group w / 9 (as the
// column-uniform layout) and stages its group's inputs for that one column.
// GROUP = true: wave w takes all nine columns of group w, one after the
// other, from one staging of the group's inputs; the nine column stores of a
// triangle then come from one wave, so its element blocks are completed in
// one L2 rather than in partial lines from several XCDs.
template <bool WET, int BS, bool GROUP>
__global__ void __launch_bounds__(BS, 2) k_jac2d_lds(ElemData D, const int *list, const int *count, int b0, int b1,
                                                  int b2, const int *blkpos, double *EM, float *EM32) {
    constexpr int WPB = BS / 64;
    __shared__ double srec[WPB][32][LREC];
    __shared__ crlog::tables slog;
    const int n = *count;
    const int ngroups = (n + 31) >> 5;
    const int nunits = GROUP ? ngroups : 9 * ngroups;
    if ((int)blockIdx.x * WPB >= nunits) return;   // the whole block has no unit (uniform exit)
    // the friction law's log table (4 KB) copied to LDS: its lookup is the
    // one data-dependent load of the element code
    {
        const double *src = (const double *)D.logt;
        double *dst = (double *)&slog;
        for (int i = threadIdx.x; i < (int)(sizeof(crlog::tables) / sizeof(double)); i += BS) dst[i] = src[i];
        __syncthreads();
    }
    ElemData DL = D;
    DL.logt = &slog;
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool minus = lane >= 32;
    double *rec = &srec[wib][lane & 31][0];
    const int wstride = gridDim.x * WPB;
    // software pipeline: the triangle id two units ahead and its node ids one
    // unit ahead are loaded while the current unit is computed
    auto tri_of = [&](int w) {
        const int g = GROUP ? w : w / 9, l = g * 32 + (lane & 31);
        return list[l < n ? l : g * 32];
    };
    int w = blockIdx.x * WPB + wib;
    int ie_a = w < nunits ? tri_of(w) : 0;
    int ie_b = w + wstride < nunits ? tri_of(w + wstride) : 0;
    int nd_a[3] = {0, 0, 0};
    if (w < nunits)
        for (int k = 0; k < 3; k++) nd_a[k] = D.e2d[3 * ie_a + k];
    for (; w < nunits; w += wstride) {
        const int grp = GROUP ? w : w / 9;
        const int li = grp * 32 + (lane & 31);
        const bool act = li < n;   // idle lanes of a last group recompute its first triangle, store nothing
        const int ie = ie_a;
        const int nodes[3] = {nd_a[0], nd_a[1], nd_a[2]};
        ie_a = ie_b;
        ie_b = w + 2 * wstride < nunits ? tri_of(w + 2 * wstride) : 0;
        if (w + wstride < nunits)
            for (int k = 0; k < 3; k++) nd_a[k] = D.e2d[3 * ie_a + k];
        // the previous group's reads are done before the records are rewritten
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        stage_triangle(D, ie, nodes, minus ? 1 : 0, rec);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if constexpr (GROUP) {
            // the unperturbed drag coefficient, once for the six u and v columns
            const double rough0 = unperturbed_roughness<WET>(DL, LdsSrc{rec});
            if (!minus) rec[LR_ROUGH0] = rough0;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        const LdsSrcT<GROUP> S{rec};
        if constexpr (GROUP) {
#pragma unroll 1
            for (int q = 0; q < 9; q++) {
                // keeps the record's reads inside the column loop (hoisted
                // out of it they would hold the whole record in registers)
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                lds_column<WET>(DL, S, ie, nodes, q / 3, q % 3, minus, act, b0, b1, b2, blkpos, EM, EM32);
            }
        } else {
            const int q = w % 9;
            lds_column<WET>(DL, S, ie, nodes, q / 3, q % 3, minus, act, b0, b1, b2, blkpos, EM, EM32);
        }
    }
}

__global__ void __launch_bounds__(128) k_jac1d(ElemData D, int B, int b0, int b1, const int *blkpos1, double *EM,
                                               float *EM32) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * B) return;
    int ie = t / 6, q = t % 6, j = q / 3, var = q % 3;
    double col[9];
    jac_column<2>(D, ie, j, var, col);
    const int nodes[2] = {D.e1d[2 * ie], D.e1d[2 * ie + 1]};
    const int *lbi = D.loc_bc_inst ? D.loc_bc_inst + 3 * D.edge_inst[ie] : nullptr;
    const int lb[2] = {lbi ? lbi[0] : b0, lbi ? lbi[1] : b1};
    jac_column_db<2>(D, nodes, lb, j, var, col);
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const size_t o = 9 * (size_t)blkpos1[4 * ie + 2 * i + j];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            if (EM32) EM32[o + 3 * r + var] = (float)col[3 * i + r];
            else EM[o + 3 * r + var] = col[3 * i + r];
        }
    }
}

// Splits the triangles into those whose three depths stay positive under the
// -eps perturbation of the FD Jacobian (h - eps(h) > 0, the exact operation
// the perturbed residual performs) and the rest, as two ordered lists: block
// counts, a one-block scan, then the ordered scatter (no same-address
// atomics, which serialise across the whole grid).
__device__ __forceinline__ bool tri_stays_wet(int ie, const int *e2d, const double *head) {
    bool w = true;
    for (int i = 0; i < 3; i++) {
        const double h = head[e2d[3 * ie + i]];
        if (!(h + (-1.0) * num_diff_eps(h) > 0.0)) w = false;
    }
    return w;
}

// Triangle classes for the element kernels: "wet" (stays wet under the -eps
// perturbation: the variant without clipping code), "general", or skipped
// (an ensemble model that is not being iterated).  bcount holds the per-block
// wet counts, then the per-block general counts.
__device__ __forceinline__ void split_class(int ie, int E, const int *e2d, const double *head, const int *elem_inst,
                                            const int *mask, bool &w, bool &g) {
    const bool act = ie < E && (!mask || mask[elem_inst[ie]]);
    w = act && tri_stays_wet(ie, e2d, head);
    g = act && !w;
}

__global__ void __launch_bounds__(256) k_split_count(int E, const int *e2d, const double *head, const int *elem_inst,
                                                     const int *mask, int *bcount) {
    __shared__ int wc[2][4];
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    bool w, g;
    split_class(ie, E, e2d, head, elem_inst, mask, w, g);
    const unsigned long long bw = __ballot(w), bg = __ballot(g);
    if ((threadIdx.x & 63) == 0) {
        wc[0][threadIdx.x >> 6] = __popcll(bw);
        wc[1][threadIdx.x >> 6] = __popcll(bg);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bcount[blockIdx.x] = wc[0][0] + wc[0][1] + wc[0][2] + wc[0][3];
        bcount[gridDim.x + blockIdx.x] = wc[1][0] + wc[1][1] + wc[1][2] + wc[1][3];
    }
}

// exclusive scans of the two block-count arrays (one block of 1024 threads,
// each owning a contiguous run of up to SCAN_R counts of both arrays, held in
// registers from one load to the write-back, so the kernel waits on memory
// once); counts[0] = wet total, counts[1] = general total.  Longer runs
// (more than 8,192 count blocks, i.e. over 2M triangles) loop over chunks.
constexpr int SCAN_T = 1024, SCAN_R = 8;
__device__ __forceinline__ void scan_pair(int s0, int s1, int &run0, int &run1, int &tot0, int &tot1) {
    __shared__ int wsum[2][SCAN_T / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x0 = s0, x1 = s1;   // inclusive scans within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int y0 = __shfl_up(x0, o, 64), y1 = __shfl_up(x1, o, 64);
        if (lane >= o) { x0 += y0; x1 += y1; }
    }
    if (lane == 63) { wsum[0][wid] = x0; wsum[1][wid] = x1; }
    __syncthreads();
    run0 = x0 - s0; run1 = x1 - s1; tot0 = 0; tot1 = 0;
    for (int w = 0; w < SCAN_T / 64; w++) {
        if (w < wid) { run0 += wsum[0][w]; run1 += wsum[1][w]; }
        tot0 += wsum[0][w];
        tot1 += wsum[1][w];
    }
}

__global__ void __launch_bounds__(SCAN_T) k_split_scan(int nb, int *bcount, int *counts) {
    int *bc0 = bcount, *bc1 = bcount + nb;
    const int per = (nb + SCAN_T - 1) / SCAN_T, b0 = min(nb, (int)threadIdx.x * per), b1 = min(nb, b0 + per);
    int run0, run1, tot0, tot1;
    if (per <= SCAN_R) {
        int v0[SCAN_R], v1[SCAN_R], s0 = 0, s1 = 0;
#pragma unroll
        for (int u = 0; u < SCAN_R; u++) {
            v0[u] = b0 + u < b1 ? bc0[b0 + u] : 0;
            v1[u] = b0 + u < b1 ? bc1[b0 + u] : 0;
        }
#pragma unroll
        for (int u = 0; u < SCAN_R; u++) { s0 += v0[u]; s1 += v1[u]; }
        scan_pair(s0, s1, run0, run1, tot0, tot1);
#pragma unroll
        for (int u = 0; u < SCAN_R; u++)
            if (b0 + u < b1) {
                bc0[b0 + u] = run0; run0 += v0[u];
                bc1[b0 + u] = run1; run1 += v1[u];
            }
    } else {
        int s0 = 0, s1 = 0;
        for (int b = b0; b < b1; b += SCAN_R) {
            int v0[SCAN_R], v1[SCAN_R];
#pragma unroll
            for (int u = 0; u < SCAN_R; u++) {
                v0[u] = b + u < b1 ? bc0[b + u] : 0;
                v1[u] = b + u < b1 ? bc1[b + u] : 0;
            }
#pragma unroll
            for (int u = 0; u < SCAN_R; u++) { s0 += v0[u]; s1 += v1[u]; }
        }
        scan_pair(s0, s1, run0, run1, tot0, tot1);
        for (int b = b0; b < b1; b += SCAN_R) {
            int v0[SCAN_R], v1[SCAN_R];
#pragma unroll
            for (int u = 0; u < SCAN_R; u++) {
                v0[u] = b + u < b1 ? bc0[b + u] : 0;
                v1[u] = b + u < b1 ? bc1[b + u] : 0;
            }
#pragma unroll
            for (int u = 0; u < SCAN_R; u++)
                if (b + u < b1) {
                    bc0[b + u] = run0; run0 += v0[u];
                    bc1[b + u] = run1; run1 += v1[u];
                }
        }
    }
    if (threadIdx.x == 0) { counts[0] = tot0; counts[1] = tot1; }
}

__global__ void __launch_bounds__(256) k_split_write(int E, const int *e2d, const double *head, const int *elem_inst,
                                                     const int *mask, const int *boff, int *wet, int *gen) {
    __shared__ int wc[2][4];
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    bool w, g;
    split_class(ie, E, e2d, head, elem_inst, mask, w, g);
    const unsigned long long bw = __ballot(w), bg = __ballot(g);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        wc[0][wid] = __popcll(bw);
        wc[1][wid] = __popcll(bg);
    }
    __syncthreads();
    int bef_w = 0, bef_g = 0;  // triangles of each class in earlier waves of this block
    for (int k = 0; k < wid; k++) { bef_w += wc[0][k]; bef_g += wc[1][k]; }
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (w) wet[boff[blockIdx.x] + bef_w + __popcll(bw & below)] = ie;
    if (g) gen[boff[gridDim.x + blockIdx.x] + bef_g + __popcll(bg & below)] = ie;
}

// ordered wet / general triangle lists for the current depths (and, for
// ensembles, the models being iterated)
static void launch_split(Ctx *c) {
    if (!c->E) return;
    const int nb = nblocks_for(c->E, 256);
    const int *mask = c->ninst ? c->inst_mask : nullptr;
    hipLaunchKernelGGL(k_split_count, dim3(nb), dim3(256), 0, c->stream, c->E, c->e2d, c->head, c->elem_inst, mask,
                       c->esplit);
    hipLaunchKernelGGL(k_split_scan, dim3(1), dim3(SCAN_T), 0, c->stream, nb, c->esplit, c->ecount);
    hipLaunchKernelGGL(k_split_write, dim3(nb), dim3(256), 0, c->stream, c->E, c->e2d, c->head, c->elem_inst, mask,
                       c->esplit, c->elist_wet, c->elist_gen);
    c->split_valid = 1;
}

// Grid-stride list launches.  Block size and grid size change only the
// schedule (every triangle's arithmetic and output slot are fixed), so they
// are free parameters: defaults below, or what adh_sw2_hip_autotune measured.
template <bool WET, int BS>
static void jac_list_bs(Ctx *c, const ElemData &D, const int *list, const int *count, int grid, int layout) {
    if (layout == 2) {   // column-uniform waves, inputs staged in LDS
        const int units = 9 * ((c->E + 31) / 32);
        hipLaunchKernelGGL((k_jac2d_lds<WET, BS, false>), dim3(std::min(grid, nblocks_for(units, BS / 64))), dim3(BS),
                           0, c->stream, D, list, count, c->loc_bc[0], c->loc_bc[1], c->loc_bc[2], c->blkpos, c->EM,
                           em32_of(c));
        return;
    }
    if (layout == 3) {   // a wave per triangle group, its nine columns from one staging
        const int units = (c->E + 31) / 32;
        hipLaunchKernelGGL((k_jac2d_lds<WET, BS, true>), dim3(std::min(grid, nblocks_for(units, BS / 64))), dim3(BS),
                           0, c->stream, D, list, count, c->loc_bc[0], c->loc_bc[1], c->loc_bc[2], c->blkpos, c->EM,
                           em32_of(c));
        return;
    }
    if (layout == 1) {   // column-uniform waves: 9 waves per 32 triangles
        const int units = 9 * ((c->E + 31) / 32);
        hipLaunchKernelGGL((k_jac2d_cols<WET, BS>), dim3(std::min(grid, nblocks_for(units, BS / 64))), dim3(BS), 0,
                           c->stream, D, list, count, c->loc_bc[0], c->loc_bc[1], c->loc_bc[2], c->blkpos, c->EM,
                           em32_of(c));
        return;
    }
    hipLaunchKernelGGL((k_jac2d_list<WET, BS, 0>), dim3(std::min(grid, nblocks_for(18 * c->E, BS))), dim3(BS), 0,
                       c->stream, D, list, count, c->loc_bc[0], c->loc_bc[1], c->loc_bc[2], c->blkpos, c->EM,
                       em32_of(c));
}
template <bool WET, int BS>
static void res_list_bs(Ctx *c, const ElemData &D, const int *list, const int *count, int grid) {
    hipLaunchKernelGGL((k_resid2d_list<WET, BS>), dim3(std::min(grid, nblocks_for(c->E, BS))), dim3(BS), 0, c->stream,
                       D, list, count, c->ER);
}

// phase: 0 Jacobian wet list, 1 Jacobian general list, 2 residual wet list, 3 residual general list
void launch_list(Ctx *c, const ElemData &D, int phase, int bs, int grid, int layout) {
    const bool wet = phase == 0 || phase == 2;
    const int *list = wet ? c->elist_wet : c->elist_gen;
    const int *count = wet ? c->ecount : c->ecount + 1;
#define SW2_LIST_CASE(B)                                                                      \
    case B:                                                                                    \
        if (phase == 0) jac_list_bs<true, B>(c, D, list, count, grid, layout);                    \
        else if (phase == 1) jac_list_bs<false, B>(c, D, list, count, grid, layout);              \
        else if (phase == 2) res_list_bs<true, B>(c, D, list, count, grid);                    \
        else res_list_bs<false, B>(c, D, list, count, grid);                                   \
        break;
    switch (bs) {
        SW2_LIST_CASE(64)
        SW2_LIST_CASE(256)
        default:
        SW2_LIST_CASE(128)
    }
#undef SW2_LIST_CASE
}

void ensure_lists(Ctx *c) {
    if (!c->pre_valid) launch_old_terms(c);
    if (!c->split_valid) launch_split(c);
}

// one list kernel with the element data its phase uses (autotuning)
void launch_list_phase(Ctx *c, int phase, int bs, int grid, int layout) {
    ElemData D = c->ed;
    D.first_elem = phase >= 2 ? c->first_elem : nullptr;
    launch_list(c, D, phase, bs, grid, layout);
}

static void launch_jac2d_split(Ctx *c, const ElemData &D) {
    if (!c->E) return;
    if (!c->split_valid) launch_split(c);
    launch_list(c, D, 0, c->tune.bs[0], c->tune.grid[0], c->tune.layout[0]);
    launch_list(c, D, 1, c->tune.bs[1], c->tune.grid[1], c->tune.layout[1]);
}

static void launch_resid2d_split(Ctx *c, const ElemData &D) {
    if (!c->E) return;
    if (!c->split_valid) launch_split(c);
    launch_list(c, D, 2, c->tune.bs[2], c->tune.grid[2], c->tune.layout[2]);
    launch_list(c, D, 3, c->tune.bs[3], c->tune.grid[3], c->tune.layout[3]);
}

void launch_jacobian(Ctx *c) {
    if (!c->pre_valid) launch_old_terms(c);
    ElemData D = c->ed;
    D.first_elem = nullptr;
    c->em32 = em32_of(c) != nullptr;
    launch_jac2d_split(c, D);
    if (c->B)
        hipLaunchKernelGGL(k_jac1d, dim3(nblocks_for(6 * c->B, 128)), dim3(128), 0, c->stream, D, c->B, c->loc_bc[0],
                           c->loc_bc[1], c->blkpos + 9 * (size_t)c->E, c->EM, em32_of(c));
}

// Ordered block assembly.  The element kernels store each 3x3 block at its
// contributor position (slot order; within a slot the contributors in
// ascending element order, 2D before 1D), so a block's sum reads one
// contiguous run.  Slots are the N diagonal blocks by node, then one per SELL
// position.  A workgroup takes one 64-row slice: wave 0 the diagonal blocks,
// wave w >= 1 the row slots j = w - 1 (+ stride), so every wave reads one
// contiguous stretch of element blocks and writes 64 consecutive sliced
// entries.  The diagonal block gets the floor and residual tweak of
// fe_sw2_load.c:218-236; the waves' row maxima, combined in LDS, give the
// Jacobi scale of solv_linear_sys_setup (the same values as k_scale_rows,
// which the set-up then skips).
template <typename TM>
__device__ __forceinline__ void sum_run(const int *slot_ptr, size_t slot, const TM *EM, double (&a)[9]) {
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = 0.0;
    const int p0 = slot_ptr[slot], p1 = slot_ptr[slot + 1];
    // pairs of blocks: both loaded before the in-order sums
    int p = p0;
    for (; p + 1 < p1; p += 2) {
        const TM *b0 = EM + 9 * (size_t)p, *b1 = b0 + 9;
        double x0[9], x1[9];
#pragma unroll
        for (int k = 0; k < 9; k++) { x0[k] = (double)b0[k]; x1[k] = (double)b1[k]; }
#pragma unroll
        for (int k = 0; k < 9; k++) a[k] += x0[k];
#pragma unroll
        for (int k = 0; k < 9; k++) a[k] += x1[k];
    }
    if (p < p1) {
        const TM *b = EM + 9 * (size_t)p;
#pragma unroll
        for (int k = 0; k < 9; k++) a[k] += (double)b[k];
    }
}

// TM: element-block type; TO: off-diagonal output type (float for the refined
// forms, which keep only an FP32 matrix: the set-up scales it in place).
template <typename TM, typename TO>
__global__ void __launch_bounds__(1024) k_assemble(int N, const int *slot_ptr, const int *row_ptr,
                                                   const int *sell_ptr, const TM *EM, double *diag, TO *offd,
                                                   double *R, const double *u, const double *v, double *sv) {
    __shared__ double smax[16][3][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int n = 64 * blockIdx.x + lane;
    double s[3] = {0.0, 0.0, 0.0};
    double a[9];
    if (n < N) {
        if (w == 0) {
            sum_run(slot_ptr, (size_t)n, EM, a);
            const double f = NOT_QUITE_SMALL;
            if (fabs(a[0]) < f) {
                a[0] = f;
                double &r0 = R[vidx3(n, 0)];
                if (fabs(r0) < f) r0 = -f * u[n];
            }
            if (fabs(a[4]) < f) {
                a[4] = f;
                double &r1 = R[vidx3(n, 1)];
                if (fabs(r1) < f) r1 = -f * v[n];
            }
            if (fabs(a[8]) < f) a[8] = f;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                diag[didx9(n, k)] = a[k];
                const double x = fabs(a[k]);
                if (x > s[k / 3]) s[k / 3] = x;
            }
        } else {
            const size_t q0 = 64 * (size_t)sell_ptr[n >> 6] + lane;
            const int len = row_ptr[n + 1] - row_ptr[n];
            for (int j = w - 1; j < len; j += nw - 1) {
                const size_t q = q0 + 64 * (size_t)j;
                sum_run(slot_ptr, (size_t)N + q, EM, a);
#pragma unroll
                for (int k = 0; k < 9; k++) {
                    offd[sell_vidx(q, k)] = (TO)a[k];
                    const double x = fabs(a[k]);
                    if (x > s[k / 3]) s[k / 3] = x;
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) smax[w][j][lane] = s[j];
    __syncthreads();
    if (w != 0 || n >= N) return;
    const double small_factor = 1.0 / sqrt(SOLV_TOL);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        double m = smax[0][j][lane];
        for (int x = 1; x < nw; x++) {
            const double y = smax[x][j][lane];
            if (y > m) m = y;
        }
        sv[vidx3(n, j)] = m < SOLV_TOL ? small_factor : 1.0 / sqrt(fabs(m));
    }
}

void launch_assemble(Ctx *c) {
    const int nw = 1 + std::min(std::max(c->kmax_row, 1), 15);   // waves per 64-row slice
    // refined forms with block Jacobi: element blocks stored in FP32, summed in
    // FP64, off-diagonal blocks written in FP32 (unscaled) straight into the
    // matrix the FP32 Krylov reads; otherwise FP64 throughout
    c->mat32 = c->em32 && c->opts.preconditioner != 1;
    dim3 g(nblocks_for(c->N, 64)), b(64 * nw);
    if (c->mat32)
        hipLaunchKernelGGL((k_assemble<float, float>), g, b, 0, c->stream, c->N, c->slot_ptr, c->row_ptr, c->sell_ptr,
                           (const float *)c->EM, c->diag, c->offd32, c->R, c->u, c->v, c->scale);
    else if (c->em32)
        hipLaunchKernelGGL((k_assemble<float, double>), g, b, 0, c->stream, c->N, c->slot_ptr, c->row_ptr,
                           c->sell_ptr, (const float *)c->EM, c->diag, c->offd, c->R, c->u, c->v, c->scale);
    else
        hipLaunchKernelGGL((k_assemble<double, double>), g, b, 0, c->stream, c->N, c->slot_ptr, c->row_ptr,
                           c->sell_ptr, (const double *)c->EM, c->diag, c->offd, c->R, c->u, c->v, c->scale);
    c->scale_valid = 1;
    c->sys_scaled = 0;
}

void launch_jac2d_only(Ctx *c) {
    if (!c->pre_valid) launch_old_terms(c);
    ElemData D = c->ed;
    D.first_elem = nullptr;
    launch_jac2d_split(c, D);
}

void launch_resid2d_only(Ctx *c) {
    if (!c->pre_valid) launch_old_terms(c);
    ElemData D = c->ed;
    D.first_elem = c->first_elem;
    launch_resid2d_split(c, D);
}

void launch_assemble_only(Ctx *c) { launch_assemble(c); }

// ------------------------------------------------------- init / wet-dry flags
// fe_sw2_init (fe_sw2_init.c:33-79)
__global__ void k_init_nodes(int N, const double *old_head, const double *old_u, const double *old_v, double *head,
                             double *u, double *v, const int *node_bc, const double *bcval, int *bc_mask,
                             int *wd_tmp, const int *node_inst, const int *mask) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (mask && (node_inst[n] < 0 || !mask[node_inst[n]])) return;  // gap nodes (-1) never change
    double hh = old_head[n], uu = old_u[n], vv = old_v[n];
    int m0 = ADH_SW2_NO, m1 = ADH_SW2_NO, m2 = ADH_SW2_NO;
    int bc = node_bc[n];
    if (bc == BCT_VEL_DIR) {
        m0 = m1 = ADH_SW2_YES;
        uu = bcval[3 * n]; vv = bcval[3 * n + 1];
    } else if (bc == BCT_PRS_DIR) {
        m2 = ADH_SW2_YES;
        hh = bcval[3 * n + 2];
    } else if (bc == BCT_VEL_PRS_DIR) {
        m0 = m1 = m2 = ADH_SW2_YES;
        uu = bcval[3 * n]; vv = bcval[3 * n + 1]; hh = bcval[3 * n + 2];
    }
    head[n] = hh; u[n] = uu; v[n] = vv;
    if (bc_mask) { bc_mask[3 * n] = m0; bc_mask[3 * n + 1] = m1; bc_mask[3 * n + 2] = m2; }
    wd_tmp[n] = 0;
}

// fe_sw2_wdflag pass 1 (fe_sw2_wdflag.c:41-51): all writers store the same value
__global__ void k_wd_mark(int E, const int *e2d, const double *head, int *wd_tmp, const int *elem_inst,
                          const int *mask) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    if (mask && !mask[elem_inst[e]]) return;
    int a = e2d[3 * e], b = e2d[3 * e + 1], c = e2d[3 * e + 2];
    int wet = (head[a] > 0.0) + (head[b] > 0.0) + (head[c] > 0.0);
    if (wet != 3) { wd_tmp[a] = 1; wd_tmp[b] = 1; wd_tmp[c] = 1; }
}

// fe_sw2_wdflag pass 2 (:57-73)
__global__ void k_wd_flag(int E, const int *e2d, const double *head, const double *old_head, const int *wd_tmp,
                          int *wd_flag, const int *elem_inst, const int *mask) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    if (mask && !mask[elem_inst[e]]) return;
    int a = e2d[3 * e], b = e2d[3 * e + 1], c = e2d[3 * e + 2];
    int f = (wd_tmp[a] != 0 || wd_tmp[b] != 0 || wd_tmp[c] != 0) ? 1 : 0;
    if (head[a] < 0 && head[b] < 0 && head[c] < 0 && old_head[a] < 0 && old_head[b] < 0 && old_head[c] < 0) f = 2;
    wd_flag[e] = f;
}

// old/older-state element terms, once per Newton solve (sw2_elem.h elem_old_terms)
__global__ void __launch_bounds__(128) k_old_terms(ElemData D, int E, double *pre) {
    int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= E) return;
    elem_old_terms(D, ie, pre + PRE_STRIDE * (size_t)ie);
}

void launch_old_terms(Ctx *c) {
    if (c->E)
        hipLaunchKernelGGL(k_old_terms, dim3(nblocks_for(c->E, 128)), dim3(128), 0, c->stream, c->ed, c->E, c->pre);
    c->pre_valid = 1;
}


void launch_init(Ctx *c, int *bc_mask_dev) {
    c->split_valid = 0;
    const int *mask = c->ninst ? c->inst_mask : nullptr;
    hipLaunchKernelGGL(k_init_nodes, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N, c->old_head,
                       c->old_u, c->old_v, c->head, c->u, c->v, c->node_bc, c->node_bc_val, bc_mask_dev, c->wd_tmp,
                       c->node_inst, mask);
    if (c->E) {
        hipLaunchKernelGGL(k_wd_mark, dim3(nblocks_for(c->E, 256)), dim3(256), 0, c->stream, c->E, c->e2d, c->head,
                           c->wd_tmp, c->elem_inst, mask);
        hipLaunchKernelGGL(k_wd_flag, dim3(nblocks_for(c->E, 256)), dim3(256), 0, c->stream, c->E, c->e2d, c->head,
                           c->old_head, c->wd_tmp, c->wd_flag, c->elem_inst, mask);
    }
    launch_old_terms(c);
}

// fe_sw2 (fe_sw2.c:42-50)
__global__ void k_shift(int N, double *h, double *u, double *v, double *oh, double *ou, double *ov, double *ooh,
                        double *oou, double *oov, const int *node_inst, const int *mask) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (mask && (node_inst[n] < 0 || !mask[node_inst[n]])) return;  // gap nodes (-1) never change
    ooh[n] = oh[n]; oou[n] = ou[n]; oov[n] = ov[n];
    oh[n] = h[n]; ou[n] = u[n]; ov[n] = v[n];
}

void launch_shift(Ctx *c) {
    c->pre_valid = 0;
    c->split_valid = 0;
    hipLaunchKernelGGL(k_shift, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N, c->head, c->u, c->v,
                       c->old_head, c->old_u, c->old_v, c->older_head, c->older_u, c->older_v, c->node_inst,
                       c->ninst ? c->inst_mask : nullptr);
}

// fe_sw2_inc (fe_sw2_inc.c:47-54): factor 1.0
__global__ void k_inc(int N, const double *sol, double *h, double *u, double *v, const int *node_inst,
                      const int *mask) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (mask && (node_inst[n] < 0 || !mask[node_inst[n]])) return;  // gap nodes (-1) never change
    u[n] += 1. * sol[vidx3(n, 0)];
    v[n] += 1. * sol[vidx3(n, 1)];
    h[n] += 1. * sol[vidx3(n, 2)];
}

void launch_inc(Ctx *c, double) {
    c->split_valid = 0;
    hipLaunchKernelGGL(k_inc, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N, c->sol, c->head, c->u,
                       c->v, c->node_inst, c->ninst ? c->inst_mask : nullptr);
}

// fe_sw2_inc, then sol = -sol (what the line search steps back along), in one
// pass over the nodes: the same operations as k_inc followed by k_scale_vec(-1)
__global__ void k_inc_flip(int N, double *sol, double *h, double *u, double *v, const int *node_inst,
                           const int *mask) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (mask && (node_inst[n] < 0 || !mask[node_inst[n]])) return;  // gap nodes (-1) never change
    const double s0 = sol[vidx3(n, 0)], s1 = sol[vidx3(n, 1)], s2 = sol[vidx3(n, 2)];
    u[n] += 1. * s0;
    v[n] += 1. * s1;
    h[n] += 1. * s2;
    sol[vidx3(n, 0)] = s0 * -1.0;
    sol[vidx3(n, 1)] = s1 * -1.0;
    sol[vidx3(n, 2)] = s2 * -1.0;
}

void launch_inc_flip(Ctx *c) {
    c->split_valid = 0;
    hipLaunchKernelGGL(k_inc_flip, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N, c->sol, c->head, c->u,
                       c->v, c->node_inst, c->ninst ? c->inst_mask : nullptr);
}

// x *= f over a sliced node vector; entry i belongs to node 64 (i / 192) + i mod 64
__global__ void k_scale_vec(size_t n, double *x, double f, const int *node_inst, const int *mask) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask) {
        const int m = node_inst[64 * (i / 192) + (i & 63)];
        if (m < 0 || !mask[m]) return;
    }
    x[i] *= f;
}

void launch_scale_sol(Ctx *c, double f) {
    hipLaunchKernelGGL(k_scale_vec, dim3(nblocks_for(3 * c->Np, 256)), dim3(256), 0, c->stream, (size_t)3 * c->Np,
                       c->sol, f, c->node_inst, c->ninst ? c->inst_mask : nullptr);
}

// plain streaming copy of n doubles (bandwidth calibration for the counters)
__global__ void __launch_bounds__(256) k_copy(size_t n, const double *__restrict__ src, double *__restrict__ dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

void launch_copy(Ctx *c, const double *src, double *dst, size_t n) {
    if (n) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream, n, src, dst);
}

// grid-stride: the tiled LU's bands exceed 2^31 doubles (4.7e9 at 1M triangles)
__global__ void k_zero(size_t n, double *x) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = 0.0;
}

void launch_zero(Ctx *c, double *p, size_t n) {
    const size_t blocks = std::max<size_t>(1, std::min<size_t>((n + 255) / 256, (size_t)1 << 20));
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, c->stream, n, p);
}

// rounding probe: device cr_log, sqrt and division of each input
__global__ void k_probe(const double *x, int n, const crlog::tables *t, double *lg, double *sq, double *dv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = x[i];
    lg[i] = crlog::cr_log_t(a, t);
    sq[i] = sqrt(a);
    dv[i] = 1.0 / a;
}

void launch_probe(const double *x, int n, const crlog::tables *t, double *lg, double *sq, double *dv) {
    hipLaunchKernelGGL(k_probe, dim3(nblocks_for(n, 256)), dim3(256), 0, 0, x, n, t, lg, sq, dv);
}

// ------------------------------------------------------------------- norms
// get_residual_norms (fe_newton.c:79-150) with include_node_in_residual
// (:177-187): max norms over nodes with min(h, h_old) > 0, l2 over all DOFs.
// Block partials, then a one-block finisher reduces them in block order.
__global__ void __launch_bounds__(256) k_norms(int N, const double *R, const double *sol, const double *h,
                                               const double *oh, double *part) {
    __shared__ double smax[256], ssum[256], sinc[256];
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0, s = 0.0, in = 0.0;
    if (n < N) {
        bool inc = dmin(h[n], oh[n]) > 0;
        for (int j = 0; j < 3; j++) {
            double r = R[vidx3(n, j)], q = sol[vidx3(n, j)];
            if (inc && fabs(r) > m) m = fabs(r);
            if (inc && fabs(q) > in) in = fabs(q);
            s += r * r;
        }
    }
    smax[threadIdx.x] = m; ssum[threadIdx.x] = s; sinc[threadIdx.x] = in;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            sinc[threadIdx.x] = fmax(sinc[threadIdx.x], sinc[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = smax[0];
        part[3 * blockIdx.x + 1] = ssum[0];
        part[3 * blockIdx.x + 2] = sinc[0];
    }
}

// one block: the block partials of k_norms, reduced in block order
// one block per model (blockIdx.x), over that model's 256-node blocks
__global__ void __launch_bounds__(256) k_norms_fin(int nparts, int N, const double *part, Norms *out,
                                                   const int *inst_blk0, const int *inst_nodes) {
    __shared__ double smax[256], ssum[256], sinc[256];
    const int m = blockIdx.x;
    const int k0 = inst_blk0 ? inst_blk0[m] : 0, k1 = inst_blk0 ? inst_blk0[m + 1] : nparts;
    if (inst_nodes) N = inst_nodes[m];
    out += m;
    double a = 0.0, b = 0.0, cc = 0.0;
    // 8 strided partials per thread loaded together, folded in block order
    const int bd = blockDim.x;
    for (int k = k0 + threadIdx.x; k < k1; k += 8 * bd) {
        double x[8][3];
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
            for (int e = 0; e < 3; e++) x[u][e] = k + u * bd < k1 ? part[3 * (k + u * bd) + e] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (k + u * bd < k1) {
                a = fmax(a, x[u][0]);
                b += x[u][1];
                cc = fmax(cc, x[u][2]);
            }
    }
    smax[threadIdx.x] = a; ssum[threadIdx.x] = b; sinc[threadIdx.x] = cc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            sinc[threadIdx.x] = fmax(sinc[threadIdx.x], sinc[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out->maxn = smax[0];
        out->l2 = sqrt(ssum[0] / (double)(N * 3));
        out->incn = sinc[0];
    }
}

void launch_norms(Ctx *c) {
    int nb = nblocks_for(c->N, 256);
    hipLaunchKernelGGL(k_norms, dim3(nb), dim3(256), 0, c->stream, c->N, c->R, c->sol, c->head, c->old_head,
                       c->red.part);
    hipLaunchKernelGGL(k_norms_fin, dim3(c->ninst ? c->ninst : 1), dim3(256), 0, c->stream, nb, c->N, c->red.part,
                       c->norms, c->ninst ? c->inst_blk0 : nullptr, c->ninst ? c->inst_nodes : nullptr);
}

}  // namespace sw2

// Block ILU(0) preconditioner in multicolour order (preconditioner = 3).
//
// No counterpart in the reference's serial path, which preconditions with a
// skyline LU of node blocks (solv_prec.c:11-66 -> solv_blk_solve.c); this is
// the north star's "block-Jacobi/ILU(0)" alternative, plugged in at the same
// place: M^-1 is applied to a vector in place after each SpMV of BiCGSTAB.
//
// Ordering.  The nodes are coloured greedily in index order (each node takes
// the smallest colour none of its lower-index neighbours has), so nodes of
// one colour never share a matrix block.  Rows are ordered by colour, then
// node index; each colour's rows are padded to a multiple of 64 so a
// wavefront never spans two colours.  In this order every row's off-diagonal
// blocks split into "lower" (neighbours of smaller colour) and "upper".
//
// Factorisation (IKJ variant of ILU(0) on 3x3 blocks, no fill): for row i,
// for each lower neighbour k in increasing position,
//     L_ik = A_ik U_kk^-1,   A_ij -= L_ik U_kj  for j in row i with pos(j) > pos(k)
// (j = i included), then U_ii = A_ii and U_ij = A_ij for the upper j.  Rows
// of one colour depend only on rows of smaller colours, so each colour is
// one launch with a thread per row; the (k, j) update lists are built once
// on the host.  The factors are formed in FP64 from the scaled system and
// stored in FP64 (precision 0) or FP32 (precision 1).
//
// Apply (in place on a node-ordered vector): forward sweeps z_i = v_i -
// sum L_ik z_k over the colours in order, backward sweeps
// x_i = U_ii^-1 (z_i - sum U_ij x_j) in reverse order, one launch per colour
// and sweep; the colour-ordered work vector w holds z, then x.
//
// Storage: like the SpMV's SELL-64 blocks, slices of 64 consecutive rows (in
// colour order) with a per-slice slot count; value e of slot b of the row in
// lane l sits at 9*64*(ptr[s] + b) + 64*e + l.  Padding slots have column -1.
#include <climits>

#include <algorithm>
#include <numeric>

#include "sw2_ctx.h"

namespace sw2 {

// y = A x for 3x3 blocks (row-major), accumulated in FP64
template <typename T>
__device__ __forceinline__ void blk_mul(const T *a, size_t st, const double *x, double *y) {
#pragma unroll
    for (int r = 0; r < 3; r++)
        y[r] = (double)a[st * (3 * r)] * x[0] + (double)a[st * (3 * r + 1)] * x[1] + (double)a[st * (3 * r + 2)] * x[2];
}

struct IluView {
    int Rp;                   // rows in colour order, padded
    const int *rnode;         // [Rp] node of each row (-1: padding)
    const int *lptr, *uptr;   // [Rp/64 + 1] slot offsets per slice
    const int *lcol, *ucol;   // [64 * slots] row position of each slot's column (-1: padding)
};

__device__ __forceinline__ size_t ilu_v(size_t slot_base, int e, int lane) { return 9 * 64 * slot_base + 64 * e + lane; }
__device__ __forceinline__ size_t ilu_w(int r, int k) { return 3 * (size_t)(r & ~63) + 64 * k + (r & 63); }

// Copies the scaled system into the factor work arrays (FP64).
template <typename MT>
__global__ void __launch_bounds__(256) k_ilu_load(IluView V, const int *lsrc, const int *usrc, const MT *diag,
                                                  const MT *offd, double *Lw, double *Uw, double *Dw) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= V.Rp) return;
    const int node = V.rnode[r];
    const int s = r >> 6, lane = r & 63;
    for (int e = 0; e < 9; e++) Dw[ilu_v(s, e, lane)] = node >= 0 ? (double)diag[didx9(node, e)] : (e % 4 == 0 ? 1.0 : 0.0);
    for (int b = V.lptr[s]; b < V.lptr[s + 1]; b++) {
        const int q = lsrc[64 * (size_t)b + lane];
        for (int e = 0; e < 9; e++) Lw[ilu_v(b, e, lane)] = q >= 0 ? (double)offd[sell_vidx((size_t)q, e)] : 0.0;
    }
    for (int b = V.uptr[s]; b < V.uptr[s + 1]; b++) {
        const int q = usrc[64 * (size_t)b + lane];
        for (int e = 0; e < 9; e++) Uw[ilu_v(b, e, lane)] = q >= 0 ? (double)offd[sell_vidx((size_t)q, e)] : 0.0;
    }
}

// One colour of the factorisation.  upd[tptr[r] .. tptr[r+1]) lists, in the
// order of the row's lower slots, (lower slot b, U slot of row k holding
// U_kj, target): target >= 0 an L slot of row r, -1 - t a U slot t of row r,
// and INT_MIN the diagonal.  Dinv holds U_kk^-1 of the rows already factored.
__global__ void __launch_bounds__(256) k_ilu_factor(IluView V, int r0, int r1, const int *tptr, const int4 *upd,
                                                    double *Lw, double *Uw, double *Dw, double *Dinv) {
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    const int s = r >> 6, lane = r & 63;
    int t = tptr[r];
    const int t1 = tptr[r + 1];
    if (V.rnode[r] >= 0) {
        for (int b = V.lptr[s]; b < V.lptr[s + 1]; b++) {
            const int k = V.lcol[64 * (size_t)b + lane];
            if (k < 0) continue;
            // L_ik = A_ik U_kk^-1
            double a[9], l[9];
            for (int e = 0; e < 9; e++) a[e] = Lw[ilu_v(b, e, lane)];
            const size_t dk = ilu_v(k >> 6, 0, k & 63);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    l[3 * i + j] = a[3 * i] * Dinv[dk + 64 * j] + a[3 * i + 1] * Dinv[dk + 64 * (3 + j)] +
                                   a[3 * i + 2] * Dinv[dk + 64 * (6 + j)];
            for (int e = 0; e < 9; e++) Lw[ilu_v(b, e, lane)] = l[e];
            // A_ij -= L_ik U_kj
            for (; t < t1 && upd[t].x == b; t++) {
                const int4 u = upd[t];
                const size_t src = ilu_v((size_t)u.y >> 6, 0, u.y & 63);
                double *dst;
                if (u.z == INT_MIN) dst = Dw + ilu_v(s, 0, lane);
                else if (u.z >= 0) dst = Lw + ilu_v(u.z, 0, lane);
                else dst = Uw + ilu_v(-1 - u.z, 0, lane);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        dst[64 * (3 * i + j)] -= l[3 * i] * Uw[src + 64 * j] + l[3 * i + 1] * Uw[src + 64 * (3 + j)] +
                                                 l[3 * i + 2] * Uw[src + 64 * (6 + j)];
            }
        }
    }
    double d[9], inv[9];
    for (int e = 0; e < 9; e++) d[e] = Dw[ilu_v(s, e, lane)];
    inv3(d, inv);
    for (int e = 0; e < 9; e++) Dinv[ilu_v(s, e, lane)] = inv[e];
}

template <typename FT>
__global__ void k_ilu_store(size_t n, const double *a, FT *b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (FT)a[i];
}

// forward sweep of one colour: w_r = v_node - sum_b L_b w_col
template <typename FT>
__global__ void __launch_bounds__(256) k_ilu_fwd(IluView V, int r0, int r1, const FT *L, const double *v, double *w,
                                                 const KScal *sc) {
    if (sc && sc->done) return;
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    const int node = V.rnode[r];
    if (node < 0) return;
    const int s = r >> 6, lane = r & 63;
    double z[3] = {v[vidx3(node, 0)], v[vidx3(node, 1)], v[vidx3(node, 2)]};
    for (int b = V.lptr[s]; b < V.lptr[s + 1]; b++) {
        const int k = V.lcol[64 * (size_t)b + lane];
        if (k < 0) continue;
        const double x[3] = {w[ilu_w(k, 0)], w[ilu_w(k, 1)], w[ilu_w(k, 2)]};
        double y[3];
        blk_mul(L + ilu_v(b, 0, lane), 64, x, y);
        z[0] -= y[0]; z[1] -= y[1]; z[2] -= y[2];
    }
    for (int j = 0; j < 3; j++) w[ilu_w(r, j)] = z[j];
}

// backward sweep of one colour: x_r = U_rr^-1 (w_r - sum_b U_b x_col), written
// to w and to the node-ordered output
template <typename FT>
__global__ void __launch_bounds__(256) k_ilu_bwd(IluView V, int r0, int r1, const FT *U, const FT *Dinv, double *w,
                                                 double *out, const KScal *sc) {
    if (sc && sc->done) return;
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    const int node = V.rnode[r];
    if (node < 0) return;
    const int s = r >> 6, lane = r & 63;
    double z[3] = {w[ilu_w(r, 0)], w[ilu_w(r, 1)], w[ilu_w(r, 2)]};
    for (int b = V.uptr[s]; b < V.uptr[s + 1]; b++) {
        const int k = V.ucol[64 * (size_t)b + lane];
        if (k < 0) continue;
        const double x[3] = {w[ilu_w(k, 0)], w[ilu_w(k, 1)], w[ilu_w(k, 2)]};
        double y[3];
        blk_mul(U + ilu_v(b, 0, lane), 64, x, y);
        z[0] -= y[0]; z[1] -= y[1]; z[2] -= y[2];
    }
    double x[3];
    blk_mul(Dinv + ilu_v(s, 0, lane), 64, z, x);
    for (int j = 0; j < 3; j++) {
        w[ilu_w(r, j)] = x[j];
        out[vidx3(node, j)] = x[j];
    }
}

static IluView iview(Ctx *c) { return IluView{c->ilu_Rp, c->ilu_rnode, c->ilu_lptr, c->ilu_uptr, c->ilu_lcol, c->ilu_ucol}; }

void ilu_free(Ctx *c) {
    for (void *p : c->ilu_allocs) (void)hipFree(p);
    c->ilu_allocs.clear();
    c->ilu_ready = 0;
    c->ilu_ncolor = 0;
}

// Greedy colouring, colour-ordered rows, the lower/upper slot structure and
// the update lists, built once per pattern.  Returns an error text or "".
std::string ilu_structure(Ctx *c) {
    ilu_free(c);
    const int N = c->N;
    const std::vector<int> &rp = c->h_row_ptr, &cl = c->h_col;
    std::vector<int> color(N, -1);
    int ncolor = 0;
    {
        std::vector<int> seen;
        for (int i = 0; i < N; i++) {
            seen.assign(ncolor + 2, 0);
            for (int p = rp[i]; p < rp[i + 1]; p++)
                if (cl[p] < i) seen[color[cl[p]]] = 1;
            int k = 0;
            while (seen[k]) k++;
            color[i] = k;
            ncolor = std::max(ncolor, k + 1);
        }
    }
    // rows: colour-major, node order within a colour, each colour padded to 64
    std::vector<int> cstart(ncolor + 1, 0), count(ncolor, 0);
    for (int i = 0; i < N; i++) count[color[i]]++;
    for (int k = 0; k < ncolor; k++) cstart[k + 1] = cstart[k] + ((count[k] + 63) & ~63);
    const int Rp = cstart[ncolor];
    std::vector<int> rnode(Rp, -1), pos(N), fill(ncolor, 0);
    for (int i = 0; i < N; i++) {
        const int r = cstart[color[i]] + fill[color[i]]++;
        rnode[r] = i;
        pos[i] = r;
    }
    // per-row lower / upper neighbours sorted by position, with their SELL source slots
    const int nsl = Rp / 64;
    std::vector<std::vector<std::pair<int, int>>> lo(Rp), up(Rp);   // (position, SELL slot)
    for (int r = 0; r < Rp; r++) {
        const int i = rnode[r];
        if (i < 0) continue;
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            const int pj = pos[cl[p]];
            (pj < r ? lo[r] : up[r]).push_back({pj, c->h_offd_pos[p]});
        }
        std::sort(lo[r].begin(), lo[r].end());
        std::sort(up[r].begin(), up[r].end());
    }
    std::vector<int> lptr(nsl + 1, 0), uptr(nsl + 1, 0);
    for (int s = 0; s < nsl; s++) {
        size_t kl = 0, ku = 0;
        for (int l = 0; l < 64; l++) {
            kl = std::max(kl, lo[64 * s + l].size());
            ku = std::max(ku, up[64 * s + l].size());
        }
        lptr[s + 1] = lptr[s] + (int)kl;
        uptr[s + 1] = uptr[s] + (int)ku;
    }
    const size_t nl = 64 * (size_t)lptr[nsl], nu = 64 * (size_t)uptr[nsl];
    std::vector<int> lcol(std::max<size_t>(nl, 1), -1), ucol(std::max<size_t>(nu, 1), -1);
    std::vector<int> lsrc(std::max<size_t>(nl, 1), -1), usrc(std::max<size_t>(nu, 1), -1);
    // slot index of (row, b) for the upper part, and lookup of a row's upper slot by column position
    for (int r = 0; r < Rp; r++) {
        const int s = r >> 6, l = r & 63;
        for (size_t b = 0; b < lo[r].size(); b++) {
            lcol[64 * (size_t)(lptr[s] + b) + l] = lo[r][b].first;
            lsrc[64 * (size_t)(lptr[s] + b) + l] = lo[r][b].second;
        }
        for (size_t b = 0; b < up[r].size(); b++) {
            ucol[64 * (size_t)(uptr[s] + b) + l] = up[r][b].first;
            usrc[64 * (size_t)(uptr[s] + b) + l] = up[r][b].second;
        }
    }
    auto uslot = [&](int r, int pj) -> int {   // global U slot of row r's block in column pj, or -1
        const int s = r >> 6, l = r & 63;
        for (size_t b = 0; b < up[r].size(); b++)
            if (up[r][b].first == pj) return 64 * (uptr[s] + (int)b) + l;
        return -1;
    };
    std::vector<int> tptr(Rp + 1, 0);
    std::vector<int4> upd;
    for (int r = 0; r < Rp; r++) {
        const int s = r >> 6, l = r & 63;
        for (size_t bk = 0; bk < lo[r].size(); bk++) {
            const int k = lo[r][bk].first;
            const int bslot = lptr[s] + (int)bk;   // slot base index (times 64 plus lane = storage slot)
            // j over row r's later lower slots, the diagonal, and the upper slots
            for (size_t bj = bk + 1; bj < lo[r].size(); bj++) {
                const int q = uslot(k, lo[r][bj].first);
                if (q >= 0) upd.push_back(make_int4(bslot, q, lptr[s] + (int)bj, 0));
            }
            {
                const int q = uslot(k, r);
                if (q >= 0) upd.push_back(make_int4(bslot, q, INT_MIN, 0));
            }
            for (size_t bj = 0; bj < up[r].size(); bj++) {
                const int q = uslot(k, up[r][bj].first);
                if (q >= 0) upd.push_back(make_int4(bslot, q, -1 - (uptr[s] + (int)bj), 0));
            }
        }
        tptr[r + 1] = (int)upd.size();
    }
    // the kernel addresses a slot's values by (slot base, lane): convert the
    // U source slots from storage index to (base * 64 + lane) form, which
    // they already are (64 * base + lane); L/U targets are slot bases
    auto up_dev = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (hipMalloc(dst, bytes ? bytes : 8) != hipSuccess) return false;
        c->ilu_allocs.push_back(*dst);
        return bytes == 0 || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    auto alloc = [&](void **dst, size_t bytes) -> bool {
        if (hipMalloc(dst, bytes ? bytes : 8) != hipSuccess) return false;
        c->ilu_allocs.push_back(*dst);
        return true;
    };
    const size_t fsz = c->opts.precision >= 1 ? sizeof(float) : sizeof(double);
    bool ok = up_dev((void **)&c->ilu_rnode, rnode.data(), sizeof(int) * Rp) &&
              up_dev((void **)&c->ilu_lptr, lptr.data(), sizeof(int) * lptr.size()) &&
              up_dev((void **)&c->ilu_uptr, uptr.data(), sizeof(int) * uptr.size()) &&
              up_dev((void **)&c->ilu_lcol, lcol.data(), sizeof(int) * lcol.size()) &&
              up_dev((void **)&c->ilu_ucol, ucol.data(), sizeof(int) * ucol.size()) &&
              up_dev((void **)&c->ilu_lsrc, lsrc.data(), sizeof(int) * lsrc.size()) &&
              up_dev((void **)&c->ilu_usrc, usrc.data(), sizeof(int) * usrc.size()) &&
              up_dev((void **)&c->ilu_tptr, tptr.data(), sizeof(int) * tptr.size()) &&
              up_dev((void **)&c->ilu_upd, upd.data(), sizeof(int4) * upd.size()) &&
              alloc((void **)&c->ilu_Lw, sizeof(double) * 9 * std::max<size_t>(nl, 1)) &&
              alloc((void **)&c->ilu_Uw, sizeof(double) * 9 * std::max<size_t>(nu, 1)) &&
              alloc((void **)&c->ilu_Dw, sizeof(double) * 9 * (size_t)Rp) &&
              alloc((void **)&c->ilu_Dinv, sizeof(double) * 9 * (size_t)Rp) &&
              alloc((void **)&c->ilu_w, sizeof(double) * 3 * (size_t)Rp) &&
              alloc((void **)&c->ilu_L, fsz * 9 * std::max<size_t>(nl, 1)) &&
              alloc((void **)&c->ilu_U, fsz * 9 * std::max<size_t>(nu, 1)) &&
              alloc((void **)&c->ilu_D, fsz * 9 * (size_t)Rp);
    if (!ok) {
        ilu_free(c);
        return "hipMalloc failed for the ILU(0) factors";
    }
    c->ilu_ncolor = ncolor;
    c->h_ilu_cstart = cstart;
    c->h_ilu_color = color;
    c->ilu_Rp = Rp;
    c->ilu_nl = nl;
    c->ilu_nu = nu;
    c->ilu_f32 = c->opts.precision >= 1;
    c->ilu_ready = 1;
    return "";
}

// Loads the scaled system, factors colour by colour, stores the factors in
// the apply's precision.
void launch_ilu_setup(Ctx *c) {
    IluView V = iview(c);
    const int nb = nblocks_for(c->ilu_Rp, 256);
    if (c->ilu_f32)
        hipLaunchKernelGGL(k_ilu_load<float>, dim3(nb), dim3(256), 0, c->stream, V, c->ilu_lsrc, c->ilu_usrc,
                           c->diag32, c->offd32, c->ilu_Lw, c->ilu_Uw, c->ilu_Dw);
    else
        hipLaunchKernelGGL(k_ilu_load<double>, dim3(nb), dim3(256), 0, c->stream, V, c->ilu_lsrc, c->ilu_usrc,
                           c->diag, c->offd, c->ilu_Lw, c->ilu_Uw, c->ilu_Dw);
    for (int k = 0; k < c->ilu_ncolor; k++) {
        const int r0 = c->h_ilu_cstart[k], r1 = c->h_ilu_cstart[k + 1];
        hipLaunchKernelGGL(k_ilu_factor, dim3(nblocks_for(r1 - r0, 256)), dim3(256), 0, c->stream, V, r0, r1,
                           c->ilu_tptr, c->ilu_upd, c->ilu_Lw, c->ilu_Uw, c->ilu_Dw, c->ilu_Dinv);
    }
    const size_t nl9 = 9 * c->ilu_nl, nu9 = 9 * c->ilu_nu, nd9 = 9 * (size_t)c->ilu_Rp;
    if (c->ilu_f32) {
        hipLaunchKernelGGL(k_ilu_store<float>, dim3(nblocks_for((int)nl9, 256)), dim3(256), 0, c->stream, nl9, c->ilu_Lw, (float *)c->ilu_L);
        hipLaunchKernelGGL(k_ilu_store<float>, dim3(nblocks_for((int)nu9, 256)), dim3(256), 0, c->stream, nu9, c->ilu_Uw, (float *)c->ilu_U);
        hipLaunchKernelGGL(k_ilu_store<float>, dim3(nblocks_for((int)nd9, 256)), dim3(256), 0, c->stream, nd9, c->ilu_Dinv, (float *)c->ilu_D);
    } else {
        hipLaunchKernelGGL(k_ilu_store<double>, dim3(nblocks_for((int)nl9, 256)), dim3(256), 0, c->stream, nl9, c->ilu_Lw, (double *)c->ilu_L);
        hipLaunchKernelGGL(k_ilu_store<double>, dim3(nblocks_for((int)nu9, 256)), dim3(256), 0, c->stream, nu9, c->ilu_Uw, (double *)c->ilu_U);
        hipLaunchKernelGGL(k_ilu_store<double>, dim3(nblocks_for((int)nd9, 256)), dim3(256), 0, c->stream, nd9, c->ilu_Dinv, (double *)c->ilu_D);
    }
}

template <typename FT>
static void ilu_sweeps(Ctx *c, double *x, const KScal *sc) {
    IluView V = iview(c);
    const FT *L = (const FT *)c->ilu_L, *U = (const FT *)c->ilu_U, *D = (const FT *)c->ilu_D;
    for (int k = 0; k < c->ilu_ncolor; k++) {
        const int r0 = c->h_ilu_cstart[k], r1 = c->h_ilu_cstart[k + 1];
        hipLaunchKernelGGL(k_ilu_fwd<FT>, dim3(nblocks_for(r1 - r0, 256)), dim3(256), 0, c->stream, V, r0, r1, L, x,
                           c->ilu_w, sc);
    }
    for (int k = c->ilu_ncolor - 1; k >= 0; k--) {
        const int r0 = c->h_ilu_cstart[k], r1 = c->h_ilu_cstart[k + 1];
        hipLaunchKernelGGL(k_ilu_bwd<FT>, dim3(nblocks_for(r1 - r0, 256)), dim3(256), 0, c->stream, V, r0, r1, U, D,
                           c->ilu_w, x, sc);
    }
}

// x <- M^-1 x (node-ordered, sliced layout); skipped once sc->done is set
void launch_ilu_solve(Ctx *c, double *x, const KScal *sc) {
    if (c->ilu_f32) ilu_sweeps<float>(c, x, sc);
    else ilu_sweeps<double>(c, x, sc);
}

}  // namespace sw2

// Internal state of one device-resident SW2 model and the launch helpers
// shared by the kernel and API translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/adh_sw2_hip.h"
#include "sw2_elem.h"

namespace sw2 {

constexpr double SOLV_TOL = 2.220446049250313e-16;  // SMALL, src/include/define.h:177 (Jacobi scale floor)
// 18 * triangles (the FD Jacobian's lane count) plus two grid strides must fit an int
constexpr long long MAX_TRIANGLES = 115000000LL;
constexpr int KUNI_MAX = 10;                              // longest uniform slice of the unrolled SpMV
constexpr int KGRAPH_MAX = 32;                            // longest captured Krylov chunk
constexpr int BAND_MAX_DOF = 8190;                        // LDS of one block solve (64 KB)
constexpr size_t BAND_MAX_BYTES = (size_t)8 << 30;       // band storage limit

// Device-side scalars of one BiCGSTAB solve (solv_bcgstab.c:60-70).
struct KScal {
    double rho, rhop, alpha, omega, beta, gamma;
    double rnorm, bnorm, tol;
    double atol, rtol;
    int it, max_it, done, pad;
};

// Reduction slots: per-block partials and completion counters.
struct Red {
    double *part;      // [3 * maxblocks]
    unsigned *count;   // [8] counters, one per reduction site
    int maxblocks;
};

// Newton norms (get_residual_norms, fe_newton.c:40-151)
struct Norms {
    double maxn, l2, incn, pad;
};

// One node block's band for the tiled band LU (sw2_lu.hip)
struct TBand {
    int n0;        // first node
    int m;         // DOFs
    int w;         // half bandwidth of the pattern (DOFs)
    int wp;        // storage half bandwidth: w + 64
    int nt;        // 64-DOF tiles (panels)
    int pad;       // always 0 (a sweep outside the Krylov loop reads it as its done flag)
    int inst;      // model of an ensemble (0 for a single model)
    int pad2;
    size_t off;    // first double of the band
    size_t toff;   // first tile in the per-tile inverse arrays
    size_t xoff;   // first DOF in the band-ordered vectors
};

struct Ctx {
    int N, E, B, M, nnz;
    int Np;                // N rounded up to a multiple of 64 (node-slice padding)
    adh_sw2_opts opts;
    double dt;
    hipStream_t stream;
    hipEvent_t ev[8];

    // node fields
    double *x, *y, *z;
    double *head, *u, *v, *old_head, *old_u, *old_v, *older_head, *older_u, *older_v;
    int *node_bc;
    double *node_bc_val;
    int *first_elem;
    int *wd_tmp;
    int loc_bc[3];   // node-string flags of nodes 0..2 (the reference's local-index read)
    // elements
    int *e2d, *mat, *src_flag, *wd_flag;
    double *djac, *grad, *ks, *src;
    int *e1d, *bc1d;
    double *djac1d, *nrml1d, *val1d;
    int *mat_turb, *mat_eev_mode;
    double *mat_eev_coef, *mat_ev, *mat_cor;
    double *wind, *mat_windatt;   // node wind values [2N] and attenuation [M] (models with winds)
    crlog::tables *logt;
    // adjacency
    int *n2e_ptr, *n2e;    // node -> (element*3 + local), ascending element
    int *n2b_ptr, *n2b;    // node -> (edge*2 + local)
    // pattern + assembly
    int *row_ptr, *col;    // off-diagonal blocks, insertion order
    int *slot_ptr, *slot_src;  // N diagonal slots (node order) then one slot per SELL position
    int scale_valid;           // the assembly computed `scale` for the current matrix
    int em32;                  // EM holds FP32 blocks (the last Jacobian ran in a refined form)
    int mat32;                 // the assembled off-diagonal blocks are in offd32 (unscaled until set-up)
    int sys_scaled;            // the set-up has scaled the current system in place (cleared by load/set_system)
    int *blkpos;               // element block (9e+3i+j, then 9E+4b+2i+j) -> contributor position
    std::vector<int> h_blkpos;
    std::vector<int> h_row_ptr, h_col;
    // Off-diagonal blocks in a wave-sliced ELL layout (SELL-64): block rows
    // 64s..64s+63 form slice s with K_s = sell_ptr[s+1]-sell_ptr[s] block
    // slots per row (rows shorter than K_s are padded with zero blocks whose
    // column is the row itself).  Slot q = 64*(sell_ptr[s] + j) + lane holds
    // the j-th block of row 64s+lane (the reference's insertion order);
    // its 9 values sit at sell_vidx(q, e) so that a wavefront reading entry e
    // of slot j touches 64 consecutive values.
    int nslice;
    int kuni;              // > 0: every slice holds kuni slots (sell_ptr[s] = s * kuni)
    int kmax_row;          // longest row (off-diagonal blocks)
    size_t nsell;          // 64 * sell_ptr[nslice]
    int *sell_ptr, *sell_col;   // [nslice+1], [nsell]
    std::vector<int> h_offd_pos;  // CSR off-diagonal block -> SELL position
    size_t ndslots;               // N diagonal slots + one slot per SELL position
    // work
    int *elist_wet, *elist_gen, *ecount;  // FD-Jacobian triangle lists [E], [E], counts [2]
    int *esplit;                          // per-block counts / offsets of the split [E/256 + 1]
    int split_valid;                      // the lists match the current depths
    // launch shape of the four list kernels (Jacobian wet/general, residual
    // wet/general): block size and grid-stride grid size
    struct Tuning {
        int bs[4] = {256, 256, 128, 128};
        int grid[4] = {4096, 1024, 2048, 512};
        // Jacobian lists: 0 a lane per (triangle, column, sign), 1 column-uniform
        // waves, 2 column-uniform waves reading their triangles' inputs from LDS,
        // 3 a wave per 32-triangle group doing its nine columns from one staging
        int layout[4] = {3, 3, 0, 0};
    } tune;
    double *pre;           // [PRE_STRIDE * E] old/older-state element terms
    int pre_valid;         // pre matches the current old/older state
    double *ER;            // element residuals [9E + 6B]
    double *EM;            // element 3x3 blocks [81E + 36B], in slot-ordered contributor positions
    // Node vectors and 3x3 node blocks use the same wave-sliced layout as
    // the off-diagonal blocks: nodes 64s..64s+63 form slice s, and component
    // k of node n sits at vidx3(n, k) (entry e of a block at didx9(n, e)), so
    // a wavefront touching one component of its 64 nodes reads 64
    // consecutive values.
    double *R;             // residual / rhs [3 Np]
    double *sol;           // [3 Np]
    double *diag, *offd;   // [9 Np] sliced, [9 nsell] SELL-64 values
    double *scale;         // [3 Np]
    double *dinv;          // [9 Np]
    // FP32 system for mixed precision
    float *diag32, *offd32, *dinv32;
    // Krylov vectors (FP64) and FP32 twins
    double *kb, *kr, *kp, *kAp, *ks_, *kAs, *kq, *ku;
    float *fb, *fr, *fp, *fAp, *fs, *fAs, *fq, *fu;
    KScal *ksc;            // [max(1, ninst)]
    KScal *h_ksc;          // pinned, [max(1, ninst)]
    Red red;
    Norms *norms, *h_norms;  // [max(1, ninst)]
    int *h_flag;           // pinned scratch
    ElemData ed;
    // Ensemble of independent models sharing one mesh ("batch"): model i owns
    // the node range [inst_node0[i], inst_node0[i] + inst_N[i]) starting on a
    // multiple of 256, so every 256-node block belongs to one model; the gap
    // nodes up to the next model are isolated (no elements, zero right-hand
    // side) and stay zero.  ninst = 0 for a single model.
    int ninst;
    int *blk_inst;         // [2 * (Np/256 + 1)] per 256-node block: model, end of that model's nodes
    int *inst_blk0;        // [ninst+1] first 256-node block of each model
    int *inst_nodes;       // [ninst] nodes of each model (the l2 norm's count)
    int *node_inst;        // [Np] model of each node (gap nodes: the model before them)
    int *inst_mask;        // [ninst] 1 = the masked node kernels update this model
    std::vector<int> h_inst_mask;   // its host copy
    int *elem_inst, *edge_inst;
    double *dt_inst;
    int *loc_bc_inst;
    std::vector<int> h_inst_node0, h_inst_N;
    std::vector<int> h_inst_e0, h_inst_E, h_inst_b0, h_inst_B;   // per-model triangle / edge ranges
    std::vector<double> h_inst_dt;
    std::vector<double> h_inst_tol;   // per-model nonlinear tolerance (NaN: opts.tol_nonlin)
    // captured BiCGSTAB iteration chunks (k iterations each), replayed with
    // hipGraphLaunch; rebuilt when the kernels or their arguments change
    hipGraphExec_t kgraph[KGRAPH_MAX + 1];
    int kgraph_prec, kgraph_pc;
    const double *kgraph_band;
    int last_lin_it;       // iterations of the previous solve (sizes the first chunk)
    // profile-LU (OP BLK) preconditioner, built on demand for opts.nblock
    int band_nblock;       // block count of the current band structure (0 = none)
    int band_nb;           // blocks
    int band_max_n;        // DOFs of the largest block (LDS of the solve)
    size_t band_size;      // doubles in `band`
    int *blk_n0, *blk_w, *node_blk;
    size_t *blk_off;
    double *band;
    // the same blocks factored by the tiled band LU (blocks beyond the one-
    // workgroup band LU's size, or opts.lu_tiled): band_tiled selects it
    int band_tiled, band_lu_tiled_opt;
    int tb_nband, tb_maxnt, tb_gt, tb_gf;
    std::vector<int> tb_h_nt, tb_h_inst;   // host copies: each band's panels and model
    size_t tb_size;
    TBand *tb_bands;
    double *tb_AB, *tb_LI, *tb_UI, *tb_xb, *tb_yb, *tb_panel;
    // block ILU(0) in multicolour order (preconditioner 3), built on demand
    int ilu_ready, ilu_ncolor, ilu_Rp, ilu_f32;
    size_t ilu_nl, ilu_nu;                 // lower / upper slots (64 per slice slot)
    std::vector<int> h_ilu_cstart;         // [ncolor + 1] first row of each colour (padded to 64)
    std::vector<int> h_ilu_color;          // [N] colour of each node
    std::vector<void *> ilu_allocs;
    int *ilu_rnode, *ilu_lptr, *ilu_uptr, *ilu_lcol, *ilu_ucol, *ilu_lsrc, *ilu_usrc, *ilu_tptr;
    int4 *ilu_upd;
    double *ilu_Lw, *ilu_Uw, *ilu_Dw, *ilu_Dinv, *ilu_w;   // FP64 factor work, colour-ordered vector
    void *ilu_L, *ilu_U, *ilu_D;           // the apply's factors (FP64 or FP32) and U_ii^-1
    // restarted GMRES (opts.krylov = 1), allocated on first use
    double *gm_V;          // basis, GMRES restart + 1 vectors of 3 Np
    double *gm_part;       // block partials of the multi-dots
    double *gm_dev;        // coefficients of the two Gram-Schmidt passes, ||w||^2, the update's y
    double *gm_host;       // pinned copy of gm_dev
};

// 3x3 block inverse by Gauss-Jordan with partial pivoting (block-Jacobi set-up)
__device__ __forceinline__ void inv3(const double (&d)[9], double (&inv)[9]) {
    double a[3][6];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            a[r][c] = d[3 * r + c];
            a[r][3 + c] = (r == c) ? 1.0 : 0.0;
        }
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int r = k + 1; r < 3; r++) if (fabs(a[r][k]) > fabs(a[p][k])) p = r;
        if (p != k)
            for (int c = 0; c < 6; c++) { double t = a[k][c]; a[k][c] = a[p][c]; a[p][c] = t; }
        double piv = a[k][k];
        if (fabs(piv) < SMALL) piv = (piv < 0) ? -SMALL : SMALL;
        double ip = 1.0 / piv;
        for (int c = 0; c < 6; c++) a[k][c] *= ip;
        for (int r = 0; r < 3; r++) {
            if (r == k) continue;
            double f = a[r][k];
            for (int c = 0; c < 6; c++) a[r][c] -= f * a[k][c];
        }
    }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) inv[3 * r + c] = a[r][3 + c];
}



// value index of entry e (0..8) of SELL slot q
__host__ __device__ __forceinline__ size_t sell_vidx(size_t q, int e) {
    size_t lane = q & 63;
    return 9 * (q - lane) + 64 * (size_t)e + lane;
}

// component k (0..2) of node n in a sliced node vector
__host__ __device__ __forceinline__ size_t vidx3(int n, int k) {
    size_t lane = (size_t)(n & 63);
    return 3 * ((size_t)n - lane) + 64 * (size_t)k + lane;
}

// entry e (0..8, row-major 3x3) of node n's diagonal block
__host__ __device__ __forceinline__ size_t didx9(int n, int e) { return sell_vidx((size_t)n, e); }

// XCD-aware workgroup order: the hardware deals consecutive workgroups to the
// 8 XCDs in turn; workgroup b runs logical block xcd_block(), which gives each
// XCD contiguous chunks of XCD_CHUNK blocks (in every run of 8 * XCD_CHUNK
// blocks, one chunk per XCD; a short tail is split contiguously), so data that
// neighbouring blocks share (a node's neighbours, a row's gathered columns a
// mesh column away) stays in that XCD's own L2 while all XCDs still stream
// the same region of memory.  A bijection on [0, gridDim.x): which workgroup
// computes what changes, no result.
constexpr int XCD_CHUNK = 64;
__device__ __forceinline__ int xcd_block() {
    constexpr int C = XCD_CHUNK;
    const int nb = gridDim.x, b = blockIdx.x;
    const int full = (nb / (8 * C)) * (8 * C);
    if (b < full) {
        const int sb = b / (8 * C), w = b % (8 * C);
        return sb * 8 * C + (w & 7) * C + (w >> 3);
    }
    const int rem = nb - full, bb = b - full;
    const int q = rem >> 3, r = rem & 7, x = bb & 7, i = bb >> 3;
    return full + x * q + min(x, r) + i;
}

// launchers (sw2_kernels.hip / sw2_solver.hip)
void launch_old_terms(Ctx *c);
void launch_resid(Ctx *c);
void launch_jacobian(Ctx *c);
void launch_assemble(Ctx *c);
void launch_init(Ctx *c, int *bc_mask_dev);
void launch_shift(Ctx *c);
void launch_inc(Ctx *c, double factor_on_sol);   // state += sol
void launch_scale_sol(Ctx *c, double f);          // sol *= f
void launch_inc_flip(Ctx *c);                     // state += sol, then sol = -sol
void launch_norms(Ctx *c);
void launch_setup(Ctx *c);
void launch_krylov_prologue(Ctx *c);
void launch_krylov_iters(Ctx *c, int k);
void launch_krylov_epilogue(Ctx *c, bool zero_guess);
void launch_geometry(Ctx *c);
void launch_zero(Ctx *c, double *p, size_t n);
void launch_copy(Ctx *c, const double *src, double *dst, size_t n);
void launch_k1_only(Ctx *c);
void launch_jac2d_only(Ctx *c);
void launch_resid2d_only(Ctx *c);
void launch_assemble_only(Ctx *c);
void ensure_lists(Ctx *c);
void launch_list_phase(Ctx *c, int phase, int bs, int grid, int layout);
std::string ilu_structure(Ctx *c);
void ilu_free(Ctx *c);
void launch_ilu_setup(Ctx *c);
void launch_ilu_solve(Ctx *c, double *x, const KScal *sc);
std::string gmres_alloc(Ctx *c);
void gmres_free(Ctx *c);
int launch_gmres(Ctx *c, int max_it, bool *converged);
std::string band_structure(Ctx *c, int nblock);
void band_free(Ctx *c);
void launch_band_setup(Ctx *c);
void launch_band_solve(Ctx *c, double *x, const KScal *sc);
// blocks b: nodes [s0[b], s1[b]) of model inst[b], pattern half bandwidth w[b]
std::string tband_structure(Ctx *c, const std::vector<int> &s0, const std::vector<int> &s1,
                            const std::vector<int> &inst, const std::vector<int> &w);
void tband_free(Ctx *c);
void launch_tband_setup(Ctx *c);
void launch_tband_solve(Ctx *c, double *x, const KScal *sc);
void launch_probe(const double *x, int n, const crlog::tables *t, double *lg, double *sq, double *dv);

int nblocks_for(int n, int bs);

}  // namespace sw2

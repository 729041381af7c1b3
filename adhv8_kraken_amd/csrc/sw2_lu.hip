// Tiled band LU: the profile-LU preconditioner (OP BLK) for node blocks of
// any size, up to the reference's default OP BLK 1 -- one LU of the whole
// matrix (solv_blk_set.c:5-33, solv_blk_profile.c:51-79,
// solv_blk_solve.c:5-62).  Nodes keep the reference's order and the LU has
// no pivoting; a multiplier is its entry over the pivot, or over SMALL when
// |pivot| <= SMALL (solv_blk_profile.c:66-71; formed here as the entry times
// the reciprocal, and the updates as fused multiply-adds), and the backward
// substitution divides by the pivot itself (solv_blk_solve.c:50-52).  Entries
// outside a row's profile are zero and stay zero (no fill outside the envelope
// without pivoting), so factoring the band that covers every row's profile
// gives the skyline factors; only the rounding order of the updates differs.
//
// Storage: each band (one node block) in LAPACK band layout, column-major,
// half bandwidth wp = w + TB (w: the pattern's), entry (i, j) at
// off + (wp + i - j) + j (2 wp + 1).  The TB extra diagonals hold zeros during
// the factorisation, so a panel's L block (rows up to TB + w - 1 below its
// first column) and U block are addressed without bounds tests, and after it
// the fill of the off-diagonal tiles scaled by k_tb_scale.  Seen with leading
// dimension 2 wp the band is a column-major dense matrix, so tiles are
// strided blocks.
//
// Factorisation, per panel of TB columns (two launches over all bands at once:
// a model's node blocks, or every block of every model of an ensemble):
//   k_tb_trsm  a workgroup per 64-row tile of L (rows below the panel:
//              L = A U~^-1 by substitution, U~ = U with the guarded pivots)
//              and per 64-column tile of U (right of the panel: U = L^-1 A);
//   k_tb_gemm  the trailing w x w window -= L U, 64 x 64 tiles, 2 x 4
//              register blocks per thread from LDS; the workgroup of the
//              first tile, the next panel's diagonal block, then factors it
//              (eight waves, row t in lane t, 8 columns per wave, blocked by
//              waves) while the other tiles are updated.
// k_tb_diag factors panel 0's diagonal block the same way.
// Then k_tb_inv forms the inverses of every diagonal block's unit-lower and
// upper factors (all blocks in parallel) and k_tb_scale multiplies every
// off-diagonal tile by its column's inverse, so that each step of the
// substitutions is independent matrix-vector work: the forward sweep is one
// launch per tile (block 0 y_k = L_kk^-1 x_k, the blocks below
// x_i -= (L_ik L_kk^-1) x_k), the backward sweep likewise upwards.
#include <algorithm>
#include <string>
#include <vector>

#include "sw2_ctx.h"

namespace sw2 {

constexpr int TB = 64;          // tile and panel width (DOFs)
constexpr int LW = 8;           // columns (rows) per wave in the blocked factor and solves
constexpr int NWV = TB / LW;    // waves per workgroup of those kernels
constexpr int NT = 64 * NWV;    // their threads

__device__ __forceinline__ size_t tb_idx(const TBand &B, int i, int j) {
    return B.off + (size_t)(B.wp + i - j) + (size_t)j * (2 * B.wp + 1);
}
__device__ __forceinline__ double &tb_at(double *AB, const TBand &B, int i, int j) { return AB[tb_idx(B, i, j)]; }
__device__ __forceinline__ double tb_get(const double *AB, const TBand &B, int i, int j) { return AB[tb_idx(B, i, j)]; }

// lane l's value of v, for a lane index known at compile time or uniform
__device__ __forceinline__ double lane_bcast(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// Band record b in SGPRs at once: the empty asm statement uses every field, so
// the compiler cannot sink the fields' loads past the kernel's first branches
// (one scalar round trip instead of three before the first vector load);
// `flag` (a done flag read beside it) is held the same way.
__device__ __forceinline__ TBand band_rec(const TBand *bands, int b, int flag = 0) {
    const TBand B = bands[b];
    asm volatile("" ::"s"(B.m), "s"(B.w), "s"(B.wp), "s"(B.nt), "s"(B.inst), "s"(B.off), "s"(B.toff), "s"(B.xoff),
                 "s"(flag));
    return B;
}

__device__ __forceinline__ double guarded(double piv) { return fabs(piv) > SMALL ? piv : SMALL; }

// Scaled matrix (diagonal blocks and SELL off-diagonal blocks) into the bands;
// the bands are zeroed before.  Ensembles: nodes of no block (the gaps between
// models) and models outside the mask are skipped.
__global__ void k_tb_load(int N, const int *node_blk, const TBand *bands, double *AB, const double *diag,
                          const double *offd, const int *row_ptr, const int *sell_ptr, const int *sell_col,
                          const int *mask) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int bk = node_blk[a];
    if (bk < 0) return;
    const TBand B = bands[bk];
    if (mask && !mask[B.inst]) return;
    const int n1 = B.n0 + B.m / 3;
    const int ia = 3 * (a - B.n0);
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) tb_at(AB, B, ia + r, ia + cc) = diag[didx9(a, 3 * r + cc)];
    const size_t q0 = 64 * (size_t)sell_ptr[a >> 6] + (a & 63);
    const int len = row_ptr[a + 1] - row_ptr[a];
    for (int s = 0; s < len; s++) {
        const size_t q = q0 + 64 * (size_t)s;
        const int cn = sell_col[q];
        if (cn < B.n0 || cn >= n1) continue;
        const int jc = 3 * (cn - B.n0);
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) tb_at(AB, B, ia + r, jc + cc) = offd[sell_vidx(q, 3 * r + cc)];
    }
}

// LU of a 64 x 64 diagonal block, blocked by LW = 8 columns: lane t of each of
// the eight waves holds row t (r[LW]), wave w columns 8 w .. 8 w + 7.  Wave
// P factors its 8 columns (right-looking: the pivot read from lane p with
// readlane, the multipliers as entry times the reciprocal guarded pivot),
// publishes them in LDS, and after one barrier the waves to its right update
// their columns with them (the same right-looking steps, multipliers read
// from LDS); then wave P + 1 factors its (updated) columns.  Eight barriers
// per block instead of one per pivot; the serial part of each phase, one
// wave's pivots and the next wave's update, is half what 16-column groups
// give, and every entry takes the same updates in the same order as with any
// grouping.  Then the factors go back into the band and
// [strict U by rows | strict L by columns | 1 / guarded pivots] into the
// band's panel copy P for k_tb_trsm (zeros on and across the diagonal let it
// update every column at every step).  pan: [2][TB][LW + 1] and rd: [TB] in LDS.
constexpr int TB_PANEL = 2 * TB * TB + TB;   // doubles of one band's panel copy
__device__ __forceinline__ void diag_lu_store(double (&r)[LW], double (*pan)[TB][LW + 1], double *rd, double *AB,
                                              const TBand &B, double *P, int c0, int kb, int lane, int wv) {
#pragma unroll
    for (int Q = 0; Q < NWV; Q++) {
        if (wv == Q) {
#pragma unroll
            for (int pj = 0; pj < LW; pj++) {
                const int p = LW * Q + pj;
                const double rden = 1.0 / guarded(lane_bcast(r[pj], p));
                if (lane == p) rd[p] = rden;
                const double l = lane > p ? r[pj] * rden : 0.0;
                if (lane > p) r[pj] = l;
#pragma unroll
                for (int qj = pj + 1; qj < LW; qj++) r[qj] = fma(-l, lane_bcast(r[qj], p), r[qj]);
            }
#pragma unroll
            for (int j = 0; j < LW; j++) pan[Q & 1][lane][j] = r[j];
        }
        __syncthreads();
        if (wv > Q) {
#pragma unroll
            for (int sj = 0; sj < LW; sj++) {
                const int sp = LW * Q + sj;
                const double l = lane > sp ? pan[Q & 1][lane][sj] : 0.0;
#pragma unroll
                for (int j = 0; j < LW; j++) r[j] = fma(-l, lane_bcast(r[j], sp), r[j]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LW; j++) {
        const int col = LW * wv + j;
        if (lane < kb && col < kb) tb_at(AB, B, c0 + lane, c0 + col) = r[j];
        P[TB * lane + col] = col > lane ? r[j] : 0.0;              // U strictly above the diagonal, row `lane`
        P[TB * TB + TB * col + lane] = col < lane ? r[j] : 0.0;    // L, column `col` (unit diagonal not stored)
    }
    if (wv == 0) P[2 * TB * TB + lane] = lane < kb ? rd[lane] : 0.0;
}

// Panel 0's diagonal block, one workgroup per band (every later panel's is
// factored by k_tb_gemm of the panel before, which updates it last).
__global__ void __launch_bounds__(NT) k_tb_diag(const TBand *bands, double *AB, double *panel, int k,
                                                const int *mask) {
    __shared__ double pan[2][TB][LW + 1];   // a factored column group (double-buffered by group parity)
    __shared__ double rd[TB];           // 1 / guarded pivots
    const TBand B = band_rec(bands, blockIdx.x);
    if (k >= B.nt || (mask && !mask[B.inst])) return;
    const int c0 = k * TB, kb = min(TB, B.m - c0);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tc = min(lane, kb - 1);
    double r[LW];
    // loads at clamped (in-band) positions, then zeroed outside the block: the
    // rows and columns past a short last block are zero, their pivots guarded
    // and their multipliers zero
#pragma unroll
    for (int j = 0; j < LW; j++) {
        const int col = LW * wv + j;
        const double v = tb_get(AB, B, c0 + tc, c0 + min(col, kb - 1));
        r[j] = (lane < kb && col < kb) ? v : 0.0;
    }
    diag_lu_store(r, pan, rd, AB, B, panel + (size_t)TB_PANEL * blockIdx.x, c0, kb, lane, wv);
}

// Panel k's L tiles (blockIdx.x < gl: 64 rows each below the panel) and U
// tiles (blockIdx.x >= gl: 64 columns each right of it), one workgroup of
// eight waves per tile, blocked by 8 as k_tb_diag.  L side: lane t holds row
// t of the tile, wave w its columns 8 w .. 8 w + 7; wave P forms the
// multipliers of its 8 columns, publishes them in LDS, and after one barrier
// the waves to its right take the rank-8 update.  U side: lane t holds
// column t, wave w its rows 8 w .. 8 w + 7; wave P finishes its rows and
// the waves below take their update.  The diagonal block's entries come from
// an LDS copy of the half of the panel block the side uses (broadcast reads;
// strict triangles, so every update runs over whole column groups): the
// L side the strict U rows and the reciprocal guarded pivots, the U side the
// strict L columns.  LDS: that copy, and one region that is the U side's
// transpose buffer before and after the blocked steps and the published
// column (row) groups during them -- 66.5 KB, two workgroups per CU.
__global__ void __launch_bounds__(NT) k_tb_trsm(const TBand *bands, double *AB, const double *panel, int k,
                                                int gl, const int *mask) {
    __shared__ double sm[2 * TB * (TB + 1)];
    double (*T)[TB + 1] = reinterpret_cast<double (*)[TB + 1]>(sm);          // U side: the tile, transposed
    double (*pan)[TB][LW + 1] = reinterpret_cast<double (*)[TB][LW + 1]>(sm);   // a finished group, by parity
    double *Pl = sm + TB * (TB + 1);                                         // the panel half
    const TBand B = band_rec(bands, blockIdx.y);
    if (k >= B.nt || (mask && !mask[B.inst])) return;
    const int c0 = k * TB, kb = min(TB, B.m - c0);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r0 = c0 + kb, r1 = min(B.m, r0 + B.w);
    const bool lside = (int)blockIdx.x < gl;
    const int g = lside ? blockIdx.x : blockIdx.x - gl;
    const int base = r0 + TB * g;
    if (base >= r1) return;
    // the panel half: [strict U rows | 1 / guarded pivots] (L side) or strict
    // L columns (U side); a thread's loads all issued before its LDS stores
    const double *Pg = panel + (size_t)TB_PANEL * blockIdx.y + (lside ? 0 : TB * TB);
    {
        constexpr int NP = TB * TB / NT;
        double v[NP];
#pragma unroll
        for (int u = 0; u < NP; u++) v[u] = Pg[threadIdx.x + NT * u];
        const double rv = lside && threadIdx.x < TB ? Pg[2 * TB * TB + threadIdx.x] : 0.0;
#pragma unroll
        for (int u = 0; u < NP; u++) Pl[threadIdx.x + NT * u] = v[u];
        if (lside && threadIdx.x < TB) Pl[TB * TB + threadIdx.x] = rv;
    }
    double r[LW];
    if (lside) {
        const double *U = Pl, *R = Pl + TB * TB;
        const int i = base + lane;
        const bool act = i < r1;
        const int ic = min(i, r1 - 1);
#pragma unroll
        for (int j = 0; j < LW; j++) {
            const int col = LW * wv + j;
            const double v = tb_get(AB, B, ic, c0 + min(col, kb - 1));   // clamped: in flight together
            r[j] = (act && col < kb) ? v : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int P = 0; P < NWV; P++) {
            if (wv == P) {
#pragma unroll
                for (int pj = 0; pj < LW; pj++) {
                    const int p = LW * P + pj;
                    const double l = r[pj] * R[p];
                    r[pj] = l;
#pragma unroll
                    for (int qj = pj + 1; qj < LW; qj++) r[qj] = fma(-l, U[TB * p + LW * P + qj], r[qj]);
                }
#pragma unroll
                for (int j = 0; j < LW; j++) pan[P & 1][lane][j] = r[j];
            }
            __syncthreads();
            if (wv > P) {
#pragma unroll
                for (int sj = 0; sj < LW; sj++) {
                    const double l = pan[P & 1][lane][sj];
#pragma unroll
                    for (int j = 0; j < LW; j++) r[j] = fma(-l, U[TB * (LW * P + sj) + LW * wv + j], r[j]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < LW; j++) {
            const int col = LW * wv + j;
            if (act && col < kb) tb_at(AB, B, i, c0 + col) = r[j];
        }
    } else {
        const double *L = Pl;
        const int ncol = min(TB, r1 - base);
        constexpr int NU = TB * TB / NT;
        double tv[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int e = threadIdx.x + NT * u, row = e & (TB - 1), col = e >> 6;   // down a column
            tv[u] = tb_get(AB, B, c0 + min(row, kb - 1), base + min(col, ncol - 1));
        }
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int e = threadIdx.x + NT * u, row = e & (TB - 1), col = e >> 6;
            T[row][col] = (row < kb && col < ncol) ? tv[u] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < LW; j++) r[j] = T[LW * wv + j][lane];
        __syncthreads();   // the groups below overwrite the transpose buffer
#pragma unroll
        for (int P = 0; P < NWV; P++) {
            if (wv == P) {
#pragma unroll
                for (int pj = 0; pj < LW; pj++) {
                    const int p = LW * P + pj;
#pragma unroll
                    for (int qj = pj + 1; qj < LW; qj++) r[qj] = fma(-L[TB * p + LW * P + qj], r[pj], r[qj]);
                }
#pragma unroll
                for (int j = 0; j < LW; j++) pan[P & 1][lane][j] = r[j];
            }
            __syncthreads();
            if (wv > P) {
#pragma unroll
                for (int sj = 0; sj < LW; sj++) {
                    const double u = pan[P & 1][lane][sj];
#pragma unroll
                    for (int j = 0; j < LW; j++) r[j] = fma(-L[TB * (LW * P + sj) + LW * wv + j], u, r[j]);
                }
            }
        }
        // the last group reads ended before the last barrier above
#pragma unroll
        for (int j = 0; j < LW; j++) T[LW * wv + j][lane] = r[j];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int e = threadIdx.x + NT * u, row = e & (TB - 1), col = e >> 6;
            if (row < kb && col < ncol) tb_at(AB, B, c0 + row, base + col) = T[row][col];
        }
    }
}

// Trailing window of panel k: A[r0:r1, r0:r1] -= L[r0:r1, panel] U[panel, r0:r1],
// one 64 x 64 tile per workgroup (512 threads, a 2 x 4 block each).  The
// workgroup of the window's first tile -- panel k + 1's diagonal block, which
// no later update of panel k touches -- goes on to factor that block
// (diag_lu_store), while the other tiles are updated: panel k + 1's diagonal
// LU leaves the critical path of the rank-64 update.
__global__ void __launch_bounds__(NT) k_tb_gemm(const TBand *bands, double *AB, double *panel, int k, int gt,
                                                const int *mask) {
    __shared__ double sm[TB * TB + TB * (TB + 1)];
    double (*Ls)[TB] = reinterpret_cast<double (*)[TB]>(sm);                 // Ls[p][row]
    double (*Us)[TB + 1] = reinterpret_cast<double (*)[TB + 1]>(sm + TB * TB);   // Us[p][col]
    const TBand B = band_rec(bands, blockIdx.y);
    if (k >= B.nt || (mask && !mask[B.inst])) return;
    const int c0 = k * TB, kb = min(TB, B.m - c0);
    const int r0 = c0 + kb, r1 = min(B.m, r0 + B.w);
    const int I = blockIdx.x % gt, J = blockIdx.x / gt;
    const int i0 = r0 + TB * I, j0 = r0 + TB * J;
    const bool next = I == 0 && J == 0 && k + 1 < B.nt;   // the tile is panel k + 1's diagonal block
    if (!next && (i0 >= r1 || j0 >= r1)) return;
    // the next diagonal block is read whole (rows and columns past the window
    // keep their values: their products are zero), other tiles in the window
    const int cl = next ? min(B.m, r0 + TB) : r1;
    const int t = threadIdx.x;
    const int tr = t & 31, tc = t >> 5;   // rows 2 tr, 2 tr + 1, columns 4 tc .. 4 tc + 3
    // this thread's 8 entries of the tile, fetched first (clamped positions)
    // so that their loads overlap the product
    double cv[2][4];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
        for (int u = 0; u < 2; u++)
            cv[u][v] = tb_get(AB, B, min(i0 + 2 * tr + u, cl - 1), min(j0 + 4 * tc + v, cl - 1));
    // both tiles read down their columns (consecutive addresses across threads)
    constexpr int NU = TB * TB / NT;
    double lv[NU], uv[NU];   // all loads in flight before the LDS stores
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int e = t + NT * u, lo = e & (TB - 1), hi = e >> 6;
        lv[u] = tb_get(AB, B, min(i0 + lo, r1 - 1), c0 + min(hi, kb - 1));
        uv[u] = tb_get(AB, B, c0 + min(lo, kb - 1), min(j0 + hi, r1 - 1));
    }
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int e = t + NT * u, lo = e & (TB - 1), hi = e >> 6;
        Ls[hi][lo] = (hi < kb && i0 + lo < r1) ? lv[u] : 0.0;   // row lo, panel column hi
        Us[lo][hi] = (lo < kb && j0 + hi < r1) ? uv[u] : 0.0;   // panel row lo, column hi
    }
    __syncthreads();
    double acc[2][4] = {};
#pragma unroll 8
    for (int p = 0; p < kb; p++) {
        double a[2], b[4];
#pragma unroll
        for (int u = 0; u < 2; u++) a[u] = Ls[p][2 * tr + u];
#pragma unroll
        for (int v = 0; v < 4; v++) b[v] = Us[p][4 * tc + v];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
    if (next) {
        // the updated block through LDS into the diagonal LU's layout
        double (*D)[TB + 1] = reinterpret_cast<double (*)[TB + 1]>(sm);
        __syncthreads();   // the product's LDS reads are done
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) D[2 * tr + u][4 * tc + v] = cv[u][v] - acc[u][v];
        __syncthreads();
        const int kb1 = min(TB, B.m - r0);
        const int lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
        double r[LW];
#pragma unroll
        for (int j = 0; j < LW; j++) {
            const int col = LW * wv + j;
            r[j] = (lane < kb1 && col < kb1) ? D[lane][col] : 0.0;
        }
        double (*pan)[TB][LW + 1] = reinterpret_cast<double (*)[TB][LW + 1]>(sm + TB * (TB + 1));
        double *rd = sm + TB * (TB + 1) + 2 * TB * (LW + 1);
        diag_lu_store(r, pan, rd, AB, B, panel + (size_t)TB_PANEL * blockIdx.y, r0, kb1, lane, wv);
        return;
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int j = j0 + 4 * tc + v;
        if (j >= r1) continue;
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int i = i0 + 2 * tr + u;
            if (i < r1) tb_at(AB, B, i, j) = cv[u][v] - acc[u][v];
        }
    }
}

// Inverses of every diagonal block's factors for the substitutions: LI the
// unit-lower one, UI the upper one with the pivots themselves; 64 x 64
// column-major per tile (padding rows and columns of a short last tile are
// identity).  One wave per tile: lane j forms column j.
__global__ void __launch_bounds__(64) k_tb_inv(const TBand *bands, const double *AB, double *LI, double *UI,
                                                const int *mask) {
    __shared__ double D[TB][TB + 1];
    __shared__ double T[TB][TB + 1];
    const TBand B = band_rec(bands, blockIdx.y);
    const int k = blockIdx.x;
    if (k >= B.nt || (mask && !mask[B.inst])) return;
    const int c0 = k * TB, kb = min(TB, B.m - c0), t = threadIdx.x;
    for (int j = 0; j < TB; j++)
        D[t][j] = (t < kb && j < kb) ? tb_get(AB, B, c0 + t, c0 + j) : (t == j ? 1.0 : 0.0);
    __syncthreads();
    double x[TB];
    // unit lower: x_i = delta_ij - sum_{q<i} L_iq x_q
#pragma unroll
    for (int i = 0; i < TB; i++) {
        double s = i == t ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < i; q++) s = fma(-D[i][q], x[q], s);
        x[i] = s;
    }
    const size_t o = 4096 * (B.toff + k);
#pragma unroll
    for (int i = 0; i < TB; i++) T[i][t] = x[i];
    __syncthreads();
    for (int j = 0; j < TB; j++) LI[o + 64 * j + t] = T[t][j];
    __syncthreads();
    // upper: x_i = (delta_ij - sum_{q>i} U_iq x_q) / U_ii
#pragma unroll
    for (int i = TB - 1; i >= 0; i--) {
        double s = i == t ? 1.0 : 0.0;
#pragma unroll
        for (int q = i + 1; q < TB; q++) s = fma(-D[i][q], x[q], s);
        x[i] = s / D[i][i];
    }
#pragma unroll
    for (int i = 0; i < TB; i++) T[i][t] = x[i];
    __syncthreads();
    for (int j = 0; j < TB; j++) UI[o + 64 * j + t] = T[t][j];
}

// After k_tb_inv: every off-diagonal tile of L (below a diagonal block) times
// that block's LI, and of U (above it) times its UI, in place, so that a sweep
// step's off-diagonal update reads one tile and no inverse: x_i -= (L_ik LI_k)
// x_k = L_ik y_k, y_i -= (U_ik UI_k) y_k = U_ik z_k.  The products fill at most
// the band's TB extra diagonals (entries within w + 63 of the diagonal), which
// is what they are stored for.  One workgroup of 256 threads per tile, a 4 x 4
// register block each from LDS.
__global__ void __launch_bounds__(256) k_tb_scale(const TBand *bands, double *AB, const double *LI, const double *UI,
                                                  int gf, const int *mask) {
    __shared__ double As[TB][TB];       // As[q][row]: the tile, column q
    __shared__ double Ms[TB][TB + 1];   // Ms[q][col]: the inverse
    const TBand B = band_rec(bands, blockIdx.z);
    const int k = blockIdx.x;
    if (k >= B.nt || (mask && !mask[B.inst])) return;
    const bool low = (int)blockIdx.y < gf;
    const int g = (low ? (int)blockIdx.y : (int)blockIdx.y - gf) + 1;
    const int c0 = k * TB, kb = min(TB, B.m - c0);
    // the tile's first row, and its rows inside the reach of the panel's columns
    const int tr0 = low ? c0 + TB * g : c0 - TB * g;
    const int i0 = low ? tr0 : max(tr0, max(0, c0 - B.w));
    const int i1 = low ? min(tr0 + TB, min(B.m, c0 + kb + B.w)) : tr0 + TB;
    if (i0 >= i1) return;
    const int t = threadIdx.x;
    const double *M = (low ? LI : UI) + 4096 * (B.toff + k);
    double av[16], mv[16];   // all loads in flight before the LDS stores
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int e = t + 256 * u, lo = e & (TB - 1), hi = e >> 6;
        av[u] = tb_get(AB, B, min(max(tr0 + lo, i0), i1 - 1), c0 + min(hi, kb - 1));
        mv[u] = M[e];   // column-major: row lo, column hi
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int e = t + 256 * u, lo = e & (TB - 1), hi = e >> 6;
        const int r = tr0 + lo;
        As[hi][lo] = (r >= i0 && r < i1 && hi < kb) ? av[u] : 0.0;
        Ms[lo][hi] = mv[u];
    }
    __syncthreads();
    const int tr = t & 15, tc = t >> 4;   // rows 4 tr .. 4 tr + 3, columns 4 tc .. 4 tc + 3
    double acc[4][4] = {};
#pragma unroll 8
    for (int q = 0; q < kb; q++) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; u++) { a[u] = As[q][4 * tr + u]; b[u] = Ms[q][4 * tc + u]; }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int c = 4 * tc + v;
        if (c >= kb) continue;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = tr0 + 4 * tr + u;
            if (r >= i0 && r < i1) tb_at(AB, B, r, c0 + c) = acc[u][v];
        }
    }
}

// sliced node vector <-> band-ordered vector (nodes of no block skipped; the
// scatter leaves a model that has converged, or is outside the mask, as it is)
__global__ void k_tb_gather(int N, const int *node_blk, const TBand *bands, const double *x, double *xb) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int bk = node_blk[a];
    if (bk < 0) return;
    const TBand B = bands[bk];
    for (int v = 0; v < 3; v++) xb[B.xoff + 3 * (a - B.n0) + v] = x[vidx3(a, v)];
}
__global__ void k_tb_scatter(int N, const int *node_blk, const TBand *bands, const double *xb, double *x,
                             const KScal *sc, const int *mask) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int bk = node_blk[a];
    if (bk < 0) return;
    const TBand B = bands[bk];
    if ((sc && sc[B.inst].done) || (mask && !mask[B.inst])) return;
    for (int v = 0; v < 3; v++) x[vidx3(a, v)] = xb[B.xoff + 3 * (a - B.n0) + v];
}

// Sweep kernels: four waves per workgroup split a tile's 64 columns (16
// each); lane r accumulates row r over its wave's columns (the column value
// broadcast from lane c) and the four partial sums are added in LDS in wave
// order.  Block 0 applies the diagonal block's inverse; block g >= 1 the
// off-diagonal tile scaled by it (k_tb_scale), to the same input slice, so
// the blocks of a step are independent and each reads one tile: all of a
// block's loads are issued at once after a single band-record fetch.
__device__ __forceinline__ double tile_mv_part(const double (&m)[16], double xr, int wv) {
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 16; u++) s = fma(m[u], lane_bcast(xr, 16 * wv + u), s);
    return s;
}
__device__ __forceinline__ double sum4(double (*part)[TB], int lane) {
    return ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

// Forward step k: block 0 stores y_k = LI_k x_k in yb; block g >= 1 takes
// x[tile k + g] -= (L LI_k)[tile k + g, tile k] x_k.
__global__ void __launch_bounds__(256) k_tb_fwd(const TBand *bands, const double *AB, const double *LI, double *xb,
                                                double *yb, int k, const int *done, const int *mask,
                                                const KScal *sc) {
    __shared__ double part[4][TB];
    const int stop = *done;   // the band record and the flag fetched together
    const TBand B = band_rec(bands, blockIdx.y, stop);
    // ensembles: the band's model, converged or outside the mask, is skipped
    if (stop || k >= B.nt || (mask && (!mask[B.inst] || (sc && sc[B.inst].done)))) return;
    const int c0 = k * TB, kb = min(TB, B.m - c0), g = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rmax = min(B.m, c0 + kb + B.w);
    if (g > 0 && c0 + TB * g >= rmax) return;
    const int i = c0 + TB * g + lane, ic = min(i, rmax - 1);
    const double xv = xb[B.xoff + c0 + min(lane, kb - 1)];
    double m[16], xi = 0.0;
    if (g == 0) {
        const double *M = LI + 4096 * (B.toff + k);
#pragma unroll
        for (int u = 0; u < 16; u++) m[u] = M[64 * (16 * wv + u) + lane];
    } else {
#pragma unroll
        for (int u = 0; u < 16; u++) m[u] = tb_get(AB, B, ic, c0 + min(16 * wv + u, kb - 1));
        xi = xb[B.xoff + ic];
    }
    const double xr = lane < kb ? xv : 0.0;
    part[wv][lane] = tile_mv_part(m, xr, wv);
    __syncthreads();
    if (wv != 0) return;
    if (g == 0) {
        if (lane < kb) yb[B.xoff + c0 + lane] = sum4(part, lane);
    } else if (i < rmax) {
        xb[B.xoff + i] = xi - sum4(part, lane);
    }
}

// Backward step k: block 0 stores z_k = UI_k y_k in xb; block g >= 1 takes
// y[tile k - g] -= (U UI_k)[tile k - g, tile k] y_k.
__global__ void __launch_bounds__(256) k_tb_bwd(const TBand *bands, const double *AB, const double *UI, double *xb,
                                                double *yb, int k, const int *done, const int *mask,
                                                const KScal *sc) {
    __shared__ double part[4][TB];
    const int stop = *done;   // the band record and the flag fetched together
    const TBand B = band_rec(bands, blockIdx.y, stop);
    // ensembles: the band's model, converged or outside the mask, is skipped
    if (stop || k >= B.nt || (mask && (!mask[B.inst] || (sc && sc[B.inst].done)))) return;
    const int c0 = k * TB, kb = min(TB, B.m - c0), g = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rmin = max(0, c0 - B.w);
    if (g > 0 && c0 - TB * g + TB <= rmin) return;
    const int i = c0 - TB * g + lane, ic = max(i, rmin);
    const double yv = yb[B.xoff + c0 + min(lane, kb - 1)];
    double m[16], yi = 0.0;
    if (g == 0) {
        const double *M = UI + 4096 * (B.toff + k);
#pragma unroll
        for (int u = 0; u < 16; u++) m[u] = M[64 * (16 * wv + u) + lane];
    } else {
#pragma unroll
        for (int u = 0; u < 16; u++) m[u] = tb_get(AB, B, ic, c0 + min(16 * wv + u, kb - 1));
        yi = yb[B.xoff + ic];
    }
    const double yr = lane < kb ? yv : 0.0;
    part[wv][lane] = tile_mv_part(m, yr, wv);
    __syncthreads();
    if (wv != 0) return;
    if (g == 0) {
        if (lane < kb) xb[B.xoff + c0 + lane] = sum4(part, lane);
    } else if (i >= rmin) {
        yb[B.xoff + i] = yi - sum4(part, lane);
    }
}

void tband_free(Ctx *c) {
    for (void *p : {(void *)c->tb_bands, (void *)c->tb_AB, (void *)c->tb_LI, (void *)c->tb_UI, (void *)c->tb_xb,
                    (void *)c->tb_yb, (void *)c->tb_panel})
        if (p) (void)hipFree(p);
    c->tb_bands = nullptr; c->tb_AB = nullptr; c->tb_LI = nullptr; c->tb_UI = nullptr; c->tb_panel = nullptr;
    c->tb_xb = nullptr; c->tb_yb = nullptr;
    c->tb_nband = 0; c->tb_maxnt = 0; c->tb_gt = 0; c->tb_gf = 0; c->tb_size = 0;
    c->tb_h_nt.clear(); c->tb_h_inst.clear();
}

// Bands of the node blocks (nodes [s0[b], s1[b]) of model inst[b]) with
// pattern half bandwidths w[] (DOFs).  Returns an error text or "".
std::string tband_structure(Ctx *c, const std::vector<int> &s0, const std::vector<int> &s1,
                            const std::vector<int> &inst, const std::vector<int> &w) {
    tband_free(c);
    const int nblk = (int)w.size();
    std::vector<TBand> h(nblk);
    size_t off = 0, toff = 0, xoff = 0;
    int maxnt = 0, maxw = 0;
    for (int b = 0; b < nblk; b++) {
        TBand &B = h[b];
        B.n0 = s0[b];
        B.m = 3 * (s1[b] - s0[b]);
        B.w = w[b];
        B.wp = w[b] + TB;
        B.nt = (B.m + TB - 1) / TB;
        B.pad = 0;
        B.inst = inst[b];
        B.pad2 = 0;
        B.off = off;
        B.toff = toff;
        B.xoff = xoff;
        off += (size_t)B.m * (2 * B.wp + 1);
        toff += B.nt;
        xoff += B.m;
        maxnt = std::max(maxnt, B.nt);
        maxw = std::max(maxw, B.w);
    }
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    const size_t need = sizeof(double) * (off + 2 * 4096 * toff + 2 * xoff);
    if (need + ((size_t)2 << 30) > freeb)
        return "tiled band LU: " + std::to_string(need >> 20) + " MiB of band storage do not fit in the free " +
               std::to_string(freeb >> 20) + " MiB; raise nblock" +
               (c->ninst ? ", or step the models in groups that fit" : "");
    if (hipMalloc((void **)&c->tb_bands, sizeof(TBand) * nblk) != hipSuccess ||
        hipMemcpy(c->tb_bands, h.data(), sizeof(TBand) * nblk, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void **)&c->tb_AB, sizeof(double) * off) != hipSuccess ||
        hipMalloc((void **)&c->tb_LI, sizeof(double) * 4096 * toff) != hipSuccess ||
        hipMalloc((void **)&c->tb_UI, sizeof(double) * 4096 * toff) != hipSuccess ||
        hipMalloc((void **)&c->tb_xb, sizeof(double) * xoff) != hipSuccess ||
        hipMalloc((void **)&c->tb_yb, sizeof(double) * xoff) != hipSuccess ||
        hipMalloc((void **)&c->tb_panel, sizeof(double) * TB_PANEL * nblk) != hipSuccess) {
        tband_free(c);
        return "hipMalloc failed for the tiled band LU";
    }
    c->tb_nband = nblk;
    c->tb_maxnt = maxnt;
    c->tb_h_nt.resize(nblk);
    c->tb_h_inst.resize(nblk);
    for (int b = 0; b < nblk; b++) { c->tb_h_nt[b] = h[b].nt; c->tb_h_inst[b] = h[b].inst; }
    c->tb_gt = (maxw + TB - 1) / TB;
    c->tb_gf = (maxw + 2 * TB - 2) / TB;   // tiles below (above) a panel that its L (U) reaches
    c->tb_size = off;
    return "";
}

// the models the masked launches of an ensemble work on (null: every band)
static const int *tb_mask(const Ctx *c) { return c->ninst ? c->inst_mask : nullptr; }

// panels of the largest band whose model is in the mask: the launches past
// it would find every band's model masked (an ensemble whose large models
// have converged stops paying for their panels)
static int tb_active_nt(const Ctx *c) {
    if (!c->ninst || c->h_inst_mask.empty()) return c->tb_maxnt;
    int nt = 0;
    for (int b = 0; b < c->tb_nband; b++)
        if (c->h_inst_mask[c->tb_h_inst[b]]) nt = std::max(nt, c->tb_h_nt[b]);
    return nt;
}

void launch_tband_setup(Ctx *c) {
    const int *mask = tb_mask(c);
    const int nt = tb_active_nt(c);
    if (nt == 0) return;
    launch_zero(c, c->tb_AB, c->tb_size);
    hipLaunchKernelGGL(k_tb_load, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N, c->node_blk,
                       c->tb_bands, c->tb_AB, c->diag, c->offd, c->row_ptr, c->sell_ptr, c->sell_col, mask);
    const int gt = std::max(1, c->tb_gt);
    hipLaunchKernelGGL(k_tb_diag, dim3(c->tb_nband), dim3(NT), 0, c->stream, c->tb_bands, c->tb_AB, c->tb_panel,
                       0, mask);
    for (int k = 0; k < nt; k++) {
        hipLaunchKernelGGL(k_tb_trsm, dim3(2 * gt, c->tb_nband), dim3(NT), 0, c->stream, c->tb_bands, c->tb_AB,
                           c->tb_panel, k, gt, mask);
        hipLaunchKernelGGL(k_tb_gemm, dim3(gt * gt, c->tb_nband), dim3(NT), 0, c->stream, c->tb_bands, c->tb_AB,
                           c->tb_panel, k, gt, mask);
    }
    hipLaunchKernelGGL(k_tb_inv, dim3(nt, c->tb_nband), dim3(64), 0, c->stream, c->tb_bands, c->tb_AB,
                       c->tb_LI, c->tb_UI, mask);
    const int gf = std::max(1, c->tb_gf);
    hipLaunchKernelGGL(k_tb_scale, dim3(nt, 2 * gf, c->tb_nband), dim3(256), 0, c->stream, c->tb_bands,
                       c->tb_AB, c->tb_LI, c->tb_UI, gf, mask);
}

// sc: the Krylov scalars (one per model) whose done flags end the sweeps, or
// null outside the iteration loop
void launch_tband_solve(Ctx *c, double *x, const KScal *sc) {
    const int *mask = tb_mask(c);
    const dim3 gn(nblocks_for(c->N, 256));
    hipLaunchKernelGGL(k_tb_gather, gn, dim3(256), 0, c->stream, c->N, c->node_blk, c->tb_bands, x, c->tb_xb);
    const dim3 g(1 + c->tb_gf, c->tb_nband);
    // one model: its done flag, or the first band record's pad (always 0),
    // read beside the band record; ensembles: each band's model's flag
    const int *done = (sc && !mask) ? &sc->done : &c->tb_bands[0].pad;
    const KScal *msc = mask ? sc : nullptr;
    const int nt = tb_active_nt(c);
    for (int k = 0; k < nt; k++)
        hipLaunchKernelGGL(k_tb_fwd, g, dim3(256), 0, c->stream, c->tb_bands, c->tb_AB, c->tb_LI, c->tb_xb, c->tb_yb,
                           k, done, mask, msc);
    for (int k = nt - 1; k >= 0; k--)
        hipLaunchKernelGGL(k_tb_bwd, g, dim3(256), 0, c->stream, c->tb_bands, c->tb_AB, c->tb_UI, c->tb_xb, c->tb_yb,
                           k, done, mask, msc);
    hipLaunchKernelGGL(k_tb_scatter, gn, dim3(256), 0, c->stream, c->N, c->node_blk, c->tb_bands, c->tb_xb, x, sc,
                       mask);
}

}  // namespace sw2

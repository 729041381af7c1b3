// Linear-solve kernels: symmetric Jacobi scaling (spv_jacobi), 3x3
// block-Jacobi preconditioner, and the reference's BiCGSTAB
// (src/solver/solv_bcgstab.c:39-301) as three fused kernels per iteration:
//
//   K1  Ap = M^-1 A p,                          gamma = (q, Ap), alpha = rho/gamma
//   K2  s = r - alpha Ap (formed on the fly),   As = M^-1 A s,
//       omega = (As,s)/(As,As), rho' = -omega (As,q)
//   K3  u += alpha p; u += omega s; r = s - omega As;
//       p = beta (p - omega Ap) + r with beta = (rho' alpha)/(rho omega); ||r||inf
//
// Each kernel writes block partials of its dot products, and a one-block
// finisher kernel reduces them in block order and updates the device
// scalars, so a BiCGSTAB iteration is 6 launches with no host round trip; the
// host reads the convergence flag every opts.check_every iterations and
// kernels launched past convergence return at entry.  The same kernels run in FP64 or FP32
// (precision 0), with FP32 matrix and preconditioner and FP64 vectors
// (precision 1), or all in FP32 (precision 2); dot products accumulate in FP64.
//
// HBM layout (wave-sliced, 64 nodes per slice): diagonal blocks at
// didx9(n, e), off-diagonal blocks in SELL-64 (sell_vidx), node vectors at
// vidx3(n, k); every per-node load of a wavefront is one 64-wide contiguous
// access.  SpMV reads 72 B (FP64) per block.
#include <cmath>
#include <string>
#include <type_traits>

#include "sw2_ctx.h"

namespace sw2 {

// FP64 matrix and preconditioner entries are read once per product:
// non-temporal loads keep them from displacing the gathered vector in L2
// (57 against 63 us for the FP64 K1 at 1M triangles; the FP32 streams are
// faster with ordinary loads)
template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    if constexpr (sizeof(T) == 8) return __builtin_nontemporal_load(p);
    else return *p;
}


// model tables of an ensemble (null for one model)
#define BI(c) ((c)->ninst ? (const int *)(c)->blk_inst : (const int *)nullptr)
#define IB0(c) ((c)->ninst ? (const int *)(c)->inst_blk0 : (const int *)nullptr)
#define NFIN(c) ((c)->ninst ? (c)->ninst : 1)



// ------------------------------------------------------------- Jacobi scaling
// spv_jacobi 3-DOF branch (src/solver/spv_jacobi.c:222-316)
__global__ void k_scale_rows(int N, const double *diag, const double *offd, const int *row_ptr, const int *sell_ptr,
                             double *sv) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double s[3];
    for (int j = 0; j < 3; j++) {
        double m = 0.0;
        for (int k = 0; k < 3; k++) {
            double a = fabs(diag[didx9(n, 3 * j + k)]);
            if (a > m) m = a;
        }
        s[j] = m;
    }
    const size_t q0 = 64 * (size_t)sell_ptr[n >> 6] + (n & 63);
    const int len = row_ptr[n + 1] - row_ptr[n];
    for (int b = 0; b < len; b++) {
        const size_t q = q0 + 64 * (size_t)b;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                double a = fabs(offd[sell_vidx(q, 3 * j + k)]);
                if (a > s[j]) s[j] = a;
            }
    }
    const double small_factor = 1.0 / sqrt(SOLV_TOL);
    for (int j = 0; j < 3; j++) sv[vidx3(n, j)] = s[j] < SOLV_TOL ? small_factor : 1.0 / sqrt(fabs(s[j]));
}

// Applies the scaling (A <- S A S, b <- S b, x <- x / S) and sets up the
// block-Jacobi inverses in one pass over the matrix.  F32: the scaled
// matrix and inverses are stored only as the FP32 copies the FP32 Krylov
// reads (rounded from the same FP64 values); BJ = false (profile-LU
// preconditioner): no inverses.
// K > 0: every slice has K slots (padding slots hold zeros and a valid
// column), run as a fully unrolled loop with all column scales gathered first.
// SRC32: the unscaled off-diagonal blocks are already in offd32 (the refined
// forms' assembly writes them there) and are scaled in place.
template <bool F32, bool BJ, int K, bool SRC32 = false>
__global__ void __launch_bounds__(256) k_scale_apply(int N, double *diag, double *offd, const int *row_ptr,
                                                     const int *sell_ptr, const int *sell_col, const double *sv,
                                                     double *R, double *sol, double *dinv, float *diag32,
                                                     float *offd32, float *dinv32) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double rf[3] = {sv[vidx3(n, 0)], sv[vidx3(n, 1)], sv[vidx3(n, 2)]};
    for (int j = 0; j < 3; j++) {
        R[vidx3(n, j)] *= rf[j];
        sol[vidx3(n, j)] /= rf[j];
    }
    double d[9];
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) {
            const size_t i = didx9(n, 3 * j + k);
            d[3 * j + k] = diag[i] * (rf[j] * rf[k]);
            if constexpr (F32) diag32[i] = (float)d[3 * j + k];
            else diag[i] = d[3 * j + k];
        }
    if constexpr (BJ) {
        double inv[9];
        inv3(d, inv);
        for (int e = 0; e < 9; e++) {
            if constexpr (F32) dinv32[didx9(n, e)] = (float)inv[e];
            else dinv[didx9(n, e)] = inv[e];
        }
    }
    const size_t q0 = 64 * (size_t)sell_ptr[n >> 6] + (n & 63);
    auto scale_slot = [&](size_t q, const double (&sc)[3]) {
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const size_t i = sell_vidx(q, 3 * j + k);
                const double v = (SRC32 ? (double)offd32[i] : offd[i]) * (rf[j] * sc[k]);
                if constexpr (F32) offd32[i] = (float)v;
                else offd[i] = v;
            }
    };
    if constexpr (K > 0) {
        int cn[K];
#pragma unroll
        for (int b = 0; b < K; b++) cn[b] = sell_col[q0 + 64 * b];
        double sc[K][3];
#pragma unroll
        for (int b = 0; b < K; b++)
#pragma unroll
            for (int k = 0; k < 3; k++) sc[b][k] = sv[vidx3(cn[b], k)];
#pragma unroll
        for (int b = 0; b < K; b++) scale_slot(q0 + 64 * (size_t)b, sc[b]);
    } else {
        const int len = row_ptr[n + 1] - row_ptr[n];
        for (int b = 0; b < len; b++) {
            const size_t q = q0 + 64 * (size_t)b;
            const int cn = sell_col[q];
            const double sc[3] = {sv[vidx3(cn, 0)], sv[vidx3(cn, 1)], sv[vidx3(cn, 2)]};
            scale_slot(q, sc);
        }
    }
}

__global__ void k_to_f32(size_t n, const double *a, float *b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// k_scale_apply dispatched on the uniform slice width (0 = per-row lengths)
#define SW2_SCALE_LAUNCH(F32, BJ, K)                                                                               \
    do {                                                                                                            \
        if (F32 && c->mat32)                                                                                        \
            hipLaunchKernelGGL((k_scale_apply<F32, BJ, K, true>), dim3(nb), dim3(256), 0, c->stream, c->N, c->diag, \
                               c->offd, c->row_ptr, c->sell_ptr, c->sell_col, c->scale, c->R, c->sol, c->dinv,      \
                               c->diag32, c->offd32, c->dinv32);                                                    \
        else                                                                                                        \
            hipLaunchKernelGGL((k_scale_apply<F32, BJ, K, false>), dim3(nb), dim3(256), 0, c->stream, c->N,         \
                               c->diag, c->offd, c->row_ptr, c->sell_ptr, c->sell_col, c->scale, c->R, c->sol,      \
                               c->dinv, c->diag32, c->offd32, c->dinv32);                                           \
    } while (0)
#define SW2_SCALE_SWITCH(F32, BJ)                           \
    switch (c->kuni) {                                      \
        case 5: SW2_SCALE_LAUNCH(F32, BJ, 5); break;        \
        case 6: SW2_SCALE_LAUNCH(F32, BJ, 6); break;        \
        case 7: SW2_SCALE_LAUNCH(F32, BJ, 7); break;        \
        case 8: SW2_SCALE_LAUNCH(F32, BJ, 8); break;        \
        default: SW2_SCALE_LAUNCH(F32, BJ, 0); break;       \
    }

void launch_setup(Ctx *c) {
    int nb = nblocks_for(c->N, 256);
    if (!c->scale_valid)   // a matrix the assembly did not produce: its row maxima first
        hipLaunchKernelGGL(k_scale_rows, dim3(nb), dim3(256), 0, c->stream, c->N, c->diag, c->offd, c->row_ptr,
                           c->sell_ptr, c->scale);
    c->scale_valid = 0;    // the scaling below consumes it (and rescales the FP64 matrix in place)
    c->sys_scaled = 1;
    if (c->opts.preconditioner == 1) {  // profile LU of the node blocks (FP64)
        SW2_SCALE_SWITCH(false, false)
        launch_band_setup(c);
        return;
    }
    if (c->opts.preconditioner == 3) {  // block ILU(0): scaled matrix without block inverses, then the factors
        if (c->opts.precision >= 1) { SW2_SCALE_SWITCH(true, false) }
        else { SW2_SCALE_SWITCH(false, false) }
        launch_ilu_setup(c);
        return;
    }
    if (c->opts.precision >= 1) { SW2_SCALE_SWITCH(true, true) }
    else { SW2_SCALE_SWITCH(false, true) }
}

// ------------------------------------------------------------ reductions
// Each reduction is split across a kernel boundary: every block of the
// producing kernel writes its K block partials (wavefront shuffles, then the
// block's waves in order), and a one-block finisher kernel reduces the
// partials in block order and updates the device scalars.  The kernel
// boundary publishes the partials; an in-kernel "last block" scheme would
// need a device-scope fence per block, which on a multi-XCD part writes back
// the XCD's L2 and costs more than the SpMV itself.
// A single model's Krylov kernels run their 256-node blocks in XCD-aware
// order (xcd_block): each XCD gets a contiguous range of blocks, so a row's
// neighbour gathers (a mesh column away: a few blocks) hit that XCD's own L2.
// Ensembles keep the plain order, which spreads the blocks of the models still
// iterating evenly over the XCDs.  Partial slots, the reduction order and
// every result are those of the plain order.
__device__ __forceinline__ int lblock(const int *blk_inst) { return blk_inst ? (int)blockIdx.x : xcd_block(); }

template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], const int (&op)[K], double *part, int lb) {
    __shared__ double sh[K][256 / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) {
            double y = __shfl_xor(x, o, 64);
            x = op[k] ? fmax(x, y) : x + y;
        }
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double x = sh[k][0];
            for (int w = 1; w < nw; w++) x = op[k] ? fmax(x, sh[k][w]) : x + sh[k][w];
            part[K * lb + k] = x;
        }
    }
}

// One-block reduction of nparts block partials: thread-strided sums in
// fixed order, then a fixed tree, so the result is deterministic.
template <int K>
__device__ __forceinline__ void reduce_partials(int k0, int k1, const double *part, const int (&op)[K],
                                                double (&out)[K]) {
    __shared__ double sh[K][256 / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.0;
    // U strided partials per thread are loaded together (one memory round
    // trip instead of U), then folded in ascending block order
    constexpr int U = 8;
    const int bd = blockDim.x;
    for (int b0 = k0 + threadIdx.x; b0 < k1; b0 += U * bd) {
        double x[U][K];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int k = 0; k < K; k++) x[u][k] = b0 + u * bd < k1 ? part[K * (b0 + u * bd) + k] : 0.0;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (b0 + u * bd < k1)
#pragma unroll
                for (int k = 0; k < K; k++) acc[k] = op[k] ? fmax(acc[k], x[u][k]) : acc[k] + x[u][k];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        double x = acc[k];
        for (int o = 32; o > 0; o >>= 1) {
            double y = __shfl_xor(x, o, 64);
            x = op[k] ? fmax(x, y) : x + y;
        }
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double x = sh[k][0];
            for (int w = 1; w < nw; w++) x = op[k] ? fmax(x, sh[k][w]) : x + sh[k][w];
            out[k] = x;
        }
    }
}

// finisher stages of the BiCGSTAB reductions (solv_bcgstab.c:117-290)
enum { FIN_P1, FIN_P2, FIN_K1, FIN_K2, FIN_K3 };

// one block per model (blockIdx.x), over that model's 256-node blocks
template <int STAGE>
__global__ void __launch_bounds__(256) k_fin(int nparts, const double *part, KScal *sc, const int *inst_blk0) {
    const int m = blockIdx.x;
    const int k0 = inst_blk0 ? inst_blk0[m] : 0, k1 = inst_blk0 ? inst_blk0[m + 1] : nparts;
    sc += m;
    if ((STAGE == FIN_K1 || STAGE == FIN_K2 || STAGE == FIN_K3) && sc->done) return;
    constexpr int K = STAGE == FIN_P2 ? 2 : (STAGE == FIN_K2 ? 3 : 1);
    int op[K];
    if constexpr (STAGE == FIN_P1 || STAGE == FIN_K3) op[0] = 1;
    else if constexpr (STAGE == FIN_P2) { op[0] = 1; op[1] = 0; }
    else if constexpr (STAGE == FIN_K1) op[0] = 0;
    else { op[0] = 0; op[1] = 0; op[2] = 0; }
    double out[K];
    reduce_partials<K>(k0, k1, part, op, out);
    if (threadIdx.x != 0) return;
    if constexpr (STAGE == FIN_P1) {
        sc->bnorm = out[0];
    } else if constexpr (STAGE == FIN_P2) {
        sc->rnorm = out[0];
        sc->rho = out[1];
        sc->rhop = 1.0;
        sc->alpha = 1.0;
        sc->omega = 1.0;
        sc->tol = sc->atol + sc->bnorm * sc->rtol;
        sc->it = 0;
        sc->done = !(sc->rnorm > sc->tol && 0 < sc->max_it);
    } else if constexpr (STAGE == FIN_K1) {
        sc->gamma = out[0];
        sc->alpha = sc->rho / sc->gamma;
    } else if constexpr (STAGE == FIN_K2) {
        double omega = out[0] / out[1];
        sc->omega = omega;
        sc->rhop = sc->rho;
        sc->rho = -1.0 * omega * out[2];
    } else {
        sc->rnorm = out[0];
        sc->it += 1;
        sc->done = !(sc->rnorm > sc->tol && sc->it < sc->max_it);
    }
}

// ------------------------------------------------------------ SpMV + M^-1
// y = M^-1 (A x) for block row n; A = diag + offd in the reference's
// accumulation order (solv_amult.c:99-123: diagonal block, then the row's
// blocks in insertion order summed into a separate partial).
// Off-diagonal blocks come from the SELL-64 slice of this row's wavefront:
// every lane runs the slice's K block slots (zero padding past the row's
// length), and entry e of slot j is one coalesced 64-value read per wave.
// Precision tags of the Krylov solve: matrix/preconditioner storage type M
// and vector/arithmetic type V (a float matrix entry times a double vector
// entry is formed in double).
struct PrecF64 { using M = double; using V = double; };   // precision 0
struct PrecMix { using M = float; using V = double; };    // precision 1: FP32 matrix, FP64 vectors
struct PrecF32 { using M = float; using V = float; };     // precision 2: all FP32

template <typename P, int K, typename XF, bool BJ = true>
__device__ __forceinline__ void amult_prec_row(int n, const typename P::M *diag, const typename P::M *offd,
                                               const int *sell_ptr, const int *sell_col, const typename P::M *dinv,
                                               XF xf, typename P::V (&y)[3]) {
    using T = typename P::V;
    using MT = typename P::M;
    const MT *D = diag + didx9(n, 0);  // entry e at D[64 e]
    T x0 = xf(vidx3(n, 0)), x1 = xf(vidx3(n, 1)), x2 = xf(vidx3(n, 2));
    T p0 = 0, p1 = 0, p2 = 0;
    const int s = n >> 6, lane = n & 63;
    if constexpr (K > 0) {
        // uniform slices: all K slots' columns, gathers and blocks are
        // independent loads, issued before any of them is consumed
        const size_t q0 = 64 * (size_t)s * K + lane;
        const MT *A = offd + 9 * 64 * (size_t)s * K + lane;
        int cn[K];
#pragma unroll
        for (int b = 0; b < K; b++) cn[b] = sell_col[q0 + 64 * b];
        T g[K][3];
#pragma unroll
        for (int b = 0; b < K; b++) {
            const size_t cc = vidx3(cn[b], 0);
            g[b][0] = xf(cc);
            g[b][1] = xf(cc + 64);
            g[b][2] = xf(cc + 128);
        }
#pragma unroll
        for (int b = 0; b < K; b++) {
            const MT *a = A + 576 * b;
            p0 += ldnt(a) * g[b][0] + ldnt(a + 64) * g[b][1] + ldnt(a + 128) * g[b][2];
            p1 += ldnt(a + 192) * g[b][0] + ldnt(a + 256) * g[b][1] + ldnt(a + 320) * g[b][2];
            p2 += ldnt(a + 384) * g[b][0] + ldnt(a + 448) * g[b][1] + ldnt(a + 512) * g[b][2];
        }
    } else {
        const int k0 = sell_ptr[s], k1 = sell_ptr[s + 1];
        for (int b = k0; b < k1; b++) {
            const size_t q = 64 * (size_t)b + lane;
            const MT *a = offd + 9 * 64 * (size_t)b + lane;
            const int cn = sell_col[q];
            const size_t cc = vidx3(cn, 0);
            T y0 = xf(cc), y1 = xf(cc + 64), y2 = xf(cc + 128);
            p0 += a[0] * y0 + a[64] * y1 + a[128] * y2;
            p1 += a[192] * y0 + a[256] * y1 + a[320] * y2;
            p2 += a[384] * y0 + a[448] * y1 + a[512] * y2;
        }
    }
    T a0 = ldnt(D) * x0 + ldnt(D + 64) * x1 + ldnt(D + 128) * x2;
    T a1 = ldnt(D + 192) * x0 + ldnt(D + 256) * x1 + ldnt(D + 320) * x2;
    T a2 = ldnt(D + 384) * x0 + ldnt(D + 448) * x1 + ldnt(D + 512) * x2;
    a0 += p0; a1 += p1; a2 += p2;
    if constexpr (!BJ) {
        y[0] = a0; y[1] = a1; y[2] = a2;
        return;
    }
    const MT *M = dinv + didx9(n, 0);
    y[0] = ldnt(M) * a0 + ldnt(M + 64) * a1 + ldnt(M + 128) * a2;
    y[1] = ldnt(M + 192) * a0 + ldnt(M + 256) * a1 + ldnt(M + 320) * a2;
    y[2] = ldnt(M + 384) * a0 + ldnt(M + 448) * a1 + ldnt(M + 512) * a2;
}

template <typename T>
struct VecX {
    const T *x;
    __device__ T operator()(size_t i) const { return x[i]; }
};
template <typename T>
struct VecS {  // s = r + (-alpha) Ap, the reference's copy + daxpy
    const T *r, *ap;
    T nalpha;
    __device__ T operator()(size_t i) const { return r[i] + nalpha * ap[i]; }
};

template <typename P>
struct KVecs {
    const typename P::M *diag, *offd, *dinv;
    typename P::V *b, *r, *p, *Ap, *s, *As, *q, *u;
};

// prologue part 1: b' = M^-1 b, u = b', ||b'||inf  (solv_bcgstab.c:117-137)
template <typename P>
__global__ void __launch_bounds__(256) k_bcg_p1(int N, KVecs<P> V, const double *rhs, double *part, const int *blk_inst,
                                                KScal *sc) {
    using T = typename P::V;
    using MT = typename P::M;
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    int n = lb * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (n < nend) {
        const MT *M = V.dinv + didx9(n, 0);
        T b0 = (T)rhs[vidx3(n, 0)], b1 = (T)rhs[vidx3(n, 1)], b2 = (T)rhs[vidx3(n, 2)];
        T y[3];
        y[0] = M[0] * b0 + M[64] * b1 + M[128] * b2;
        y[1] = M[192] * b0 + M[256] * b1 + M[320] * b2;
        y[2] = M[384] * b0 + M[448] * b1 + M[512] * b2;
        for (int j = 0; j < 3; j++) {
            V.b[vidx3(n, j)] = y[j];
            V.u[vidx3(n, j)] = y[j];
            v[0] = fmax(v[0], fabs((double)y[j]));
        }
    }
    const int op[1] = {1};
    block_partials<1>(v, op, part, lb);
}

// prologue part 2: r = b' - M^-1 A u, q = r, p = r, rho = (r,q), ||r||inf  (:138-188)
template <typename P, int K>
__global__ void __launch_bounds__(256) k_bcg_p2(int N, KVecs<P> V, const int *sell_ptr, const int *sell_col, double *part, const int *blk_inst, KScal *sc) {
    using T = typename P::V;
    using MT = typename P::M;
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    int n = lb * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (n < nend) {
        T y[3];
        amult_prec_row<P, K>(n, V.diag, V.offd, sell_ptr, sell_col, V.dinv, VecX<T>{V.u}, y);
        for (int j = 0; j < 3; j++) {
            const size_t i = vidx3(n, j);
            T r = V.b[i] + (T)(-1.0) * y[j];
            V.r[i] = r;
            V.q[i] = r;
            V.p[i] = r;
            v[0] = fmax(v[0], fabs((double)r));
            v[1] += (double)r * (double)r;
        }
    }
    const int op[2] = {1, 0};
    block_partials<2>(v, op, part, lb);
}

template <typename P, int K>
__global__ void __launch_bounds__(256) k_bcg_k1(int N, KVecs<P> V, const int *sell_ptr, const int *sell_col, double *part, const int *blk_inst, KScal *sc) {
    using T = typename P::V;
    using MT = typename P::M;
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    if (sc->done) return;
    int n = lb * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (n < nend) {
        T y[3];
        amult_prec_row<P, K>(n, V.diag, V.offd, sell_ptr, sell_col, V.dinv, VecX<T>{V.p}, y);
        for (int j = 0; j < 3; j++) {
            const size_t i = vidx3(n, j);
            V.Ap[i] = y[j];
            v[0] += (double)V.q[i] * (double)y[j];
        }
    }
    const int op[1] = {0};
    block_partials<1>(v, op, part, lb);
}

template <typename P, int K>
__global__ void __launch_bounds__(256) k_bcg_k2(int N, KVecs<P> V, const int *sell_ptr, const int *sell_col, double *part, const int *blk_inst, KScal *sc) {
    using T = typename P::V;
    using MT = typename P::M;
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    if (sc->done) return;
    int n = lb * blockDim.x + threadIdx.x;
    const T nalpha = (T)(-sc->alpha);
    double v[3] = {0.0, 0.0, 0.0};
    if (n < nend) {
        VecS<T> sf{V.r, V.Ap, nalpha};
        T y[3];
        amult_prec_row<P, K>(n, V.diag, V.offd, sell_ptr, sell_col, V.dinv, sf, y);
        for (int j = 0; j < 3; j++) {
            const size_t i = vidx3(n, j);
            T s = sf(i);
            V.s[i] = s;
            V.As[i] = y[j];
            v[0] += (double)y[j] * (double)s;
            v[1] += (double)y[j] * (double)y[j];
            v[2] += (double)y[j] * (double)V.q[i];
        }
    }
    const int op[3] = {0, 0, 0};
    block_partials<3>(v, op, part, lb);
}

template <typename P>
__global__ void __launch_bounds__(256) k_bcg_k3(int N, KVecs<P> V, double *part, const int *blk_inst, KScal *sc) {
    using T = typename P::V;
    using MT = typename P::M;
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    if (sc->done) return;
    int n = lb * blockDim.x + threadIdx.x;
    const double alpha = sc->alpha, omega = sc->omega;
    const double beta = (sc->rho * alpha) / (sc->rhop * omega);
    const T ta = (T)alpha, to = (T)omega, tno = (T)(-omega), tb = (T)beta;
    double v[1] = {0.0};
    if (n < nend) {
        for (int j = 0; j < 3; j++) {
            const size_t i = vidx3(n, j);
            T s = V.s[i], p = V.p[i];
            T u = V.u[i];
            u += ta * p;
            u += to * s;
            V.u[i] = u;
            T r = s + tno * V.As[i];
            V.r[i] = r;
            p += tno * V.Ap[i];
            p *= tb;
            p += (T)1.0 * r;
            V.p[i] = p;
            v[0] = fmax(v[0], fabs((double)r));
        }
    }
    const int op[1] = {1};
    block_partials<1>(v, op, part, lb);
}

// ---------------------------------------------------------------------------
// BiCGSTAB with a preconditioner applied after the SpMV: the profile LU
// (FP64, sw2_blk.hip) or the block ILU(0) (FP64 or FP32 factors,
// sw2_ilu.hip).  Same recurrences as K1..K3 above; the product M^-1 A x is
// split into a plain SpMV (the matrix in the form's storage, FP64 vectors)
// and the preconditioner's triangular solves, so each half-iteration is
// SpMV, solve, then a reduction kernel.  K3 is shared.

// y = A x with x = p (S = false) or x = s = r - alpha Ap (S = true, s stored)
template <typename P, int K, bool S>
__global__ void __launch_bounds__(256) k_pl_spmv(int N, KVecs<P> V, const int *sell_ptr, const int *sell_col,
                                                 const double *x, double *y, const int *blk_inst, const KScal *sc,
                                                 int check_done) {
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    if (blk_inst) { sc += blk_inst[2 * blockIdx.x]; nend = blk_inst[2 * blockIdx.x + 1]; }
    if (check_done && sc->done) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nend) return;
    double a[3];
    if constexpr (S) {
        VecS<double> sf{V.r, V.Ap, -sc->alpha};
        amult_prec_row<P, K, VecS<double>, false>(n, V.diag, V.offd, sell_ptr, sell_col, nullptr, sf, a);
        for (int j = 0; j < 3; j++) V.s[vidx3(n, j)] = sf(vidx3(n, j));
    } else {
        amult_prec_row<P, K, VecX<double>, false>(n, V.diag, V.offd, sell_ptr, sell_col, nullptr,
                                                 VecX<double>{x}, a);
    }
    for (int j = 0; j < 3; j++) y[vidx3(n, j)] = a[j];
}

// b = rhs (the reference solves in place on its right-hand side)
__global__ void k_pl_copy(size_t n, const double *src, double *dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// u = b' (= M^-1 b), ||b'||inf
template <typename P>
__global__ void __launch_bounds__(256) k_pl_p1(int N, KVecs<P> V, double *part, const int *blk_inst, KScal *sc) {
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    const int n = lb * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (n < nend)
        for (int j = 0; j < 3; j++) {
            const size_t i = vidx3(n, j);
            const double y = V.b[i];
            V.u[i] = y;
            v[0] = fmax(v[0], fabs(y));
        }
    const int op[1] = {1};
    block_partials<1>(v, op, part, lb);
}

// r = b' - y (y = M^-1 A u in As), q = p = r, rho, ||r||inf, scalars
template <typename P>
__global__ void __launch_bounds__(256) k_pl_p2(int N, KVecs<P> V, double *part, const int *blk_inst, KScal *sc) {
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    const int n = lb * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (n < nend)
        for (int j = 0; j < 3; j++) {
            const size_t i = vidx3(n, j);
            const double r = V.b[i] + (-1.0) * V.As[i];
            V.r[i] = r;
            V.q[i] = r;
            V.p[i] = r;
            v[0] = fmax(v[0], fabs(r));
            v[1] += r * r;
        }
    const int op[2] = {1, 0};
    block_partials<2>(v, op, part, lb);
}

// gamma = (q, Ap), alpha = rho / gamma
template <typename P>
__global__ void __launch_bounds__(256) k_pl_gamma(int N, KVecs<P> V, double *part, const int *blk_inst, KScal *sc) {
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    if (sc->done) return;
    const int n = lb * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (n < nend)
        for (int j = 0; j < 3; j++) v[0] += V.q[vidx3(n, j)] * V.Ap[vidx3(n, j)];
    const int op[1] = {0};
    block_partials<1>(v, op, part, lb);
}

// omega = (As, s) / (As, As), rho' = -omega (As, q)
template <typename P>
__global__ void __launch_bounds__(256) k_pl_omega(int N, KVecs<P> V, double *part, const int *blk_inst, KScal *sc) {
    int nend = N;  // ensembles: the block's model and that model's last node + 1
    const int lb = lblock(blk_inst);   // XCD-aware block order
    if (blk_inst) { sc += blk_inst[2 * lb]; nend = blk_inst[2 * lb + 1]; }
    if (sc->done) return;
    const int n = lb * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (n < nend)
        for (int j = 0; j < 3; j++) {
            const size_t i = vidx3(n, j);
            const double y = V.As[i];
            v[0] += y * V.s[i];
            v[1] += y * y;
            v[2] += y * V.q[i];
        }
    const int op[3] = {0, 0, 0};
    block_partials<3>(v, op, part, lb);
}

// solv_return_old (src/solver/solv_return_old.c:23-30): sol = (u + u0) * scale,
// u0 the initial guess held in sol.  ZERO: the caller zeroed the guess (the
// Newton step does), so u0 = +0 enters as a constant and the kernel gives the
// same sol however often it runs.
template <typename T, bool ZERO>
__global__ void k_bcg_epi(int N3, const T *u, const double *scale, double *sol) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N3) return;
    double x = (double)u[i];
    x += ZERO ? 0.0 : sol[i];
    x *= scale[i];
    sol[i] = x;
}

template <typename P>
static KVecs<P> vecs(Ctx *c);
template <>
KVecs<PrecF64> vecs<PrecF64>(Ctx *c) {
    return KVecs<PrecF64>{c->diag, c->offd, c->dinv, c->kb, c->kr, c->kp, c->kAp, c->ks_, c->kAs, c->kq, c->ku};
}
template <>
KVecs<PrecMix> vecs<PrecMix>(Ctx *c) {
    return KVecs<PrecMix>{c->diag32, c->offd32, c->dinv32, c->kb, c->kr, c->kp, c->kAp, c->ks_, c->kAs, c->kq, c->ku};
}
template <>
KVecs<PrecF32> vecs<PrecF32>(Ctx *c) {
    return KVecs<PrecF32>{c->diag32, c->offd32, c->dinv32, c->fb, c->fr, c->fp, c->fAp, c->fs, c->fAs, c->fq, c->fu};
}

template <typename P, int K>
static void prologue_k(Ctx *c) {
    int nb = nblocks_for(c->N, 256);
    KVecs<P> V = vecs<P>(c);
    hipLaunchKernelGGL(k_bcg_p1<P>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->R, c->red.part, BI(c),
                       c->ksc);
    hipLaunchKernelGGL(k_fin<FIN_P1>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
    hipLaunchKernelGGL((k_bcg_p2<P, K>), dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr, c->sell_col,
                       c->red.part, BI(c), c->ksc);
    hipLaunchKernelGGL(k_fin<FIN_P2>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
}

template <typename P, int K>
static void iters_k(Ctx *c, int k) {
    int nb = nblocks_for(c->N, 256);
    KVecs<P> V = vecs<P>(c);
    for (int i = 0; i < k; i++) {
        hipLaunchKernelGGL((k_bcg_k1<P, K>), dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr, c->sell_col,
                           c->red.part, BI(c), c->ksc);
        hipLaunchKernelGGL(k_fin<FIN_K1>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
        hipLaunchKernelGGL((k_bcg_k2<P, K>), dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr, c->sell_col,
                           c->red.part, BI(c), c->ksc);
        hipLaunchKernelGGL(k_fin<FIN_K2>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
        hipLaunchKernelGGL(k_bcg_k3<P>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->red.part, BI(c),
                           c->ksc);
        hipLaunchKernelGGL(k_fin<FIN_K3>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
    }
}

template <typename P, int K>
static void k1_only_k(Ctx *c) {
    int nb = nblocks_for(c->N, 256);
    hipLaunchKernelGGL((k_bcg_k1<P, K>), dim3(nb), dim3(256), 0, c->stream, c->N, vecs<P>(c), c->sell_ptr,
                       c->sell_col, c->red.part, BI(c), c->ksc);
}

// dispatch on the uniform slice width (0 = per-slice widths, looped)
#define SW2_KSWITCH(FN, T, ...)                          \
    switch (c->kuni) {                                   \
        case 3: FN<T, 3>(__VA_ARGS__); break;            \
        case 4: FN<T, 4>(__VA_ARGS__); break;            \
        case 5: FN<T, 5>(__VA_ARGS__); break;            \
        case 6: FN<T, 6>(__VA_ARGS__); break;            \
        case 7: FN<T, 7>(__VA_ARGS__); break;            \
        case 8: FN<T, 8>(__VA_ARGS__); break;            \
        case 9: FN<T, 9>(__VA_ARGS__); break;            \
        case 10: FN<T, 10>(__VA_ARGS__); break;          \
        default: FN<T, 0>(__VA_ARGS__); break;           \
    }

// the preconditioner applied after an SpMV: profile LU or ILU(0), in place
static void prec_solve(Ctx *c, double *x, const KScal *sc) {
    if (c->opts.preconditioner == 3) launch_ilu_solve(c, x, sc);
    else launch_band_solve(c, x, sc);
}

template <typename P, int K>
static void pl_prologue_k(Ctx *c) {
    const int nb = nblocks_for(c->N, 256);
    KVecs<P> V = vecs<P>(c);
    const size_t n3 = 3 * (size_t)c->Np;
    hipLaunchKernelGGL(k_pl_copy, dim3(nblocks_for((int)n3, 256)), dim3(256), 0, c->stream, n3, c->R, V.b);
    prec_solve(c, V.b, nullptr);
    hipLaunchKernelGGL(k_pl_p1<P>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->red.part, BI(c), c->ksc);
    hipLaunchKernelGGL(k_fin<FIN_P1>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
    hipLaunchKernelGGL((k_pl_spmv<P, K, false>), dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr,
                       c->sell_col, V.u, V.As, BI(c), c->ksc, 0);
    prec_solve(c, V.As, nullptr);
    hipLaunchKernelGGL(k_pl_p2<P>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->red.part, BI(c), c->ksc);
    hipLaunchKernelGGL(k_fin<FIN_P2>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
}

template <typename P, int K>
static void pl_iters_k(Ctx *c, int k) {
    const int nb = nblocks_for(c->N, 256);
    KVecs<P> V = vecs<P>(c);
    for (int i = 0; i < k; i++) {
        hipLaunchKernelGGL((k_pl_spmv<P, K, false>), dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr,
                           c->sell_col, V.p, V.Ap, BI(c), c->ksc, 1);
        prec_solve(c, V.Ap, c->ksc);
        hipLaunchKernelGGL(k_pl_gamma<P>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->red.part, BI(c),
                           c->ksc);
        hipLaunchKernelGGL(k_fin<FIN_K1>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
        hipLaunchKernelGGL((k_pl_spmv<P, K, true>), dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr,
                           c->sell_col, nullptr, V.As, BI(c), c->ksc, 1);
        prec_solve(c, V.As, c->ksc);
        hipLaunchKernelGGL(k_pl_omega<P>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->red.part, BI(c),
                           c->ksc);
        hipLaunchKernelGGL(k_fin<FIN_K2>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
        hipLaunchKernelGGL(k_bcg_k3<P>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->red.part, BI(c), c->ksc);
        hipLaunchKernelGGL(k_fin<FIN_K3>, dim3(NFIN(c)), dim3(256), 0, c->stream, nb, c->red.part, c->ksc, IB0(c));
    }
}

// the post-SpMV preconditioners run with FP64 Krylov vectors only (the
// options refuse them with precision 2)
template <typename P>
static void prologue(Ctx *c) {
    if constexpr (std::is_same<typename P::V, double>::value) {
        if (c->opts.preconditioner != 2) {
            SW2_KSWITCH(pl_prologue_k, P, c)
            return;
        }
    }
    SW2_KSWITCH(prologue_k, P, c)
}

template <typename P>
static void iters(Ctx *c, int k) {
    if constexpr (std::is_same<typename P::V, double>::value) {
        if (c->opts.preconditioner != 2) {
            SW2_KSWITCH(pl_iters_k, P, c, k)
            return;
        }
    }
    SW2_KSWITCH(iters_k, P, c, k)
}

void launch_k1_only(Ctx *c) {
    if (c->opts.precision == 2) { SW2_KSWITCH(k1_only_k, PrecF32, c) }
    else if (c->opts.precision == 1) { SW2_KSWITCH(k1_only_k, PrecMix, c) }
    else { SW2_KSWITCH(k1_only_k, PrecF64, c) }
}

void launch_krylov_prologue(Ctx *c) {
    if (c->opts.precision == 2) prologue<PrecF32>(c);
    else if (c->opts.precision == 1) prologue<PrecMix>(c);
    else prologue<PrecF64>(c);
}

void launch_krylov_iters(Ctx *c, int k) {
    if (c->opts.precision == 2) iters<PrecF32>(c, k);
    else if (c->opts.precision == 1) iters<PrecMix>(c, k);
    else iters<PrecF64>(c, k);
}

void launch_krylov_epilogue(Ctx *c, bool zero_guess) {
    int n3 = 3 * c->Np;
    dim3 g(nblocks_for(n3, 256));
    if (c->opts.precision == 2) {
        if (zero_guess) hipLaunchKernelGGL((k_bcg_epi<float, true>), g, dim3(256), 0, c->stream, n3, c->fu, c->scale, c->sol);
        else hipLaunchKernelGGL((k_bcg_epi<float, false>), g, dim3(256), 0, c->stream, n3, c->fu, c->scale, c->sol);
    } else {
        if (zero_guess) hipLaunchKernelGGL((k_bcg_epi<double, true>), g, dim3(256), 0, c->stream, n3, c->ku, c->scale, c->sol);
        else hipLaunchKernelGGL((k_bcg_epi<double, false>), g, dim3(256), 0, c->stream, n3, c->ku, c->scale, c->sol);
    }
}

}  // namespace sw2

// ===========================================================================
// Restarted GMRES(m), left-preconditioned (opts.krylov = 1).  No counterpart
// in the reference's serial path, which solves with BiCGSTAB only
// (solv_linear_sys_solve.c -> solv_bcgstab.c); it is the north star's
// alternative Krylov method and uses the same SpMV, preconditioners and
// scaled system.  Arnoldi with classical Gram-Schmidt applied twice (CGS2):
// each pass is one multi-dot kernel (all (v_i, w) of the basis at once, block
// partials reduced in block order by a one-block finisher) and one update
// kernel, so an iteration reads the basis twice instead of once per vector.
// The small Hessenberg least-squares problem (Givens rotations) is solved on
// the host from the coefficients copied back once per iteration.  Stop: the
// 2-norm of the preconditioned residual |g_{j+1}| <= lin_atol +
// lin_rtol * ||M^-1 b||_2 (BiCGSTAB's rule with the Arnoldi residual in place
// of its max-norm), or lin_it_cap iterations in total.
namespace sw2 {

constexpr int GM_RESTART = 30;              // Krylov vectors per cycle
constexpr int GM_MAX = GM_RESTART + 1;      // basis vectors stored

// (v_i, w) for i < nv: per-block partials, i at part[GM_MAX * block + i]
__global__ void __launch_bounds__(256) k_gm_mdot(int N, const double *V, size_t ld, int nv, const double *w,
                                                 double *part) {
    __shared__ double sh[GM_MAX][4];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double w0 = 0.0, w1 = 0.0, w2 = 0.0;
    if (n < N) { w0 = w[vidx3(n, 0)]; w1 = w[vidx3(n, 1)]; w2 = w[vidx3(n, 2)]; }
    for (int i = 0; i < nv; i++) {
        double x = 0.0;
        if (n < N) {
            const double *v = V + ld * i;
            x = v[vidx3(n, 0)] * w0;
            x += v[vidx3(n, 1)] * w1;
            x += v[vidx3(n, 2)] * w2;
        }
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) sh[i][wid] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        double s = sh[threadIdx.x][0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) s += sh[threadIdx.x][k];
        part[GM_MAX * (size_t)blockIdx.x + threadIdx.x] = s;
    }
}

// out[i] = sum of the block partials of value i, in block order (one block)
__global__ void __launch_bounds__(256) k_gm_fin(int nparts, const double *part, int nv, double *out) {
    __shared__ double sh[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = 0; i < nv; i++) {
        double x = 0.0;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) x += part[GM_MAX * (size_t)b + i];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) sh[wid] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = sh[0];
            for (int k = 1; k < (int)(blockDim.x >> 6); k++) s += sh[k];
            out[i] = s;
        }
        __syncthreads();
    }
}

// w -= sum_{i < nv} h_i v_i; with `norm`, also the block partials of (w, w) at value 0
__global__ void __launch_bounds__(256) k_gm_update(int N, const double *V, size_t ld, int nv, const double *h,
                                                   double *w, double *part, int norm) {
    __shared__ double sh[4];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    double x = 0.0;
    if (n < N) {
        for (int k = 0; k < 3; k++) {
            const size_t q = vidx3(n, k);
            double a = w[q];
            for (int i = 0; i < nv; i++) a -= h[i] * V[ld * i + q];
            w[q] = a;
            x += a * a;
        }
    }
    if (!norm) return;
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = sh[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) s += sh[k];
        part[GM_MAX * (size_t)blockIdx.x] = s;
    }
}

// v = w / sqrt(nrm2[0])  (the next basis vector)
__global__ void k_gm_scale(size_t n3, const double *w, const double *nrm2, double *v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) v[i] = w[i] / sqrt(nrm2[0]);
}

// r = b - y (restart residual), y in place
__global__ void k_gm_sub(size_t n3, const double *b, double *y) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) y[i] = b[i] - y[i];
}

// x += sum_{i < nv} y_i v_i
__global__ void k_gm_xupd(size_t n3, const double *V, size_t ld, int nv, const double *y, double *x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double a = x[i];
    for (int k = 0; k < nv; k++) a += y[k] * V[ld * k + i];
    x[i] = a;
}

// x <- D^-1 x with the block-Jacobi inverses (storage type MT)
template <typename MT>
__global__ void k_bj_apply(int N, const MT *dinv, double *x) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double a0 = x[vidx3(n, 0)], a1 = x[vidx3(n, 1)], a2 = x[vidx3(n, 2)];
    const MT *M = dinv + didx9(n, 0);
    x[vidx3(n, 0)] = M[0] * a0 + M[64] * a1 + M[128] * a2;
    x[vidx3(n, 1)] = M[192] * a0 + M[256] * a1 + M[320] * a2;
    x[vidx3(n, 2)] = M[384] * a0 + M[448] * a1 + M[512] * a2;
}

static void gm_prec(Ctx *c, double *x) {
    if (c->opts.preconditioner == 2) {
        if (c->opts.precision >= 1)
            hipLaunchKernelGGL(k_bj_apply<float>, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N,
                               c->dinv32, x);
        else
            hipLaunchKernelGGL(k_bj_apply<double>, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N,
                               c->dinv, x);
    } else {
        prec_solve(c, x, nullptr);
    }
}

template <typename P, int K>
static void gm_spmv_k(Ctx *c, const double *x, double *y) {
    hipLaunchKernelGGL((k_pl_spmv<P, K, false>), dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N,
                       vecs<P>(c), c->sell_ptr, c->sell_col, x, y, nullptr, c->ksc, 0);
}

// y = M^-1 A x
static void gm_op(Ctx *c, const double *x, double *y) {
    if (c->opts.precision >= 1) { SW2_KSWITCH(gm_spmv_k, PrecMix, c, x, y) }
    else { SW2_KSWITCH(gm_spmv_k, PrecF64, c, x, y) }
    gm_prec(c, y);
}

// ||w||_2^2 of a vector into dev[0] (via the update kernel with no basis vectors)
static void gm_norm2(Ctx *c, double *w, double *dev) {
    const int nb = nblocks_for(c->N, 256);
    hipLaunchKernelGGL(k_gm_update, dim3(nb), dim3(256), 0, c->stream, c->N, (const double *)nullptr, (size_t)0, 0,
                       (const double *)nullptr, w, c->gm_part, 1);
    hipLaunchKernelGGL(k_gm_fin, dim3(1), dim3(256), 0, c->stream, nb, c->gm_part, 1, dev);
}

std::string gmres_alloc(Ctx *c) {
    if (c->gm_V) return "";
    const size_t ld = 3 * (size_t)c->Np;
    const int nb = nblocks_for(c->N, 256);
    if (hipMalloc((void **)&c->gm_V, sizeof(double) * ld * GM_MAX) != hipSuccess ||
        hipMalloc((void **)&c->gm_part, sizeof(double) * GM_MAX * (size_t)(nb + 8)) != hipSuccess ||
        hipMalloc((void **)&c->gm_dev, sizeof(double) * 4 * GM_MAX) != hipSuccess ||
        hipHostMalloc((void **)&c->gm_host, sizeof(double) * 4 * GM_MAX) != hipSuccess) {
        gmres_free(c);
        return "hipMalloc failed for the GMRES basis";
    }
    (void)hipMemset(c->gm_V, 0, sizeof(double) * ld * GM_MAX);
    return "";
}

void gmres_free(Ctx *c) {
    if (c->gm_V) (void)hipFree(c->gm_V);
    if (c->gm_part) (void)hipFree(c->gm_part);
    if (c->gm_dev) (void)hipFree(c->gm_dev);
    if (c->gm_host) (void)hipHostFree(c->gm_host);
    c->gm_V = c->gm_part = c->gm_dev = c->gm_host = nullptr;
}

// Solves the scaled system into ku (x, zero initial guess); returns the
// iteration count and whether the stop rule was met.  hip errors surface
// through the caller's hipGetLastError / synchronisation.
int launch_gmres(Ctx *c, int max_it, bool *converged) {
    const size_t ld = 3 * (size_t)c->Np, n3 = ld;
    const int nb = nblocks_for(c->N, 256), nbv = nblocks_for((int)n3, 256);
    double *V = c->gm_V, *w = c->kAp, *bp = c->kb, *x = c->ku;
    double *c1 = c->gm_dev, *c2 = c->gm_dev + GM_MAX, *nrm = c->gm_dev + 2 * GM_MAX, *ydev = c->gm_dev + 3 * GM_MAX;
    double *hh = c->gm_host;
    // b' = M^-1 b, x = 0
    hipLaunchKernelGGL(k_pl_copy, dim3(nbv), dim3(256), 0, c->stream, n3, c->R, bp);
    gm_prec(c, bp);
    launch_zero(c, x, n3);
    gm_norm2(c, bp, nrm);
    (void)hipMemcpyAsync(hh, nrm, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    const double bnorm = std::sqrt(hh[0]);
    const double tol = c->opts.lin_atol + c->opts.lin_rtol * bnorm;
    *converged = bnorm <= tol;
    int it = 0;
    double H[GM_MAX][GM_RESTART], cs[GM_RESTART], sn[GM_RESTART], g[GM_MAX];
    bool first = true;
    while (!*converged && it < max_it) {
        // r = b' - M^-1 A x (the first cycle: r = b'), v0 = r / ||r||
        if (first) {
            hipLaunchKernelGGL(k_pl_copy, dim3(nbv), dim3(256), 0, c->stream, n3, bp, w);
        } else {
            gm_op(c, x, w);
            hipLaunchKernelGGL(k_gm_sub, dim3(nbv), dim3(256), 0, c->stream, n3, bp, w);
        }
        gm_norm2(c, w, nrm);
        hipLaunchKernelGGL(k_gm_scale, dim3(nbv), dim3(256), 0, c->stream, n3, w, nrm, V);
        (void)hipMemcpyAsync(hh, nrm, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        (void)hipStreamSynchronize(c->stream);
        const double beta = first ? bnorm : std::sqrt(hh[0]);
        first = false;
        if (beta <= tol) { *converged = true; break; }
        for (int i = 0; i < GM_MAX; i++) g[i] = 0.0;
        g[0] = beta;
        int j = 0;
        for (; j < GM_RESTART && it < max_it; j++) {
            const int nv = j + 1;
            gm_op(c, V + ld * j, w);
            hipLaunchKernelGGL(k_gm_mdot, dim3(nb), dim3(256), 0, c->stream, c->N, V, ld, nv, w, c->gm_part);
            hipLaunchKernelGGL(k_gm_fin, dim3(1), dim3(256), 0, c->stream, nb, c->gm_part, nv, c1);
            hipLaunchKernelGGL(k_gm_update, dim3(nb), dim3(256), 0, c->stream, c->N, V, ld, nv, c1, w, c->gm_part, 0);
            hipLaunchKernelGGL(k_gm_mdot, dim3(nb), dim3(256), 0, c->stream, c->N, V, ld, nv, w, c->gm_part);
            hipLaunchKernelGGL(k_gm_fin, dim3(1), dim3(256), 0, c->stream, nb, c->gm_part, nv, c2);
            hipLaunchKernelGGL(k_gm_update, dim3(nb), dim3(256), 0, c->stream, c->N, V, ld, nv, c2, w, c->gm_part, 1);
            hipLaunchKernelGGL(k_gm_fin, dim3(1), dim3(256), 0, c->stream, nb, c->gm_part, 1, nrm);
            hipLaunchKernelGGL(k_gm_scale, dim3(nbv), dim3(256), 0, c->stream, n3, w, nrm, V + ld * (j + 1));
            (void)hipMemcpyAsync(hh, c1, sizeof(double) * (2 * GM_MAX + 1), hipMemcpyDeviceToHost, c->stream);
            (void)hipStreamSynchronize(c->stream);
            it++;
            for (int i = 0; i < nv; i++) H[i][j] = hh[i] + hh[GM_MAX + i];
            const double hn = std::sqrt(hh[2 * GM_MAX]);   // ||w|| after orthogonalisation
            H[nv][j] = hn;
            for (int i = 0; i < j; i++) {   // previous rotations
                const double a = H[i][j], b = H[i + 1][j];
                H[i][j] = cs[i] * a + sn[i] * b;
                H[i + 1][j] = -sn[i] * a + cs[i] * b;
            }
            const double a = H[j][j], b = H[j + 1][j];
            const double r = std::hypot(a, b);
            cs[j] = r > 0.0 ? a / r : 1.0;
            sn[j] = r > 0.0 ? b / r : 0.0;
            H[j][j] = r;
            H[j + 1][j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            // stop on the tolerance, or on a breakdown (w in the span: the
            // subspace holds the solution)
            const bool ok = std::fabs(g[j + 1]) <= tol;
            if (ok || hn == 0.0) {
                *converged = ok || hn == 0.0;
                j++;
                break;
            }
        }
        // x += V y, H y = g (upper triangular, j columns)
        double y[GM_RESTART];
        for (int i = j - 1; i >= 0; i--) {
            double s = g[i];
            for (int k = i + 1; k < j; k++) s -= H[i][k] * y[k];
            y[i] = s / H[i][i];
        }
        (void)hipMemcpyAsync(ydev, y, sizeof(double) * j, hipMemcpyHostToDevice, c->stream);
        hipLaunchKernelGGL(k_gm_xupd, dim3(nbv), dim3(256), 0, c->stream, n3, V, ld, j, ydev, x);
        (void)hipStreamSynchronize(c->stream);
    }
    return it;
}

}  // namespace sw2

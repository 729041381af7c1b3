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
// Each kernel ends in a single-pass reduction (block partials, the last block
// to finish reduces them in block order and updates the device scalars), so a
// BiCGSTAB iteration is 3 launches with no host round trip; the host reads
// the convergence flag every opts.check_every iterations and kernels launched
// past convergence return at entry.  The same kernels run in FP64 or FP32
// (matrix, preconditioner and vectors), with dot products accumulated in FP64.
//
// HBM layout: 3x3 diagonal blocks diag[9N] and off-diagonal blocks offd[9*nnz]
// (block rows in the reference's insertion order), col[nnz], row_ptr[N+1],
// vectors [3N] node-major.  SpMV reads 72 B (FP64) per block.
#include "sw2_ctx.h"

namespace sw2 {

constexpr double SOLV_TOL = SMALL;  // src/include/define.h:177

// ------------------------------------------------------------- Jacobi scaling
// spv_jacobi 3-DOF branch (src/solver/spv_jacobi.c:222-316)
__global__ void k_scale_rows(int N, const double *diag, const double *offd, const int *row_ptr, const int *sell_ptr,
                             double *sv) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double *D = diag + 9 * (size_t)n;
    double s[3];
    for (int j = 0; j < 3; j++) {
        double m = 0.0;
        for (int k = 0; k < 3; k++) if (fabs(D[3 * j + k]) > m) m = fabs(D[3 * j + k]);
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
    for (int j = 0; j < 3; j++) sv[3 * n + j] = s[j] < SOLV_TOL ? small_factor : 1.0 / sqrt(fabs(s[j]));
}

__global__ void k_scale_apply(int N, double *diag, double *offd, const int *row_ptr, const int *sell_ptr,
                              const int *sell_col, const double *sv, double *R, double *sol) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double rf[3] = {sv[3 * n], sv[3 * n + 1], sv[3 * n + 2]};
    for (int j = 0; j < 3; j++) {
        R[3 * n + j] *= rf[j];
        sol[3 * n + j] /= rf[j];
    }
    double *D = diag + 9 * (size_t)n;
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) D[3 * j + k] *= rf[j] * rf[k];
    const size_t q0 = 64 * (size_t)sell_ptr[n >> 6] + (n & 63);
    const int len = row_ptr[n + 1] - row_ptr[n];
    for (int b = 0; b < len; b++) {
        const size_t q = q0 + 64 * (size_t)b;
        const double *sc = sv + 3 * (size_t)sell_col[q];
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) offd[sell_vidx(q, 3 * j + k)] *= rf[j] * sc[k];
    }
}

// 3x3 block inverse by Gauss-Jordan with partial pivoting (block-Jacobi set-up)
__global__ void k_bj_setup(int N, const double *diag, double *dinv, float *dinv32) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double a[3][6];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            a[r][c] = diag[9 * (size_t)n + 3 * r + c];
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
        for (int c = 0; c < 3; c++) {
            dinv[9 * (size_t)n + 3 * r + c] = a[r][3 + c];
            if (dinv32) dinv32[9 * (size_t)n + 3 * r + c] = (float)a[r][3 + c];
        }
}

__global__ void k_to_f32(size_t n, const double *a, float *b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

void launch_setup(Ctx *c) {
    int nb = nblocks_for(c->N, 256);
    hipLaunchKernelGGL(k_scale_rows, dim3(nb), dim3(256), 0, c->stream, c->N, c->diag, c->offd, c->row_ptr,
                       c->sell_ptr, c->scale);
    hipLaunchKernelGGL(k_scale_apply, dim3(nb), dim3(256), 0, c->stream, c->N, c->diag, c->offd, c->row_ptr,
                       c->sell_ptr, c->sell_col, c->scale, c->R, c->sol);
    bool f32 = c->opts.precision == 1;
    hipLaunchKernelGGL(k_bj_setup, dim3(nb), dim3(256), 0, c->stream, c->N, c->diag, c->dinv, f32 ? c->dinv32 : nullptr);
    if (f32) {
        size_t n9 = 9 * (size_t)c->N, m9 = 9 * c->nsell;
        hipLaunchKernelGGL(k_to_f32, dim3(nblocks_for((int)n9, 256)), dim3(256), 0, c->stream, n9, c->diag, c->diag32);
        if (m9)
            hipLaunchKernelGGL(k_to_f32, dim3(nblocks_for((int)m9, 256)), dim3(256), 0, c->stream, m9, c->offd,
                               c->offd32);
    }
}

// ------------------------------------------------------------ reductions
// Block-level reduction of K values (op: 0 sum, 1 max) followed by the
// last-block pass.  Returns true in the last block; `out` valid in thread 0.
template <int K>
__device__ __forceinline__ bool reduce_last(double (&v)[K], const int (&op)[K], double *part, unsigned *count,
                                            double (&out)[K]) {
    __shared__ double sh[K][256 / 64];
    __shared__ bool last;
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
            part[K * blockIdx.x + k] = x;
        }
        __threadfence();
        unsigned prev = atomicAdd(count, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = op[k] ? 0.0 : 0.0;
    // thread-strided partial sums in fixed order, then a fixed tree: deterministic
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < K; k++) {
            double x = __builtin_nontemporal_load(part + K * b + k);
            acc[k] = op[k] ? fmax(acc[k], x) : acc[k] + x;
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
        *count = 0;
    }
    return true;
}

// ------------------------------------------------------------ SpMV + M^-1
// y = M^-1 (A x) for block row n; A = diag + offd in the reference's
// accumulation order (solv_amult.c:99-123: diagonal block, then the row's
// blocks in insertion order summed into a separate partial).
// Off-diagonal blocks come from the SELL-64 slice of this row's wavefront:
// every lane runs the slice's K block slots (zero padding past the row's
// length), and entry e of slot j is one coalesced 64-value read per wave.
template <typename T, typename XF>
__device__ __forceinline__ void amult_prec_row(int n, const T *diag, const T *offd, const int *sell_ptr,
                                               const int *sell_col, const T *dinv, XF xf, T (&y)[3]) {
    const T *D = diag + 9 * (size_t)n;
    T x0 = xf(3 * n), x1 = xf(3 * n + 1), x2 = xf(3 * n + 2);
    T a0 = D[0] * x0 + D[1] * x1 + D[2] * x2;
    T a1 = D[3] * x0 + D[4] * x1 + D[5] * x2;
    T a2 = D[6] * x0 + D[7] * x1 + D[8] * x2;
    T p0 = 0, p1 = 0, p2 = 0;
    const int s = n >> 6, lane = n & 63;
    const int k0 = sell_ptr[s], k1 = sell_ptr[s + 1];
    for (int b = k0; b < k1; b++) {
        const size_t q = 64 * (size_t)b + lane;
        const T *a = offd + 9 * 64 * (size_t)b + lane;
        int cc = 3 * sell_col[q];
        T y0 = xf(cc), y1 = xf(cc + 1), y2 = xf(cc + 2);
        p0 += a[0] * y0 + a[64] * y1 + a[128] * y2;
        p1 += a[192] * y0 + a[256] * y1 + a[320] * y2;
        p2 += a[384] * y0 + a[448] * y1 + a[512] * y2;
    }
    a0 += p0; a1 += p1; a2 += p2;
    const T *M = dinv + 9 * (size_t)n;
    y[0] = M[0] * a0 + M[1] * a1 + M[2] * a2;
    y[1] = M[3] * a0 + M[4] * a1 + M[5] * a2;
    y[2] = M[6] * a0 + M[7] * a1 + M[8] * a2;
}

template <typename T>
struct VecX {
    const T *x;
    __device__ T operator()(int i) const { return x[i]; }
};
template <typename T>
struct VecS {  // s = r + (-alpha) Ap, the reference's copy + daxpy
    const T *r, *ap;
    T nalpha;
    __device__ T operator()(int i) const { return r[i] + nalpha * ap[i]; }
};

template <typename T>
struct KVecs {
    const T *diag, *offd, *dinv;
    T *b, *r, *p, *Ap, *s, *As, *q, *u;
};

// prologue part 1: b' = M^-1 b, u = b', ||b'||inf  (solv_bcgstab.c:117-137)
template <typename T>
__global__ void __launch_bounds__(256) k_bcg_p1(int N, KVecs<T> V, const double *rhs, double *part, unsigned *count,
                                                KScal *sc) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (n < N) {
        const T *M = V.dinv + 9 * (size_t)n;
        T b0 = (T)rhs[3 * n], b1 = (T)rhs[3 * n + 1], b2 = (T)rhs[3 * n + 2];
        T y[3];
        y[0] = M[0] * b0 + M[1] * b1 + M[2] * b2;
        y[1] = M[3] * b0 + M[4] * b1 + M[5] * b2;
        y[2] = M[6] * b0 + M[7] * b1 + M[8] * b2;
        for (int j = 0; j < 3; j++) {
            V.b[3 * n + j] = y[j];
            V.u[3 * n + j] = y[j];
            v[0] = fmax(v[0], fabs((double)y[j]));
        }
    }
    const int op[1] = {1};
    double out[1];
    if (reduce_last<1>(v, op, part, count, out) && threadIdx.x == 0) sc->bnorm = out[0];
}

// prologue part 2: r = b' - M^-1 A u, q = r, p = r, rho = (r,q), ||r||inf  (:138-188)
template <typename T>
__global__ void __launch_bounds__(256) k_bcg_p2(int N, KVecs<T> V, const int *sell_ptr, const int *sell_col, double *part,
                                                unsigned *count, KScal *sc) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (n < N) {
        T y[3];
        amult_prec_row<T>(n, V.diag, V.offd, sell_ptr, sell_col, V.dinv, VecX<T>{V.u}, y);
        for (int j = 0; j < 3; j++) {
            T r = V.b[3 * n + j] + (T)(-1.0) * y[j];
            V.r[3 * n + j] = r;
            V.q[3 * n + j] = r;
            V.p[3 * n + j] = r;
            v[0] = fmax(v[0], fabs((double)r));
            v[1] += (double)r * (double)r;
        }
    }
    const int op[2] = {1, 0};
    double out[2];
    if (reduce_last<2>(v, op, part, count, out) && threadIdx.x == 0) {
        sc->rnorm = out[0];
        sc->rho = out[1];
        sc->rhop = 1.0;
        sc->alpha = 1.0;
        sc->omega = 1.0;
        sc->tol = sc->atol + sc->bnorm * sc->rtol;
        sc->it = 0;
        sc->done = !(sc->rnorm > sc->tol && 0 < sc->max_it);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bcg_k1(int N, KVecs<T> V, const int *sell_ptr, const int *sell_col, double *part,
                                                unsigned *count, KScal *sc) {
    if (sc->done) return;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (n < N) {
        T y[3];
        amult_prec_row<T>(n, V.diag, V.offd, sell_ptr, sell_col, V.dinv, VecX<T>{V.p}, y);
        for (int j = 0; j < 3; j++) {
            V.Ap[3 * n + j] = y[j];
            v[0] += (double)V.q[3 * n + j] * (double)y[j];
        }
    }
    const int op[1] = {0};
    double out[1];
    if (reduce_last<1>(v, op, part, count, out) && threadIdx.x == 0) {
        sc->gamma = out[0];
        sc->alpha = sc->rho / sc->gamma;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bcg_k2(int N, KVecs<T> V, const int *sell_ptr, const int *sell_col, double *part,
                                                unsigned *count, KScal *sc) {
    if (sc->done) return;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    const T nalpha = (T)(-sc->alpha);
    double v[3] = {0.0, 0.0, 0.0};
    if (n < N) {
        VecS<T> sf{V.r, V.Ap, nalpha};
        T y[3];
        amult_prec_row<T>(n, V.diag, V.offd, sell_ptr, sell_col, V.dinv, sf, y);
        for (int j = 0; j < 3; j++) {
            T s = sf(3 * n + j);
            V.s[3 * n + j] = s;
            V.As[3 * n + j] = y[j];
            v[0] += (double)y[j] * (double)s;
            v[1] += (double)y[j] * (double)y[j];
            v[2] += (double)y[j] * (double)V.q[3 * n + j];
        }
    }
    const int op[3] = {0, 0, 0};
    double out[3];
    if (reduce_last<3>(v, op, part, count, out) && threadIdx.x == 0) {
        double omega = out[0] / out[1];
        sc->omega = omega;
        sc->rhop = sc->rho;
        sc->rho = -1.0 * omega * out[2];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bcg_k3(int N, KVecs<T> V, double *part, unsigned *count, KScal *sc) {
    if (sc->done) return;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    const double alpha = sc->alpha, omega = sc->omega;
    const double beta = (sc->rho * alpha) / (sc->rhop * omega);
    const T ta = (T)alpha, to = (T)omega, tno = (T)(-omega), tb = (T)beta;
    double v[1] = {0.0};
    if (n < N) {
        for (int j = 0; j < 3; j++) {
            int i = 3 * n + j;
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
    double out[1];
    if (reduce_last<1>(v, op, part, count, out) && threadIdx.x == 0) {
        sc->rnorm = out[0];
        sc->it += 1;
        sc->done = !(sc->rnorm > sc->tol && sc->it < sc->max_it);
    }
}

// solv_return_old (src/solver/solv_return_old.c:23-30): sol = (u + u0) * scale
template <typename T>
__global__ void k_bcg_epi(int N3, const T *u, const double *scale, double *sol) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N3) return;
    double x = (double)u[i];
    x += sol[i];
    x *= scale[i];
    sol[i] = x;
}

template <typename T>
static KVecs<T> vecs(Ctx *c);
template <>
KVecs<double> vecs<double>(Ctx *c) {
    return KVecs<double>{c->diag, c->offd, c->dinv, c->kb, c->kr, c->kp, c->kAp, c->ks_, c->kAs, c->kq, c->ku};
}
template <>
KVecs<float> vecs<float>(Ctx *c) {
    return KVecs<float>{c->diag32, c->offd32, c->dinv32, c->fb, c->fr, c->fp, c->fAp, c->fs, c->fAs, c->fq, c->fu};
}

template <typename T>
static void prologue(Ctx *c) {
    int nb = nblocks_for(c->N, 256);
    KVecs<T> V = vecs<T>(c);
    hipLaunchKernelGGL(k_bcg_p1<T>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->R, c->red.part, c->red.count + 0,
                       c->ksc);
    hipLaunchKernelGGL(k_bcg_p2<T>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr, c->sell_col, c->red.part,
                       c->red.count + 1, c->ksc);
}

template <typename T>
static void iters(Ctx *c, int k) {
    int nb = nblocks_for(c->N, 256);
    KVecs<T> V = vecs<T>(c);
    for (int i = 0; i < k; i++) {
        hipLaunchKernelGGL(k_bcg_k1<T>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr, c->sell_col, c->red.part,
                           c->red.count + 2, c->ksc);
        hipLaunchKernelGGL(k_bcg_k2<T>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->sell_ptr, c->sell_col, c->red.part,
                           c->red.count + 3, c->ksc);
        hipLaunchKernelGGL(k_bcg_k3<T>, dim3(nb), dim3(256), 0, c->stream, c->N, V, c->red.part, c->red.count + 4,
                           c->ksc);
    }
}

void launch_k1_only(Ctx *c) {
    int nb = nblocks_for(c->N, 256);
    if (c->opts.precision == 1)
        hipLaunchKernelGGL(k_bcg_k1<float>, dim3(nb), dim3(256), 0, c->stream, c->N, vecs<float>(c), c->sell_ptr,
                           c->sell_col, c->red.part, c->red.count + 2, c->ksc);
    else
        hipLaunchKernelGGL(k_bcg_k1<double>, dim3(nb), dim3(256), 0, c->stream, c->N, vecs<double>(c), c->sell_ptr,
                           c->sell_col, c->red.part, c->red.count + 2, c->ksc);
}

void launch_krylov_prologue(Ctx *c) {
    if (c->opts.precision == 1) prologue<float>(c);
    else prologue<double>(c);
}

void launch_krylov_iters(Ctx *c, int k) {
    if (c->opts.precision == 1) iters<float>(c, k);
    else iters<double>(c, k);
}

void launch_krylov_epilogue(Ctx *c) {
    int n3 = 3 * c->N;
    if (c->opts.precision == 1)
        hipLaunchKernelGGL(k_bcg_epi<float>, dim3(nblocks_for(n3, 256)), dim3(256), 0, c->stream, n3, c->fu, c->scale,
                           c->sol);
    else
        hipLaunchKernelGGL(k_bcg_epi<double>, dim3(nblocks_for(n3, 256)), dim3(256), 0, c->stream, n3, c->ku, c->scale,
                           c->sol);
}

}  // namespace sw2

// Profile-LU block preconditioner (OP BLK): the reference's solv_blk_set /
// solv_blk_load_profile / solv_blk_factor_profile / solv_blk_solve
// (src/solver/solv_blk_*.c, node blocks from node_order.c:64-104) on the
// device.  Nodes are cut into `nblock` consecutive ranges; each block's
// diagonal submatrix (Jacobi-scaled, as the reference factors it) is
// factored exactly, without pivoting and with the reference's SMALL guard
// on the pivot, in a dense band that covers the block's profile (so the
// factors are those of the skyline LU; only the rounding order of the
// updates differs).
//
// One workgroup factors one block (right-looking band update, a barrier per
// pivot); one wavefront solves one block (forward unit-lower, backward
// upper), keeping the block's right-hand side in LDS.  This is the robust
// preconditioner for steps with large Courant numbers, where 3x3 block
// Jacobi stalls; its cost grows with block size times bandwidth squared.
#include <algorithm>
#include <string>

#include "sw2_ctx.h"

namespace sw2 {

// Band of block b: local DOF i in [0, nb), entry (i, j) with |i - j| <= w at
// band[off + i * (2w + 1) + (j - i + w)].
struct BandView {
    const int *n0;         // [nblk + 1] first node of each block
    const int *w;          // [nblk] half bandwidth (DOFs)
    const size_t *off;     // [nblk] band offsets (doubles)
    double *band;
};

__global__ void k_band_load(int N, const int *node_blk, BandView B, const double *diag, const double *offd,
                            const int *row_ptr, const int *sell_ptr, const int *sell_col) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int b = node_blk[a], n0 = B.n0[b], n1 = B.n0[b + 1], w = B.w[b], ld = 2 * w + 1;
    double *A = B.band + B.off[b];
    const int ia = 3 * (a - n0);
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) {
            const int i = ia + r, j = ia + cc;
            A[(size_t)i * ld + (j - i + w)] = diag[didx9(a, 3 * r + cc)];
        }
    const size_t q0 = 64 * (size_t)sell_ptr[a >> 6] + (a & 63);
    const int len = row_ptr[a + 1] - row_ptr[a];
    for (int s = 0; s < len; s++) {
        const size_t q = q0 + 64 * (size_t)s;
        const int c = sell_col[q];
        if (c < n0 || c >= n1) continue;
        const int jc = 3 * (c - n0);
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                const int i = ia + r, j = jc + cc;
                A[(size_t)i * ld + (j - i + w)] = offd[sell_vidx(q, 3 * r + cc)];
            }
    }
}

// Right-looking band LU of one block per workgroup.  Multipliers are divided
// by the pivot, or by SMALL when |pivot| <= SMALL (solv_blk_load_profile.c,
// factor loop); U keeps the diagonal.
__global__ void __launch_bounds__(256) k_band_factor(BandView B) {
    const int b = blockIdx.x;
    const int n = 3 * (B.n0[b + 1] - B.n0[b]), w = B.w[b], ld = 2 * w + 1;
    double *A = B.band + B.off[b];
    for (int k = 0; k < n; k++) {
        const double piv = A[(size_t)k * ld + w];
        const double den = fabs(piv) > SMALL ? piv : SMALL;
        const int m = min(n - 1, k + w) - k;  // rows (and columns) below/right of the pivot in the band
        for (int t = threadIdx.x; t < m; t += blockDim.x) {
            const int i = k + 1 + t;
            A[(size_t)i * ld + (k - i + w)] /= den;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
            const int i = k + 1 + t / m, j = k + 1 + t % m;
            A[(size_t)i * ld + (j - i + w)] -= A[(size_t)i * ld + (k - i + w)] * A[(size_t)k * ld + (j - k + w)];
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// x <- (LU)^-1 x for one block per 64-thread workgroup; x is a sliced node
// vector.  Skipped once the Krylov solve has converged (done flag).
__global__ void __launch_bounds__(64) k_band_solve(BandView B, double *x, const KScal *sc) {
    extern __shared__ double ys[];
    if (sc && sc->done) return;
    const int b = blockIdx.x, lane = threadIdx.x;
    const int node0 = B.n0[b], n = 3 * (B.n0[b + 1] - node0), w = B.w[b], ld = 2 * w + 1;
    const double *A = B.band + B.off[b];
    for (int i = lane; i < n; i += 64) ys[i] = x[vidx3(node0 + i / 3, i % 3)];
    __syncthreads();
    for (int i = 0; i < n; i++) {  // forward, unit lower
        double s = 0.0;
        for (int j = max(0, i - w) + lane; j < i; j += 64) s += A[(size_t)i * ld + (j - i + w)] * ys[j];
        s = wave_sum(s);
        if (lane == 0) ys[i] -= s;
        __syncthreads();
    }
    for (int i = n - 1; i >= 0; i--) {  // backward, upper with diagonal
        double s = 0.0;
        const int jend = min(n - 1, i + w);
        for (int j = i + 1 + lane; j <= jend; j += 64) s += A[(size_t)i * ld + (j - i + w)] * ys[j];
        s = wave_sum(s);
        if (lane == 0) ys[i] = (ys[i] - s) / A[(size_t)i * ld + w];
        __syncthreads();
    }
    for (int i = lane; i < n; i += 64) x[vidx3(node0 + i / 3, i % 3)] = ys[i];
}

static BandView view(Ctx *c) { return BandView{c->blk_n0, c->blk_w, c->blk_off, c->band}; }

void band_free(Ctx *c) {
    tband_free(c);
    c->band_tiled = 0;
    for (void *p : {(void *)c->blk_n0, (void *)c->blk_w, (void *)c->blk_off, (void *)c->node_blk, (void *)c->band})
        if (p) (void)hipFree(p);
    c->blk_n0 = nullptr; c->blk_w = nullptr; c->blk_off = nullptr; c->node_blk = nullptr; c->band = nullptr;
    c->band_nblock = 0;
    c->band_nb = 0;
    c->band_size = 0;
    c->band_max_n = 0;
}

// Builds the block partition and band offsets for `nblock` blocks
// (node_order.c:64-104 serial partition: N/nblock consecutive nodes per
// block, the last block taking the rest).  An ensemble partitions each model's
// node range so, as that model alone would be (the gap nodes between models
// belong to no block), and always takes the tiled band LU, whose launches
// cover every model's bands.  Returns an error text or "".
std::string band_structure(Ctx *c, int nblock) {
    band_free(c);
    const int N = c->N;
    if (nblock < 1) return "nblock must be >= 1";
    nblock = std::min(nblock, N);
    std::vector<int> r0, r1;   // node ranges to partition
    if (c->ninst) {
        for (int i = 0; i < c->ninst; i++) {
            r0.push_back(c->h_inst_node0[i]);
            r1.push_back(c->h_inst_node0[i] + c->h_inst_N[i]);
        }
    } else {
        r0.push_back(0);
        r1.push_back(N);
    }
    std::vector<int> s0, s1, binst;   // block b: nodes [s0[b], s1[b]) of model binst[b]
    for (size_t i = 0; i < r0.size(); i++) {
        const int ni = r1[i] - r0[i], nbi = std::min(nblock, ni);
        const int inc = std::max(1, ni / nbi);
        int last = r0[i];
        for (int b = 0; b < nbi; b++) {
            s0.push_back(last);
            last = (b == nbi - 1) ? r1[i] : last + inc;
            s1.push_back(last);
            binst.push_back((int)i);
        }
    }
    const int nbk = (int)s0.size();
    std::vector<int> w(nbk), node_blk(N, -1);
    std::vector<size_t> off(nbk + 1, 0);
    int max_n = 0;
    for (int b = 0; b < nbk; b++) {
        int maxd = 0;
        for (int a = s0[b]; a < s1[b]; a++) {
            node_blk[a] = b;
            for (int p = c->h_row_ptr[a]; p < c->h_row_ptr[a + 1]; p++) {
                const int cn = c->h_col[p];
                if (cn >= s0[b] && cn < s1[b]) maxd = std::max(maxd, std::abs(cn - a));
            }
        }
        w[b] = 3 * maxd + 2;
        const int nb = 3 * (s1[b] - s0[b]);
        max_n = std::max(max_n, nb);
        off[b + 1] = off[b] + (size_t)nb * (2 * w[b] + 1);
    }
    // blocks beyond the one-workgroup band LU (its solve keeps a block's
    // right-hand side in LDS), ensembles, or when asked for: the tiled band LU
    if (c->ninst || max_n > BAND_MAX_DOF || off[nbk] * sizeof(double) > BAND_MAX_BYTES || c->opts.lu_tiled) {
        if (hipMalloc((void **)&c->node_blk, sizeof(int) * N) != hipSuccess ||
            hipMemcpy(c->node_blk, node_blk.data(), sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess) {
            band_free(c);
            return "hipMalloc failed for the profile-LU blocks";
        }
        std::string err = tband_structure(c, s0, s1, binst, w);
        if (!err.empty()) {
            band_free(c);
            return err;
        }
        c->band_tiled = 1;
        c->band_nblock = nblock;
        c->band_nb = nbk;
        c->band_max_n = max_n;
        return "";
    }
    // one model: consecutive blocks, n0[b + 1] = the end of block b
    std::vector<int> n0(s0);
    n0.push_back(N);
    nblock = nbk;
    auto up = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (hipMalloc(dst, bytes ? bytes : 8) != hipSuccess) return false;
        return bytes == 0 || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up((void **)&c->blk_n0, n0.data(), sizeof(int) * n0.size()) ||
        !up((void **)&c->blk_w, w.data(), sizeof(int) * w.size()) ||
        !up((void **)&c->blk_off, off.data(), sizeof(size_t) * off.size()) ||
        !up((void **)&c->node_blk, node_blk.data(), sizeof(int) * node_blk.size()) ||
        hipMalloc((void **)&c->band, sizeof(double) * (off[nblock] ? off[nblock] : 1)) != hipSuccess) {
        band_free(c);
        return "hipMalloc failed for the profile-LU band";
    }
    c->band_nblock = nblock;
    c->band_nb = nblock;
    c->band_size = off[nblock];
    c->band_max_n = max_n;
    return "";
}

// Loads the scaled system into the bands and factors every block.
void launch_band_setup(Ctx *c) {
    if (c->band_tiled) {
        launch_tband_setup(c);
        return;
    }
    launch_zero(c, c->band, c->band_size);
    hipLaunchKernelGGL(k_band_load, dim3(nblocks_for(c->N, 256)), dim3(256), 0, c->stream, c->N, c->node_blk,
                       view(c), c->diag, c->offd, c->row_ptr, c->sell_ptr, c->sell_col);
    hipLaunchKernelGGL(k_band_factor, dim3(c->band_nb), dim3(256), 0, c->stream, view(c));
}

void launch_band_solve(Ctx *c, double *x, const KScal *sc) {
    if (c->band_tiled) {
        launch_tband_solve(c, x, sc);
        return;
    }
    hipLaunchKernelGGL(k_band_solve, dim3(c->band_nb), dim3(64), sizeof(double) * c->band_max_n, c->stream, view(c), x,
                       sc);
}

}  // namespace sw2

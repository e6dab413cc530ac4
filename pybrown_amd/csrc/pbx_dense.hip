// Dense FP64 linear algebra on gfx950 for the 3N x 3N diffusion and
// resistance matrices: blocked Cholesky (potrf), triangular inverse and the
// SPD inverse built from them.
//
// Matrices are row-major, order np (a multiple of 64) with leading dimension
// ld; only the lower triangle is read.  Work is cut into 64 x 64 tiles; every
// tile product runs on the FP64 matrix cores (v_mfma_f64_16x16x4_f64): a
// 256-thread workgroup owns one 64 x 64 output tile, each of its four waves a
// 32 x 32 quadrant held as 2 x 2 MFMA accumulators (pbx_tile.h).
//
// Cholesky (variants 1 and 2), right-looking with look-ahead fused into the
// update kernel:
//   potrf_diag(0)
//   for k = 0 .. T-2:
//     trsm(k):   L_ik = A_ik * inv(L_kk)^T           (one workgroup per tile row)
//     syrk(k):   A_ij -= L_ik L_jk^T  for k < j <= i   (one workgroup per tile)
//                and the workgroup owning tile (k+1,k+1) factors it in LDS
// Variant 3 is the single-launch tile graph of pbx_chol_dag.hip.
// Variant 4 (default) folds trsm(k+1) into syrk(k): the workgroups of tile
// column k+1 keep their updated tile in registers, wait (bounded) for the
// look-ahead workgroup to publish inv(L_{k+1,k+1}), and solve in place, so
// each panel costs one launch; its diagonal tiles use the wave-specialised
// factorisation (tile_factor_ws).
// inv(L_kk) is produced with each diagonal factorisation and kept in `work`
// (T tiles of 64 x 64), where the SPD inverse reuses it.
#include "pbx_common.h"
#include "pbx_tile.h"

#define TRY_RC(x)            \
    do {                     \
        int rc_ = (x);       \
        if (rc_) return rc_; \
    } while (0)

using namespace pbx_tile;

namespace {

// Column-by-column factorisation of the LDS tile with inv(L) formed row by
// row (two workgroup barriers per column).  Kept as variant 1: the simplest
// statement of the tile step, used to cross-check the panel version.
__device__ void tile_factor_columns(double* sL, double* sW, double* sd, int col0, int* info) {
    const int tid = threadIdx.x;
    for (int e = tid; e < T64 * SP; e += blockDim.x) sW[e] = 0.0;
    __syncthreads();
    for (int j = 0; j < T64; ++j) {
        const double djj = sL[j * SP + j];
        const double dj = sqrt(djj);
        if (tid < 64) {
            int i = tid;
            if (i > j) sL[i * SP + j] = sL[i * SP + j] / dj;
            if (i == j) {
                sd[j] = dj;
                if (!(djj > 0.0)) atomicCAS(info, 0, col0 + j + 1);
            }
        } else if (tid < 128) {
            int c = tid - 64;
            if (c <= j) {
                double s0 = (c == j) ? 1.0 : 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                int k = c;
                for (; k + 3 < j; k += 4) {
                    s0 -= sL[j * SP + k] * sW[k * SP + c];
                    s1 -= sL[j * SP + k + 1] * sW[(k + 1) * SP + c];
                    s2 -= sL[j * SP + k + 2] * sW[(k + 2) * SP + c];
                    s3 -= sL[j * SP + k + 3] * sW[(k + 3) * SP + c];
                }
                for (; k < j; ++k) s0 -= sL[j * SP + k] * sW[k * SP + c];
                sW[j * SP + c] = ((s0 + s1) + (s2 + s3)) / dj;
            }
        }
        __syncthreads();
        for (int e = tid; e < T64 * T64; e += blockDim.x) {
            int i = e >> 6, c = e & 63;
            if (c > j && c <= i) sL[i * SP + c] -= sL[i * SP + j] * sL[c * SP + j];
        }
        __syncthreads();
    }
    if (tid < 64) sL[tid * SP + tid] = sd[tid];
    __syncthreads();
}

template <int V>
__device__ __forceinline__ void factor_tile(double* sL, double* sW, double* sd, int col0, int* info) {
    if (V == 5)
        tile_factor_ws(sL, sW, sd, col0, info);
    else if (V == 2)
        tile_factor(sL, sW, sd, col0, info);
    else
        tile_factor_columns(sL, sW, sd, col0, info);
}

// Write the factored diagonal tile (lower incl. diagonal) and its inverse.
__device__ void store_diag(const double* sL, const double* sW, double* A, int ld, int k, double* work) {
    double* Akk = A + (size_t)k * T64 * ld + (size_t)k * T64;
    double* Wk = work + (size_t)k * T64 * T64;
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        int r = e >> 6, c = e & 63;
        if (c <= r) Akk[(size_t)r * ld + c] = sL[r * SP + c];
        Wk[e] = sW[r * SP + c];
    }
}

// Look-ahead state words (variant 4), in the work area after the diagonal
// inverses: [LA_FLAG] = index of the last diagonal tile published,
// [LA_ERR] = a dependency wait ran out of budget.  (Words 2 and 3 of the
// single-launch graph's state block, which that variant does not use.)
constexpr int LA_FLAG = 2, LA_ERR = 3;

template <int V>
__global__ __launch_bounds__(256) void potrf_diag_kernel(double* A, int ld, int k, double* work, int* info,
                                                         int* la) {
    __shared__ double sL[T64 * SP];
    __shared__ double sW[T64 * SP];
    __shared__ double sd[T64];
    if (la && threadIdx.x == 0) {
        la[LA_FLAG] = 0;
        la[LA_ERR] = 0;
    }
    stage<false>(sL, A + (size_t)k * T64 * ld + (size_t)k * T64, ld);
    __syncthreads();
    factor_tile<V>(sL, sW, sd, k * T64, info);
    store_diag(sL, sW, A, ld, k, work);
}

// L_ik = A_ik * W_k^T for tile rows i = k+1+blockIdx.x
__global__ __launch_bounds__(256) void trsm_kernel(double* A, int ld, int k, const double* work) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    const int i = k + 1 + blockIdx.x;
    double* Aik = A + (size_t)i * T64 * ld + (size_t)k * T64;
    stage<true>(sA, Aik, ld);                             // op(A) = A_ik
    stage<true>(sB, work + (size_t)k * T64 * T64, T64);   // op(B)(kk,c) = W[c][kk]
    __syncthreads();
    v4d acc[2][2];
    acc_zero(acc);
    mma64(sA, sB, acc);
    acc_for_each(acc, [&](int r, int c, double v) { Aik[(size_t)r * ld + c] = v; });
}

__device__ __forceinline__ void tri_decode(int t, int& ii, int& jj) {
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    ii = r;
    jj = t - r * (r + 1) / 2;
}

// A_ij -= L_ik L_jk^T over the trailing lower tiles; the workgroup owning
// (k+1, k+1) (blockIdx 0) then factors that tile.
template <int V>
__global__ __launch_bounds__(256, 2) void syrk_kernel(double* A, int ld, int k, double* work, int* info) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double sd[T64];
    int ii, jj;
    tri_decode(blockIdx.x, ii, jj);
    const int i = k + 1 + ii, j = k + 1 + jj;
    double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
    double cv[16];
    acc_load(Aij, ld, cv);  // issued early: overlaps the staging below
    stage<true>(sA, A + (size_t)i * T64 * ld + (size_t)k * T64, ld);
    if (i != j) stage<true>(sB, A + (size_t)j * T64 * ld + (size_t)k * T64, ld);
    __syncthreads();
    v4d acc[2][2];
    acc_zero(acc);
    mma64(sA, i == j ? sA : sB, acc);
    if (blockIdx.x != 0) {
        int q = 0;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v, ++q) {
                    const int r = acc_row(x, v), c = acc_col(y);
                    if (i != j || c <= r) Aij[(size_t)r * ld + c] = cv[q] - acc[x][y][v];
                }
        return;
    }
    // look-ahead: update, factor and invert tile (k+1, k+1) in LDS
    __syncthreads();  // all waves done reading sA
    double* sL = sA;
    double* sW = sB;
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v, ++q) sL[acc_row(x, v) * SP + acc_col(y)] = cv[q] - acc[x][y][v];
    __syncthreads();
    factor_tile<V>(sL, sW, sd, (k + 1) * T64, info);
    store_diag(sL, sW, A, ld, k + 1, work);
}

// Wait until *flag >= target (thread 0 polls with device-scope acquire),
// bounded; false if the budget ran out or another wait already failed.
__device__ bool wait_ge_agent(const int* flag, int target, int* err) {
    __shared__ int ok;
    if (threadIdx.x == 0) {
        int good = 1;
        long long spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > (1 << 22)) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// syrk(k) with trsm(k+1) folded in (variant 4).  Workgroup 0 updates and
// factors tile (k+1, k+1) and publishes it; workgroups of tiles (i, k+1),
// i > k+1, wait for that and write L_{i,k+1} = C W_{k+1}^T instead of C.
// Only T-k-2 workgroups ever wait, and only on workgroup 0, which waits on
// nothing, so every wait ends.
template <int V>
__global__ __launch_bounds__(256, 2) void syrk_trsm_kernel(double* A, int ld, int k, double* work, int* info,
                                                           int* la) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double sd[T64];
    int ii, jj;
    tri_decode(blockIdx.x, ii, jj);
    const int i = k + 1 + ii, j = k + 1 + jj;
    double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
    double cv[16];
    acc_load(Aij, ld, cv);
    stage<true>(sA, A + (size_t)i * T64 * ld + (size_t)k * T64, ld);
    if (i != j) stage<true>(sB, A + (size_t)j * T64 * ld + (size_t)k * T64, ld);
    __syncthreads();
    v4d acc[2][2];
    acc_zero(acc);
    mma64(sA, i == j ? sA : sB, acc);
    if (blockIdx.x == 0) {
        __syncthreads();
        double* sL = sA;
        double* sW = sB;
        int q = 0;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v, ++q) sL[acc_row(x, v) * SP + acc_col(y)] = cv[q] - acc[x][y][v];
        __syncthreads();
        factor_tile<V>(sL, sW, sd, (k + 1) * T64, info);
        store_diag(sL, sW, A, ld, k + 1, work);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(la + LA_FLAG, k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    if (jj != 0) {
        int q = 0;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v, ++q) {
                    const int r = acc_row(x, v), c = acc_col(y);
                    if (i != j || c <= r) Aij[(size_t)r * ld + c] = cv[q] - acc[x][y][v];
                }
        return;
    }
    // tile (i, k+1): op(A)(r, kk) = C[r][kk] in sA, then C W^T
    __syncthreads();
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v, ++q) sA[acc_col(y) * SP + acc_row(x, v)] = cv[q] - acc[x][y][v];
    if (!wait_ge_agent(la + LA_FLAG, k + 1, la + LA_ERR)) return;
    stage<true>(sB, work + (size_t)(k + 1) * T64 * T64, T64);  // op(B)(kk, c) = W[c][kk]
    __syncthreads();
    v4d out[2][2];
    acc_zero(out);
    mma64(sA, sB, out);
    acc_for_each(out, [&](int r, int c, double v) { Aij[(size_t)r * ld + c] = v; });
}

// An exhausted look-ahead wait becomes the factorisation status word.
__global__ void la_finish_kernel(const int* la, int* info) {
    if (threadIdx.x == 0 && la[LA_ERR] != 0) atomicCAS(info, 0, PBX_INFO_DAG_WAIT);
}

// ---------------------------------------------------------------------------
// Triangular inverse X = L^{-1} (lower) from a factored A and its diagonal
// inverses W_i: X_ii = W_i; X_ij = -W_i * sum_{j<=m<i} L_im X_mj  (i > j).
// One launch per tile row i, one workgroup per j < i.

__global__ void copy_diag_inv_kernel(double* X, int ldx, const double* work) {
    const int i = blockIdx.x;
    double* Xii = X + (size_t)i * T64 * ldx + (size_t)i * T64;
    const double* Wi = work + (size_t)i * T64 * T64;
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        int r = e >> 6, c = e & 63;
        Xii[(size_t)r * ldx + c] = Wi[e];
    }
}

__global__ __launch_bounds__(256) void trtri_row_kernel(const double* A, int ld, double* X, int ldx, int i,
                                                        const double* work) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    const int j = blockIdx.x;  // j < i
    v4d acc[2][2];
    acc_zero(acc);
    for (int m = j; m < i; ++m) {
        stage<true>(sA, A + (size_t)i * T64 * ld + (size_t)m * T64, ld);     // op(A) = L_im
        stage<false>(sB, X + (size_t)m * T64 * ldx + (size_t)j * T64, ldx);  // op(B) = X_mj
        __syncthreads();
        mma64(sA, sB, acc);
        __syncthreads();
    }
    // X_ij = -W_i * S
    acc_for_each(acc, [&](int r, int c, double v) { sB[r * SP + c] = v; });  // s[k=r][c] = S[r][c]
    stage<true>(sA, work + (size_t)i * T64 * T64, T64);                        // op(A) = W_i
    __syncthreads();
    v4d out[2][2];
    acc_zero(out);
    mma64(sA, sB, out);
    double* Xij = X + (size_t)i * T64 * ldx + (size_t)j * T64;
    acc_for_each(out, [&](int r, int c, double v) { Xij[(size_t)r * ldx + c] = -v; });
}

// Y = scale * X^T X, lower tiles computed, both triangles written:
// Y_ij = sum_{m >= i} X_mi^T X_mj for i >= j.
__global__ __launch_bounds__(256) void lauum_kernel(const double* X, int ldx, double* Y, int ld, int T,
                                                    double scale) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    int i, j;
    tri_decode(blockIdx.x, i, j);
    v4d acc[2][2];
    acc_zero(acc);
    for (int m = i; m < T; ++m) {
        stage<false>(sA, X + (size_t)m * T64 * ldx + (size_t)i * T64, ldx);  // op(A)(r,k) = X_mi[k][r]
        stage<false>(sB, X + (size_t)m * T64 * ldx + (size_t)j * T64, ldx);  // op(B)(k,c) = X_mj[k][c]
        __syncthreads();
        mma64(sA, sB, acc);
        __syncthreads();
    }
    double* Yij = Y + (size_t)i * T64 * ld + (size_t)j * T64;
    double* Yji = Y + (size_t)j * T64 * ld + (size_t)i * T64;
    acc_for_each(acc, [&](int r, int c, double v) {
        double s = v * scale;
        if (i != j || c <= r) {
            Yij[(size_t)r * ld + c] = s;
            Yji[(size_t)c * ld + r] = s;
        }
    });
}

// ---------------------------------------------------------------------------
// utilities

// Identity in the pad: rows/cols n..np-1 of an np x np matrix.
__global__ void fill_pad_kernel(double* A, int n, int np, int ld) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)np * np;
    if (t >= total) return;
    int r = (int)(t / np), c = (int)(t % np);
    if (r < n && c < n) return;
    A[(size_t)r * ld + c] = (r == c) ? 1.0 : 0.0;
}

// dst = lower tiles of src (diagonal tiles copied whole)
__global__ void copy_lower_kernel(const double* src, double* dst, int np, int ld) {
    int ti, tj;
    tri_decode(blockIdx.x, ti, tj);
    const double* s = src + (size_t)ti * T64 * ld + (size_t)tj * T64;
    double* d = dst + (size_t)ti * T64 * ld + (size_t)tj * T64;
    v2d v[8];
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int e = threadIdx.x + it * 256;
        v[it] = *(const v2d*)(s + (size_t)(e >> 5) * ld + (e & 31) * 2);
    }
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int e = threadIdx.x + it * 256;
        *(v2d*)(d + (size_t)(e >> 5) * ld + (e & 31) * 2) = v[it];
    }
}

// C = A B^T for an M x K and an N x K matrix, all dims multiples of 64
// (test entry for the MFMA tile path).
__global__ __launch_bounds__(256) void gemm_nt_kernel(const double* A, const double* B, double* C, int M, int N,
                                                      int K) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    const int ti = blockIdx.y, tj = blockIdx.x;
    v4d acc[2][2];
    acc_zero(acc);
    for (int k = 0; k < K; k += T64) {
        stage<true>(sA, A + (size_t)ti * T64 * K + k, K);
        stage<true>(sB, B + (size_t)tj * T64 * K + k, K);
        __syncthreads();
        mma64(sA, sB, acc);
        __syncthreads();
    }
    double* Ct = C + (size_t)ti * T64 * N + (size_t)tj * T64;
    acc_for_each(acc, [&](int r, int c, double v) { Ct[(size_t)r * N + c] = v; });
}

}  // namespace

size_t pbx_dag_state_ints(int np);
int pbx_potrf_dag(double* A, int np, int ld, double* work, int* info_dev, int* state, int grid, hipStream_t s);
bool pbx_persist_fits(int np, int cus);
int pbx_potrf_persist(double* A, int np, int ld, double* work, int* info_dev, int* state, int cus, hipStream_t s);

// compute units of the current device (queried once)
static int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            cus = 1;
    }
    return cus;
}

// work = [T inverse tiles of the diagonal factor][single-launch DAG state]
size_t pbx_dense_work_doubles(int np) { return (size_t)np * T64 + (pbx_dag_state_ints(np) + 1) / 2; }

// Cholesky variant: 1 = launch per panel, column-by-column diagonal tiles;
// 2 = launch per panel, 8-column panel diagonal tiles; 3 = single persistent
// launch over the tile dependency graph (pbx_chol_dag.hip); 4 = one launch
// per panel with the solve folded in and wave-specialised diagonal tiles;
// 5 = persistent launch with tiles owned in registers (pbx_chol_persist.hip;
// falls back to 4 when the tiles do not fit the compute units).
static int g_potrf_variant = 4;
extern "C" int pbx_set_potrf_variant(int v) {
    if (v < 1 || v > 5) return PBX_ERR_ARG;
    g_potrf_variant = v;
    return PBX_OK;
}

int pbx_fill_pad(double* A, int n, int np, int ld, hipStream_t s) {
    if (np == n) return PBX_OK;
    size_t total = (size_t)np * np;
    hipLaunchKernelGGL(fill_pad_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, A, n, np, ld);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_copy_lower(const double* src, double* dst, int np, int ld, hipStream_t s) {
    int T = np / T64;
    hipLaunchKernelGGL(copy_lower_kernel, dim3(T * (T + 1) / 2), dim3(256), 0, s, src, dst, np, ld);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// In-place lower Cholesky of the np x np matrix A; inverses of the diagonal
// factor tiles are left in work (pbx_dense_work_doubles(np) doubles).
// *info_dev must be zero on entry; it receives column+1 of the first
// non-positive pivot.
int pbx_potrf(double* A, int np, int ld, double* work, int* info_dev, hipStream_t s) {
    if (np % T64 != 0 || np <= 0) {
        pbx_set_error(PBX_ERR_ARG, "potrf: order %d is not a positive multiple of 64", np);
        return PBX_ERR_ARG;
    }
    const int T = np / T64;
    if (g_potrf_variant == 3) {
        int* state = (int*)(work + (size_t)np * T64);
        TRY_RC(pbx_potrf_dag(A, np, ld, work, info_dev, state, 0, s));
        return PBX_OK;
    }
    if (g_potrf_variant == 5 && pbx_persist_fits(np, device_cus())) {
        int* state = (int*)(work + (size_t)np * T64);
        TRY_RC(pbx_potrf_persist(A, np, ld, work, info_dev, state, device_cus(), s));
        return PBX_OK;
    }
    if (g_potrf_variant >= 4) {
        int* la = (int*)(work + (size_t)np * T64);
        hipLaunchKernelGGL(potrf_diag_kernel<5>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev, la);
        if (T > 1) hipLaunchKernelGGL(trsm_kernel, dim3(T - 1), dim3(256), 0, s, A, ld, 0, work);
        for (int k = 0; k + 1 < T; ++k) {
            int cnt = (T - 1 - k) * (T - k) / 2;
            hipLaunchKernelGGL(syrk_trsm_kernel<5>, dim3(cnt), dim3(256), 0, s, A, ld, k, work, info_dev, la);
        }
        hipLaunchKernelGGL(la_finish_kernel, dim3(1), dim3(64), 0, s, la, info_dev);
        PBX_HIP(hipGetLastError());
        return PBX_OK;
    }
    const bool v2 = g_potrf_variant == 2;
    if (v2)
        hipLaunchKernelGGL(potrf_diag_kernel<2>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev, nullptr);
    else
        hipLaunchKernelGGL(potrf_diag_kernel<1>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev, nullptr);
    for (int k = 0; k + 1 < T; ++k) {
        hipLaunchKernelGGL(trsm_kernel, dim3(T - 1 - k), dim3(256), 0, s, A, ld, k, work);
        int cnt = (T - 1 - k) * (T - k) / 2;
        if (v2)
            hipLaunchKernelGGL(syrk_kernel<2>, dim3(cnt), dim3(256), 0, s, A, ld, k, work, info_dev);
        else
            hipLaunchKernelGGL(syrk_kernel<1>, dim3(cnt), dim3(256), 0, s, A, ld, k, work, info_dev);
    }
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// A := scale * inv(A) for SPD A (both triangles written).  X is an np x np
// scratch matrix (leading dimension ld).
int pbx_spd_inverse(double* A, int np, int ld, double* work, double* X, int* info_dev, hipStream_t s,
                    double scale) {
    int rc = pbx_potrf(A, np, ld, work, info_dev, s);
    if (rc) return rc;
    const int T = np / T64;
    PBX_HIP(hipMemsetAsync(X, 0, sizeof(double) * (size_t)np * ld, s));
    hipLaunchKernelGGL(copy_diag_inv_kernel, dim3(T), dim3(256), 0, s, X, ld, work);
    for (int i = 1; i < T; ++i)
        hipLaunchKernelGGL(trtri_row_kernel, dim3(i), dim3(256), 0, s, A, ld, X, ld, i, work);
    hipLaunchKernelGGL(lauum_kernel, dim3(T * (T + 1) / 2), dim3(256), 0, s, X, ld, A, ld, T, scale);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_gemm_nt_test(const double* A, const double* B, double* C, int M, int N, int K, hipStream_t s) {
    if (M % 64 || N % 64 || K % 64) {
        pbx_set_error(PBX_ERR_ARG, "gemm_nt: dimensions must be multiples of 64");
        return PBX_ERR_ARG;
    }
    hipLaunchKernelGGL(gemm_nt_kernel, dim3(N / 64, M / 64), dim3(256), 0, s, A, B, C, M, N, K);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// Dense FP64 linear algebra on gfx950 for the 3N x 3N diffusion and
// resistance matrices: blocked Cholesky (potrf), triangular inverse and the
// SPD inverse built from them.
//
// Matrices are row-major, order np (a multiple of 64) with leading dimension
// ld; only the lower triangle is read.  Work is cut into 64 x 64 tiles; every
// tile product runs on the FP64 matrix cores (v_mfma_f64_16x16x4_f64): a
// 256-thread workgroup owns one 64 x 64 output tile, each of its four waves a
// 32 x 32 quadrant held as 2 x 2 MFMA accumulators.  Operands are staged in
// LDS k-major (s[k][r]) so each quarter-wave reads 16 consecutive doubles.
//
// Cholesky, right-looking with look-ahead fused into the update kernel:
//   potrf_diag(0)
//   for k = 0 .. T-2:
//     trsm(k):   L_ik = A_ik * inv(L_kk)^T           (one workgroup per tile row)
//     syrk(k):   A_ij -= L_ik L_jk^T  for k < j <= i   (one workgroup per tile)
//                and the workgroup owning tile (k+1,k+1) factors it in LDS
// inv(L_kk) is produced with each diagonal factorisation and kept in `work`
// (T tiles of 64 x 64), where the SPD inverse reuses it.
#include "pbx_common.h"

#define TRY_RC(x)            \
    do {                     \
        int rc_ = (x);       \
        if (rc_) return rc_; \
    } while (0)

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int T64 = 64;
constexpr int SP = 65;  // LDS row stride in doubles

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Stage a 64 x 64 global tile into LDS as s[k][r] = op(X)(r, k).
//   rk = true : X is indexed X[r][k] (row r, column k)  -> transposed copy
//   rk = false: X is indexed X[k][r]                     -> straight copy
__device__ __forceinline__ void stage(double* s, const double* X, int ld, bool rk) {
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        int row = e >> 6, col = e & 63;
        double v = X[(size_t)row * ld + col];
        if (rk)
            s[col * SP + row] = v;
        else
            s[row * SP + col] = v;
    }
}

// acc += op(A) op(B) for one 64-deep slab, operands staged k-major.
// acc[x][y][v] <-> C[r0 + 16x + lane/16 + 4v][c0 + 16y + lane%16] (the f64 16x16x4
// accumulator interleaves rows: register v of lane group g holds row g + 4v)
__device__ __forceinline__ void mma64(const double* sA, const double* sB, v4d acc[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
    for (int k = 0; k < T64; k += 4) {
        const double* pa = sA + (k + lk) * SP;
        const double* pb = sB + (k + lk) * SP;
        double a0 = pa[r0 + li], a1 = pa[r0 + 16 + li];
        double b0 = pb[c0 + li], b1 = pb[c0 + 16 + li];
        acc[0][0] = mfma(a0, b0, acc[0][0]);
        acc[0][1] = mfma(a0, b1, acc[0][1]);
        acc[1][0] = mfma(a1, b0, acc[1][0]);
        acc[1][1] = mfma(a1, b1, acc[1][1]);
    }
}

__device__ __forceinline__ void acc_zero(v4d acc[2][2]) {
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
}

// Visit every accumulator element with its tile coordinates.
template <class F>
__device__ __forceinline__ void acc_for_each(v4d acc[2][2], F f) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
            for (int v = 0; v < 4; ++v) f(r0 + 16 * x + lk + 4 * v, c0 + 16 * y + li, acc[x][y][v]);
}

// Factor the SPD 64 x 64 tile held in LDS (lower triangle, stride SP) and
// build the inverse of the factor in sW (lower, zero above).  Column-by-
// column right-looking elimination; row j of inv(L) is formed in the same
// step from rows < j.  The diagonal square roots are kept in sd[] until the
// end so no lane overwrites L[j][j] while others still read it.
__device__ void tile_potrf_inv(double* sL, double* sW, double* sd, int col0, int* info) {
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
            // W[j][c] for c <= j: (delta_jc - sum_{c<=k<j} L[j][k] W[k][c]) / L[j][j]
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
        // trailing update: A[i][c] -= L[i][j] * L[c][j] for j < c <= i
        for (int e = tid; e < T64 * T64; e += blockDim.x) {
            int i = e >> 6, c = e & 63;
            if (c > j && c <= i) sL[i * SP + c] -= sL[i * SP + j] * sL[c * SP + j];
        }
        __syncthreads();
    }
    if (tid < 64) sL[tid * SP + tid] = sd[tid];
    __syncthreads();
}

// Broadcast a double from a uniform lane (v_readlane on both halves).
__device__ __forceinline__ double bcast(double x, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

// 1/sqrt(d): hardware estimate refined by two Newton steps (FMA), then
// sqrt(d) = d * rsqrt(d).  Within an ulp or two of the correctly rounded
// values; non-positive or NaN d propagates NaN.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    double h = 0.5 * d;
    double r = fma(-h * y, y, 0.5);
    y = fma(y, r, y);
    r = fma(-h * y, y, 0.5);
    y = fma(y, r, y);
    return y;
}

// Same contract as tile_potrf_inv with far fewer workgroup barriers:
// 8-column panels are factored by wave 0 alone (lane = row, panel in
// registers, pivots broadcast with readlane); the trailing update
// A22 -= P P^T (K = 8) runs on the FP64 matrix cores across all four
// waves; rows of inv(L) for the panel are formed by waves 1-3 once the
// panel rows are final.
__device__ void tile_potrf_inv_v2(double* sL, double* sW, double* sd, int col0, int* info) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int e = tid; e < T64 * SP; e += blockDim.x) sW[e] = 0.0;
    for (int kb = 0; kb < 8; ++kb) {
        const int c0 = kb * 8;
        if (w == 0) {
            const bool act = lane >= c0;
            double v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = act ? sL[lane * SP + c0 + q] : 0.0;
            int first_bad = -1;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int j = c0 + q;
                const double d = bcast(v[q], j);
                if (!(d > 0.0) && first_bad < 0) first_bad = q;
                const double y = rsqrt_nr(d);
                if (lane == j) {
                    v[q] = d * y;
                    sd[j] = d * y;
                } else if (lane > j) {
                    v[q] = v[q] * y;
                }
#pragma unroll
                for (int q2 = q + 1; q2 < 8; ++q2) v[q2] -= v[q] * bcast(v[q], c0 + q2);
            }
            if (act) {
#pragma unroll
                for (int q = 0; q < 8; ++q) sL[lane * SP + c0 + q] = v[q];
            }
            if (first_bad >= 0 && lane == 0) atomicCAS(info, 0, col0 + c0 + first_bad + 1);
        }
        __syncthreads();
        // inv(L) rows c0..c0+7 (waves 1-3): W[r][c] for c <= r, sequential in r
        // within the panel; uses L rows r (final) and W rows < r.
        if (w > 0) {
            const int t = tid - 64;  // 192 threads: column c = t % 64, split = t / 64
            const int c = t & 63, part = t >> 6;
            for (int r = c0; r < c0 + 8; ++r) {
                double s = 0.0;
                if (c <= r) {
                    for (int k = c + part; k < r; k += 3) s -= sL[r * SP + k] * sW[k * SP + c];
                }
                // reduce the three partial sums through LDS scratch in sW's pad column
                __shared__ double red[3][64];
                red[part][c] = s;
                __builtin_amdgcn_s_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
                if (part == 0 && c <= r) {
                    double tot = red[0][c] + red[1][c] + red[2][c] + (c == r ? 1.0 : 0.0);
                    sW[r * SP + c] = tot / sd[r];
                }
                __builtin_amdgcn_s_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            }
        } else {
            // wave 0 meets the waves-1..3 barriers of the inverse rows
            for (int r = 0; r < 16; ++r) {
                __builtin_amdgcn_s_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            }
        }
        // trailing update of rows/cols >= c0+8 on the matrix cores
        if (kb < 7) {
            const int r0 = (w >> 1) * 32, cq = (w & 1) * 32;
            const int li = lane & 15, lk = lane >> 4;
            const int lo = c0 + 8;
            const bool live = (r0 + 31 >= lo) && !(cq > r0 + 31);
            v4d acc[2][2];
            acc_zero(acc);
            if (live) {
#pragma unroll
                for (int ks = 0; ks < 8; ks += 4) {
                    const int k = c0 + ks + lk;
                    double a0 = sL[(r0 + li) * SP + k], a1 = sL[(r0 + 16 + li) * SP + k];
                    double b0 = sL[(cq + li) * SP + k], b1 = sL[(cq + 16 + li) * SP + k];
                    acc[0][0] = mfma(a0, b0, acc[0][0]);
                    acc[0][1] = mfma(a0, b1, acc[0][1]);
                    acc[1][0] = mfma(a1, b0, acc[1][0]);
                    acc[1][1] = mfma(a1, b1, acc[1][1]);
                }
            }
            __syncthreads();
            if (live) {
                acc_for_each(acc, [&](int r, int c, double v) {
                    if (r >= lo && c >= lo && c <= r) sL[r * SP + c] -= v;
                });
            }
            __syncthreads();
        }
    }
    if (tid < 64) sL[tid * SP + tid] = sd[tid];
    __syncthreads();
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

template <int V>
__device__ __forceinline__ void factor_tile(double* sL, double* sW, double* sd, int col0, int* info) {
    if (V == 2)
        tile_potrf_inv_v2(sL, sW, sd, col0, info);
    else
        tile_potrf_inv(sL, sW, sd, col0, info);
}

template <int V>
__global__ __launch_bounds__(256) void potrf_diag_kernel(double* A, int ld, int k, double* work, int* info) {
    __shared__ double sL[T64 * SP];
    __shared__ double sW[T64 * SP];
    __shared__ double sd[T64];
    const double* Akk = A + (size_t)k * T64 * ld + (size_t)k * T64;
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        int r = e >> 6, c = e & 63;
        sL[r * SP + c] = Akk[(size_t)r * ld + c];
    }
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
    stage(sA, Aik, ld, true);                                  // op(A) = A_ik
    stage(sB, work + (size_t)k * T64 * T64, T64, true);        // op(B)(kk,c) = W[c][kk]
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
__global__ __launch_bounds__(256) void syrk_kernel(double* A, int ld, int k, double* work, int* info) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double sd[T64];
    int ii, jj;
    tri_decode(blockIdx.x, ii, jj);
    const int i = k + 1 + ii, j = k + 1 + jj;
    const double* Aik = A + (size_t)i * T64 * ld + (size_t)k * T64;
    const double* Ajk = A + (size_t)j * T64 * ld + (size_t)k * T64;
    double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
    stage(sA, Aik, ld, true);
    if (i != j) stage(sB, Ajk, ld, true);
    __syncthreads();
    v4d acc[2][2];
    acc_zero(acc);
    mma64(sA, i == j ? sA : sB, acc);
    if (blockIdx.x != 0) {
        if (i == j) {
            acc_for_each(acc, [&](int r, int c, double v) {
                if (c <= r) Aij[(size_t)r * ld + c] -= v;
            });
        } else {
            acc_for_each(acc, [&](int r, int c, double v) { Aij[(size_t)r * ld + c] -= v; });
        }
        return;
    }
    // look-ahead: update, factor and invert tile (k+1, k+1) in LDS
    __syncthreads();  // all waves done reading sA
    double* sL = sA;
    double* sW = sB;
    acc_for_each(acc, [&](int r, int c, double v) { sL[r * SP + c] = Aij[(size_t)r * ld + c] - v; });
    __syncthreads();
    factor_tile<V>(sL, sW, sd, (k + 1) * T64, info);
    store_diag(sL, sW, A, ld, k + 1, work);
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
        stage(sA, A + (size_t)i * T64 * ld + (size_t)m * T64, ld, true);    // op(A) = L_im
        stage(sB, X + (size_t)m * T64 * ldx + (size_t)j * T64, ldx, false);  // op(B) = X_mj
        __syncthreads();
        mma64(sA, sB, acc);
        __syncthreads();
    }
    // X_ij = -W_i * S
    acc_for_each(acc, [&](int r, int c, double v) { sB[r * SP + c] = v; });  // s[k=r][c] = S[r][c]
    stage(sA, work + (size_t)i * T64 * T64, T64, true);                        // op(A) = W_i
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
        stage(sA, X + (size_t)m * T64 * ldx + (size_t)i * T64, ldx, false);  // op(A)(r,k) = X_mi[k][r]
        stage(sB, X + (size_t)m * T64 * ldx + (size_t)j * T64, ldx, false);  // op(B)(k,c) = X_mj[k][c]
        __syncthreads();
        mma64(sA, sB, acc);
        __syncthreads();
    }
    double* Yij = Y + (size_t)i * T64 * ld + (size_t)j * T64;
    double* Yji = Y + (size_t)j * T64 * ld + (size_t)i * T64;
    acc_for_each(acc, [&](int r, int c, double v) {
        double s = v * scale;
        Yij[(size_t)r * ld + c] = s;
        Yji[(size_t)c * ld + r] = s;
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
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        int r = e >> 6, c = e & 63;
        d[(size_t)r * ld + c] = s[(size_t)r * ld + c];
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
        stage(sA, A + (size_t)ti * T64 * K + k, K, true);
        stage(sB, B + (size_t)tj * T64 * K + k, K, true);
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

// work = [T inverse tiles of the diagonal factor][single-launch DAG state]
size_t pbx_dense_work_doubles(int np) { return (size_t)np * T64 + (pbx_dag_state_ints(np) + 1) / 2; }

// Cholesky variant: 1 = launch per panel, column-by-column diagonal tiles;
// 2 = launch per panel, 8-column panel diagonal tiles; 3 = single persistent
// launch over the tile dependency graph (pbx_chol_dag.hip).
static int g_potrf_variant = 2;
extern "C" int pbx_set_potrf_variant(int v) {
    if (v < 1 || v > 3) return PBX_ERR_ARG;
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
// factor tiles are left in work (np*64 doubles).  *info_dev must be zero on
// entry; it receives column+1 of the first non-positive pivot.
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
    const bool v2 = g_potrf_variant == 2;
    if (v2)
        hipLaunchKernelGGL(potrf_diag_kernel<2>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev);
    else
        hipLaunchKernelGGL(potrf_diag_kernel<1>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev);
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

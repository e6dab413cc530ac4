// 64 x 64 FP64 tile primitives shared by the factorisation kernels
// (pbx_dense.hip, pbx_chol_dag.hip).  All helpers assume a 256-thread
// workgroup (four waves).
//
// LDS operand layout: s[k * SP + r] = op(X)(r, k) ("k-major"), so a
// quarter-wave reads 16 consecutive doubles for an MFMA operand.
// v_mfma_f64_16x16x4_f64 operands: lane l holds A[l%16][l/16] and
// B[l/16][l%16]; the 16x16 accumulator interleaves rows: register v of lane
// group g = l/16 holds D[g + 4v][l%16].
#pragma once
#include "pbx_common.h"

namespace pbx_tile {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

constexpr int T64 = 64;
constexpr int SP = 65;  // LDS row stride (odd: staggers rows across banks)

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Stage a 64 x 64 global tile (leading dimension ld, 16-byte aligned rows)
// into LDS.  rk = true: s[col][row] = X[row][col] (transposed); rk = false:
// s[row][col] = X[row][col].  All eight 16-byte loads per thread are issued
// before any LDS store so their latencies overlap.
template <bool RK>
__device__ __forceinline__ void stage(double* s, const double* X, int ld) {
    v2d v[8];
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int e = threadIdx.x + it * 256;  // 2048 pairs
        const int row = e >> 5, col = (e & 31) * 2;
        v[it] = *(const v2d*)(X + (size_t)row * ld + col);
    }
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int e = threadIdx.x + it * 256;
        const int row = e >> 5, col = (e & 31) * 2;
        if (RK) {
            s[col * SP + row] = v[it].x;
            s[(col + 1) * SP + row] = v[it].y;
        } else {
            s[row * SP + col] = v[it].x;
            s[row * SP + col + 1] = v[it].y;
        }
    }
}

// acc += op(A) op(B) over a 64-deep slab; wave w owns the 32 x 32 quadrant
// (w/2, w%2) as 2 x 2 MFMA blocks.
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
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
}

// Tile coordinates of accumulator element (x, y, v) of this thread.
__device__ __forceinline__ int acc_row(int x, int v) {
    return ((threadIdx.x >> 7) * 32) + 16 * x + ((threadIdx.x & 63) >> 4) + 4 * v;
}
__device__ __forceinline__ int acc_col(int y) { return (((threadIdx.x >> 6) & 1) * 32) + 16 * y + (threadIdx.x & 15); }

template <class F>
__device__ __forceinline__ void acc_for_each(v4d acc[2][2], F f) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) f(acc_row(x, v), acc_col(y), acc[x][y][v]);
}

// Read the 16 global values under this thread's accumulator positions
// (all loads issued before use).
__device__ __forceinline__ void acc_load(const double* C, int ld, double out[16]) {
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[q++] = C[(size_t)acc_row(x, v) * ld + acc_col(y)];
}

__device__ __forceinline__ double bcast(double x, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

// 1/sqrt(d): hardware estimate refined by two Newton steps (within an ulp
// or two of the correctly rounded value; d <= 0 or NaN propagates NaN).
__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    double h = 0.5 * d;
    double r = fma(-h * y, y, 0.5);
    y = fma(y, r, y);
    r = fma(-h * y, y, 0.5);
    y = fma(y, r, y);
    return y;
}

// Factor the SPD tile held in LDS (lower triangle, stride SP) in place and
// form W = inv(L) (lower, zero above) in sW.
//
// The trailing matrix and an augmented block R (initially I) live in the
// MFMA accumulators of the four waves for the whole factorisation, so the
// elimination applied to A is applied to R as well and R ends as inv(L)
// (Gauss-Jordan on [L | I]).  Eight 8-column panels:
//   1. owners copy the panel's columns of A and rows of R to LDS;
//   2. wave 0 factors the panel (lane = row, pivots broadcast with
//      readlane), then solves the panel's 8 rows of inv(L) (lane = column);
//   3. all waves apply A -= P P^T and R -= P W_p (K = 8) on the FP64 matrix
//      cores, with rows above the panel masked out through the A operand.
// info receives col0 + column + 1 of the first non-positive pivot.
__device__ inline void tile_factor(double* sL, double* sW, double* sd, int col0, int* info) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r0 = (w >> 1) * 32, cq = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    v4d accL[2][2], accR[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = acc_row(x, v), c = acc_col(y);
                accL[x][y][v] = sL[r * SP + c];
                accR[x][y][v] = (r == c) ? 1.0 : 0.0;
            }
    __syncthreads();
    for (int p = 0; p < 8; ++p) {
        const int c0 = 8 * p;
        // 1. panel columns of A (rows >= c0) and panel rows of R to LDS
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = acc_row(x, v), c = acc_col(y);
                    if (c >= c0 && c < c0 + 8 && r >= c0) sL[r * SP + c] = accL[x][y][v];
                    if (r >= c0 && r < c0 + 8) sW[r * SP + c] = accR[x][y][v];
                }
        __syncthreads();
        // 2. wave 0: panel factorisation, then the panel rows of inv(L)
        if (w == 0) {
            double v[8], yq[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = (lane >= c0) ? sL[lane * SP + c0 + q] : 0.0;
            int first_bad = -1;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int j = c0 + q;
                const double d = bcast(v[q], j);
                if (!(d > 0.0) && first_bad < 0) first_bad = q;
                const double y = rsqrt_nr(d);
                yq[q] = y;
                const double lj = d * y;
                const double vs = v[q] * y;
                v[q] = (lane == j) ? lj : ((lane > j) ? vs : v[q]);
#pragma unroll
                for (int q2 = q + 1; q2 < 8; ++q2) v[q2] = fma(-v[q], bcast(v[q], c0 + q2), v[q2]);
            }
            if (lane >= c0) {
#pragma unroll
                for (int q = 0; q < 8; ++q) sL[lane * SP + c0 + q] = v[q];
            }
            // diagonal of the panel
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (lane == c0 + q) sd[c0 + q] = v[q];
            if (first_bad >= 0 && lane == 0) atomicCAS(info, 0, col0 + c0 + first_bad + 1);
            // W[c0+q][c] = (R[c0+q][c] - sum_{q'<q} L[c0+q][c0+q'] W[c0+q'][c]) / L[c0+q][c0+q]
            double wv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) wv[q] = sW[(c0 + q) * SP + lane];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
#pragma unroll
                for (int q2 = 0; q2 < q; ++q2) wv[q] = fma(-bcast(v[q2], c0 + q), wv[q2], wv[q]);
                wv[q] = wv[q] * yq[q];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) sW[(c0 + q) * SP + lane] = wv[q];
        }
        __syncthreads();
        // 3. trailing updates of A and R on the matrix cores
        if (p < 7) {
            const int lo = c0 + 8;
            const bool rows_live = (r0 + 31 >= lo) && !(cq > r0 + 31);
            const bool liveL = rows_live && (cq + 31 >= lo);
            if (rows_live) {
#pragma unroll
                for (int ks = 0; ks < 8; ks += 4) {
                    const int k = c0 + ks + lk;
                    const int ra = r0 + li, rb = r0 + 16 + li;
                    const double a0 = (ra >= lo) ? -sL[ra * SP + k] : 0.0;
                    const double a1 = (rb >= lo) ? -sL[rb * SP + k] : 0.0;
                    const double w0 = sW[k * SP + cq + li], w1 = sW[k * SP + cq + 16 + li];
                    accR[0][0] = mfma(a0, w0, accR[0][0]);
                    accR[0][1] = mfma(a0, w1, accR[0][1]);
                    accR[1][0] = mfma(a1, w0, accR[1][0]);
                    accR[1][1] = mfma(a1, w1, accR[1][1]);
                    if (liveL) {
                        const double b0 = sL[(cq + li) * SP + k], b1 = sL[(cq + 16 + li) * SP + k];
                        accL[0][0] = mfma(a0, b0, accL[0][0]);
                        accL[0][1] = mfma(a0, b1, accL[0][1]);
                        accL[1][0] = mfma(a1, b0, accL[1][0]);
                        accL[1][1] = mfma(a1, b1, accL[1][1]);
                    }
                }
            }
        }
    }
    __syncthreads();
}

}  // namespace pbx_tile

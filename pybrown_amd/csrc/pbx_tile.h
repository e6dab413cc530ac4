// 64 x 64 FP64 tile primitives shared by the factorisation kernels
// (pbx_dense.hip).  All helpers assume a 256-thread
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
// The two halves are also available separately (stage_fetch into registers,
// stage_put into LDS) so a caller can fetch the next operand while the
// matrix cores work on the current one.
__device__ __forceinline__ void stage_fetch(v2d v[8], const double* X, int ld) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int e = threadIdx.x + it * 256;  // 2048 pairs
        const int row = e >> 5, col = (e & 31) * 2;
        v[it] = *(const v2d*)(X + (size_t)row * ld + col);
    }
}

template <bool RK>
__device__ __forceinline__ void stage_put(double* s, const v2d v[8]) {
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

template <bool RK>
__device__ __forceinline__ void stage(double* s, const double* X, int ld) {
    v2d v[8];
    stage_fetch(v, X, ld);
    stage_put<RK>(s, v);
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

// ---------------------------------------------------------------------------
// FP32 tile products (v_mfma_f32_16x16x4_f32) for the mixed-precision bulk
// updates.  Operands are rounded to FP32 while staged; products accumulate in
// FP32.  Operand lanes are as for the FP64 instruction, but the FP32
// accumulator does not interleave rows: register v of lane group g holds
// D[4g + v][l%16] (checked against the FP64 layout on the device).
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int SPF = 65;  // FP32 LDS row stride

__device__ __forceinline__ v4f mfma32(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// s[col * SPF + row] = (float)X[row][col] (op-layout staging of an FP64
// tile fetched with stage_fetch)
__device__ __forceinline__ void stage_put_f32t(float* s, const v2d v[8]) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int e = threadIdx.x + it * 256;
        const int row = e >> 5, col = (e & 31) * 2;
        s[col * SPF + row] = (float)v[it].x;
        s[(col + 1) * SPF + row] = (float)v[it].y;
    }
}

__device__ __forceinline__ void stage_f32t(float* s, const double* X, int ld) {
    v2d v[8];
    stage_fetch(v, X, ld);
    stage_put_f32t(s, v);
}

// s[row * SPF + col] = (float)X[row][col] (the tile as the B operand, or its
// transpose as the A operand)
__device__ __forceinline__ void stage_put_f32n(float* s, const v2d v[8]) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int e = threadIdx.x + it * 256;
        const int row = e >> 5, col = (e & 31) * 2;
        s[row * SPF + col] = (float)v[it].x;
        s[row * SPF + col + 1] = (float)v[it].y;
    }
}

__device__ __forceinline__ void mma64_f32(const float* sA, const float* sB, v4f acc[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
    for (int k = 0; k < T64; k += 4) {
        const float* pa = sA + (k + lk) * SPF;
        const float* pb = sB + (k + lk) * SPF;
        const float a0 = pa[r0 + li], a1 = pa[r0 + 16 + li];
        const float b0 = pb[c0 + li], b1 = pb[c0 + 16 + li];
        acc[0][0] = mfma32(a0, b0, acc[0][0]);
        acc[0][1] = mfma32(a0, b1, acc[0][1]);
        acc[1][0] = mfma32(a1, b0, acc[1][0]);
        acc[1][1] = mfma32(a1, b1, acc[1][1]);
    }
}

__device__ __forceinline__ void acc32_zero(v4f acc[2][2]) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (v4f){0.f, 0.f, 0.f, 0.f};
}

// Tile row of FP32 accumulator element (x, v) of this thread (columns as acc_col).
__device__ __forceinline__ int acc32_row(int x, int v) {
    return ((threadIdx.x >> 7) * 32) + 16 * x + 4 * ((threadIdx.x & 63) >> 4) + v;
}

// FP64 sums of FP32 slab products, in the FP32 accumulator's layout: each
// 64-deep slab accumulates in FP32 on the matrix cores and is then added in
// FP64, so accumulation rounding grows with the slab depth, not with the
// whole inner dimension.
__device__ __forceinline__ void acc32_add_to(const v4f acc[2][2], v4d sum[2][2]) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) sum[x][y][v] += (double)acc[x][y][v];
}

template <class F>
__device__ __forceinline__ void acc32d_for_each(v4d sum[2][2], F f) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) f(acc32_row(x, v), acc_col(y), sum[x][y][v]);
}

__device__ __forceinline__ void acc32_load(const double* C, int ld, double out[16]) {
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[q++] = C[(size_t)acc32_row(x, v) * ld + acc_col(y)];
}

__device__ __forceinline__ double bcast(double x, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

// 1/sqrt(d): the hardware estimate y0 (relative error below 6e-8 on
// gfx950) corrected to second order, y0 (1 + e/2 + 3 e^2 / 8) with
// e = 1 - d y0^2: the truncation term is O(e^3), far below an ulp, and the
// result is within one ulp of 1/sqrt(d) (as two Newton steps are), with four
// dependent operations after the estimate instead of six.  d <= 0 or NaN
// propagates NaN.
__device__ __forceinline__ double rsqrt_nr(double d) {
    const double y = __builtin_amdgcn_rsq(d);
    const double e = fma(-d * y, y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
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

// ---------------------------------------------------------------------------
// Wave-specialised tile factorisation: same contract as tile_factor (L in sL,
// inv(L) in sW, diagonal in sd, status in info), with the four waves working
// concurrently on one 64 x 64 tile instead of in lock step:
//   wave 0  the serial L chain: factors 8-column panel p (lane = row),
//           publishes it, then applies panel p to panel p+1's columns
//           itself (row entries read as LDS broadcasts).  Every lane first
//           factors the panel's 8 x 8 diagonal block in its own registers
//           (read once from LDS) and then solves for its row, so no column
//           waits on a cross-lane broadcast.  The operations and their order
//           are those of a column-by-column chain with the pivots and
//           multipliers broadcast from their lanes, and give the same bits;
//   waves 2, 3  hold the trailing L blocks in MFMA accumulators, apply each
//           published panel and stage the columns of panel p+2 (updated
//           through p) one panel ahead of wave 0;
//   wave 1  forms the rows of inv(L) a panel behind the L chain
//           (Gauss-Jordan on [L | I]: W rows of panel p by substitution with
//           the panel's 8 x 8 diagonal block, then R -= L_panel W_p).
// Hand-offs are LDS counters with workgroup-scope release/acquire; every
// wait targets a counter released by an earlier step of another wave, so the
// waits always end (each is also bounded).  Needs exactly 256 threads.
//
// FUSE (the first tile of a two-column launch): waves 2 and 3 also stage,
// once another workgroup has published them and between their own steps
// (never waiting for them), the tiles the products after the factorisation
// need: C = tile (c0+1, c0) into fx->sC in operand layout and T = tile
// (c0+1, c0+1) into fx->sT row-major, 32 rows each.  The result is true when
// both halves were staged.

// Inputs of the products between the two factorisations of a two-column
// launch (see fused_between in pbx_dense.hip).
struct FuseCtx {
    const double* C;  // tile (c0+1, c0) through the lagging panels (global)
    int ldc;
    const double* T;  // tile (c0+1, c0+1) through the lagging panels (global)
    int ldt;
    const int* ready; // both are published once *ready >= target
    int target;
    double* sC;       // LDS, 64 x SP: C, sC[k * SP + r] = C[r][k]
    double* sT;       // LDS, 64 x SP: T, row-major
};

// Bounded wait on an LDS counter; an exhausted budget is reported through
// the status word (the factorisation is then not valid).
__device__ __forceinline__ void lds_wait_ge(int* f, int target, int* info) {
    int it = 0;
    while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
        if (++it >= (1 << 24)) {
            atomicCAS(info, 0, PBX_INFO_DAG_WAIT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ void lds_release_add(int* f, int v) {
    __hip_atomic_fetch_add(f, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// FUSE: wave w (2 or 3) stages its 32 rows of C (operand layout) and T
// (row-major) once their producers have published them
__device__ __forceinline__ void fuse_stage(const FuseCtx* fx, int w, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int r0 = 32 * (w - 2);
#pragma unroll
    for (int i0 = 0; i0 < 32; i0 += 8) {
        double cv[8], tv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            cv[i] = fx->C[(size_t)(r0 + i0 + i) * fx->ldc + lane];
            tv[i] = fx->T[(size_t)(r0 + i0 + i) * fx->ldt + lane];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            fx->sC[lane * SP + r0 + i0 + i] = cv[i];
            fx->sT[(r0 + i0 + i) * SP + lane] = tv[i];
        }
    }
}

// Wg (optional): wave 1 also writes each panel's rows of inv(L) there
// (64 x 64, row-major) with device-scope stores as it forms them, so a
// publication after the factorisation finds them mostly written.
template <bool FUSE = false>
__device__ __forceinline__ bool tile_factor_ws(double* sL, double* sW, double* sd, int col0, int* info,
                                      const FuseCtx* fx = nullptr, double* Wg = nullptr) {
    // [0]: panels published by wave 0; [1], [2]: panels whose next-but-one
    // columns wave 2, wave 3 has staged.  One counter per producer: a shared
    // count of both waves' stagings could reach its target while one wave was
    // a panel ahead and the other still staging.  [3] (FUSE): halves of C
    // and T staged.
    __shared__ int fl[4];
    __shared__ double sy[64];    // 1 / L_jj
    __shared__ double spiv[64];  // pivots before the root (checked by wave 1)
    __shared__ double sB[64];  // the next panel's diagonal block, row-major 8 x 8
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    // trailing L blocks (block row X, block column Y) of waves 2 and 3; the
    // split balances the number of panel updates each block receives
    const int BX2[5] = {3, 2, 3, 1, 3}, BY2[5] = {3, 1, 1, 0, 0};
    const int BX3[5] = {2, 3, 1, 0, 2}, BY3[5] = {2, 2, 1, 0, 0};
    v4d acc[10];
    int bx[5], by[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        bx[s] = (w == 2) ? BX2[s] : BX3[s];
        by[s] = (w == 2) ? BY2[s] : BY3[s];
    }
    if (tid == 0) {
        fl[0] = 0;
        fl[1] = 0;
        fl[2] = 0;
        fl[3] = 0;
    }
    if (w >= 2) {
#pragma unroll
        for (int s = 0; s < 5; ++s)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[s][v] = sL[(16 * bx[s] + lk + 4 * v) * SP + 16 * by[s] + li];
    } else if (w == 1) {
        // R = I; lower block (X, Y) is acc[X (X + 1) / 2 + Y]
#pragma unroll
        for (int X = 0; X < 4; ++X)
#pragma unroll
            for (int Y = 0; Y <= X; ++Y)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[X * (X + 1) / 2 + Y][v] = (X == Y && lk + 4 * v == li) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (w == 0) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = sL[lane * SP + q];
        for (int p = 0; p < 8; ++p) {
            const int c0 = 8 * p;
            double yq[8], piv[8];
            // Dm: lower 8 x 8 diagonal block, packed (k, m) -> k (k + 1) / 2 + m,
            // from sL for the first panel, else as the previous step left it in sB
            double Dm[36];
            const double* src = (p == 0) ? sL : sB;
            const int lds = (p == 0) ? SP : 8;
#pragma unroll
            for (int k = 0; k < 8; ++k)
#pragma unroll
                for (int m = 0; m <= k; ++m) Dm[k * (k + 1) / 2 + m] = src[k * lds + m];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double d = Dm[q * (q + 1) / 2 + q];
                piv[q] = d;
                const double y = rsqrt_nr(d);
                yq[q] = y;
#pragma unroll
                for (int k = q; k < 8; ++k) Dm[k * (k + 1) / 2 + q] *= y;
#pragma unroll
                for (int m = q + 1; m < 8; ++m)
#pragma unroll
                    for (int k = m; k < 8; ++k)
                        Dm[k * (k + 1) / 2 + m] =
                            fma(-Dm[k * (k + 1) / 2 + q], Dm[m * (m + 1) / 2 + q], Dm[k * (k + 1) / 2 + m]);
            }
            // this lane's row of the panel (rows above c0 are dead)
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                double t = v[q];
#pragma unroll
                for (int k = 0; k < q; ++k) t = fma(-v[k], Dm[q * (q + 1) / 2 + k], t);
                v[q] = t * yq[q];
            }
            if (lane >= c0) {
#pragma unroll
                for (int q = 0; q < 8; ++q) sL[lane * SP + c0 + q] = v[q];
            }
            // the block's diagonal, reciprocal roots and pivots are the same in
            // every lane (the diagonal equals lane c0+q's own solve to the
            // bit): one lane stores them; wave 1 checks the pivots
            if (lane == 0) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    sd[c0 + q] = Dm[q * (q + 1) / 2 + q];
                    sy[c0 + q] = yq[q];
                    spiv[c0 + q] = piv[q];
                }
                lds_release_add(&fl[0], 1);
            }
            if (p == 7) break;
            // panel p+1's columns (updated through p-1 by waves 2 and 3) less
            // panel p's contribution
            const int n0 = c0 + 8;
            // panel p's L at rows n0..n0+7 (this wave's own stores above),
            // all requested before the wait so their latency overlaps it
            // (read as needed, each LDS round trip was exposed in turn)
            double lr[64];
#pragma unroll
            for (int i = 0; i < 64; ++i) lr[i] = sL[(n0 + (i >> 3)) * SP + c0 + (i & 7)];
            __builtin_amdgcn_sched_barrier(0);
            if (p > 0) {
                lds_wait_ge(&fl[1], p, info);
                lds_wait_ge(&fl[2], p, info);
            }
            double nv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double t = sL[lane * SP + n0 + q];  // (a plain load: no branch per element)
                nv[q] = (lane >= n0) ? t : 0.0;
            }
#pragma unroll
            for (int q2 = 0; q2 < 8; ++q2) {
                double t = nv[q2];
#pragma unroll
                for (int q = 0; q < 8; ++q) t = fma(-v[q], lr[q2 * 8 + q], t);
                nv[q2] = t;
            }
            if (lane >= n0 && lane < n0 + 8) {
#pragma unroll
                for (int q = 0; q < 8; ++q) sB[(lane - n0) * 8 + q] = nv[q];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = nv[q];
            __builtin_amdgcn_wave_barrier();
        }
    } else if (w >= 2) {
        bool staged = false;  // FUSE: this wave's halves of C and T
        for (int p = 0; p < 7; ++p) {
            const int c0 = 8 * p, lo = c0 + 8;
            // FUSE: the readiness word, requested before the wait so its
            // latency overlaps it
            const int rd = (FUSE && !staged)
                               ? __hip_atomic_load(fx->ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : 0;
            lds_wait_ge(&fl[0], p + 1, info);
            const int n2 = c0 + 16;  // first column of panel p+2
            // pass 0: blocks holding panel p+2's columns, which are staged;
            // pass 1: the rest
#pragma unroll
            for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    const int X = bx[s], Y = by[s];
                    const bool live = 16 * Y + 15 >= lo;
                    const bool early = (n2 < 64) && (16 * Y <= n2) && (n2 < 16 * Y + 16);
                    if (!live || early != (pass == 0)) continue;
#pragma unroll
                    for (int ks = 0; ks < 8; ks += 4) {
                        const int k = c0 + ks + lk;
                        const int ra = 16 * X + li;
                        const double ta = sL[ra * SP + k];
                        const double a0 = (ra >= lo) ? -ta : 0.0;
                        const double b0 = sL[(16 * Y + li) * SP + k];
                        acc[s] = mfma(a0, b0, acc[s]);
                    }
                    if (early) {
                        const int c = 16 * Y + li;
                        if (c >= n2 && c < n2 + 8)
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int r = 16 * X + lk + 4 * v;
                                if (r >= n2) sL[r * SP + c] = acc[s][v];
                            }
                    }
                }
                if (pass == 0 && lane == 0) lds_release_add(&fl[w - 1], 1);
            }
            if (FUSE && !staged && rd >= fx->target) {
                fuse_stage(fx, w, lane);
                staged = true;
                if (lane == 0) lds_release_add(&fl[3], 1);
            }
        }
        // one more look once this wave's part is done (waves 0 and 1 still
        // have their last panels ahead).  (Polling here until wave 1 finished
        // caught every launch's jobs but made the factorisation 3 us slower.)
        if (FUSE && !staged &&
            __hip_atomic_load(fx->ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= fx->target) {
            fuse_stage(fx, w, lane);
            if (lane == 0) lds_release_add(&fl[3], 1);
        }
    } else {
        for (int p = 0; p < 8; ++p) {
            const int c0 = 8 * p, lo = c0 + 8;
            lds_wait_ge(&fl[0], p + 1, info);
            // the panel's first non-positive pivot (NaN included), if any
            if (lane < 8) {
                const unsigned long long bad = __ballot(!(spiv[c0 + lane] > 0.0));
                if (lane == 0 && bad) atomicCAS(info, 0, col0 + c0 + __builtin_ctzll(bad) + 1);
            }
            // R rows c0..c0+7 (columns <= c0+7 lie in block row X = c0 / 16)
            const int X = c0 >> 4;
#pragma unroll
            for (int XX = 0; XX < 4; ++XX) {
                if (XX != X) continue;
#pragma unroll
                for (int Y = 0; Y <= XX; ++Y)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = 16 * XX + lk + 4 * v, c = 16 * Y + li;
                        if (r >= c0 && r < c0 + 8) sW[r * SP + c] = acc[XX * (XX + 1) / 2 + Y][v];
                    }
            }
            __builtin_amdgcn_wave_barrier();
            // W rows: lane = column, substitution with the panel's diagonal block
            double wv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double t = sW[(c0 + q) * SP + lane];
                wv[q] = (lane < c0 + 8) ? t : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double* Lr = sL + (c0 + q) * SP + c0;
                double t = wv[q];
#pragma unroll
                for (int q2 = 0; q2 < q; ++q2) t = fma(-Lr[q2], wv[q2], t);
                wv[q] = t * sy[c0 + q];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) sW[(c0 + q) * SP + lane] = wv[q];
            if (Wg) {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    __hip_atomic_store(Wg + (c0 + q) * 64 + lane, wv[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __builtin_amdgcn_wave_barrier();
            if (p == 7) break;
            // R[rows >= lo] -= L[rows, panel] W_p over blocks with columns <= c0+7
#pragma unroll
            for (int XB = 0; XB < 4; ++XB)
#pragma unroll
                for (int YB = 0; YB <= XB; ++YB) {
                    if (16 * XB + 15 < lo || 16 * YB > c0 + 7) continue;
#pragma unroll
                    for (int ks = 0; ks < 8; ks += 4) {
                        const int k = c0 + ks + lk;
                        const int ra = 16 * XB + li;
                        const double ta = sL[ra * SP + k];
                        const double a0 = (ra >= lo) ? -ta : 0.0;
                        const double b0 = sW[k * SP + 16 * YB + li];
                        acc[XB * (XB + 1) / 2 + YB] = mfma(a0, b0, acc[XB * (XB + 1) / 2 + YB]);
                    }
                }
        }
    }
    __syncthreads();
    return !FUSE || fl[3] == 2;
}

}  // namespace pbx_tile

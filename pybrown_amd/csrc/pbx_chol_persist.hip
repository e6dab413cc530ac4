// Persistent tiled Cholesky on gfx950 (variant 5): one launch, one
// workgroup per CU, tiles owned for the whole factorisation.
//
// Workgroup 0 walks the diagonal.  At step k it takes tile (k, k-1) and tile
// (k, k), both updated through panel k-2 by their owners, forms
// L_{k,k-1} = A_{k,k-1} inv(L_{k-1,k-1})^T, applies panel k-1 to (k, k),
// factors it in LDS with the wave-specialised tile step and publishes
// L_kk and inv(L_kk).  Its chain per step is one tile solve, one tile product
// and one tile factorisation; everything else runs beside it.
//
// Every other lower tile (i, j) belongs to one worker workgroup (tiles dealt
// round-robin in row-major order), which applies its updates in place:
//   diagonal (j, j):   panels 0 .. j-2, then handed to workgroup 0;
//   sub-diagonal (j+1, j):  panels 0 .. j-1, then handed to workgroup 0;
//   other (i, j):      panels 0 .. j-1, then L_ij = C inv(L_jj)^T.
// Workers advance panel by panel: at step k they first finish their tiles
// whose updates are complete (solves waiting for inv(L_kk)), then apply panel
// k to their remaining tiles (rows in increasing order, so tiles near the
// diagonal come first) and hand tiles over as soon as their last update is in.
// Compared with one launch per panel, no launch boundary separates panels:
// the HBM-bound trailing updates of early panels run beside the latency-bound
// diagonal chain instead of alternating with it.  Measured at n = 3000 it is
// slower than variant 4 (2.2 ms against 1.4 ms): each worker applies its
// updates one at a time behind two device-scope flag acquisitions, so early
// panels are paced by the workers rather than by the diagonal chain.
//
// Flags done[i][j]: 1 = handed to workgroup 0, 2 = final L_ij published.
// Producers store the tile, release at device scope, then set the flag;
// consumers acquire it.  Deadlock freedom: all workgroups are resident (grid
// <= CUs); a wait at step k depends only on tiles of panels <= k whose
// producers never wait on the waiting workgroup's later work.  (Residency:
// one workgroup per CU.  The kernel is built with the whole register file
// so nothing spills: a build limited to half the file spilled three
// registers and returned wrong factors, so that configuration is not used.)  Every wait is
// bounded: on exhaustion the error word is raised, all other waits return at
// once, the kernel drains and the host reports the failure.
#include "pbx_common.h"
#include "pbx_tile.h"


using namespace pbx_tile;

namespace {

constexpr int PER_CU = 1;    // resident workgroups per compute unit
constexpr int ST_ERR = 4;    // state word: a wait ran out of budget
constexpr int ST_FLAGS = 16; // state words: T x T completion flags, then T column counters

__device__ bool wait_flag(const int* flag, int target, int* err) {
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

__device__ void publish_flag(int* flag, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void tile_of(int t, int& i, int& j) {
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    i = r;
    j = t - r * (r + 1) / 2;
}

__global__ void persist_reset_kernel(int* state, int T) {
    const int n = ST_FLAGS + T * T + T;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) state[e] = 0;
}

__global__ void persist_finish_kernel(const int* state, int* info) {
    if (threadIdx.x == 0 && state[ST_ERR] != 0) atomicCAS(info, 0, PBX_INFO_DAG_WAIT);
}

__global__ __launch_bounds__(256, 1) void potrf_persist_kernel(double* A, int ld, int T, double* work, int* info,
                                                            int* state) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double sd[T64];
    int* err = state + ST_ERR;
    int* done = state + ST_FLAGS;
    int* colcnt = done + T * T;  // L tiles published below the diagonal, per column
    const int W = gridDim.x - 1;
    const int ntiles = T * (T + 1) / 2;

    if (blockIdx.x == 0) {
        // ---- diagonal chain
        for (int k = 0; k < T; ++k) {
            double* Akk = A + (size_t)k * T64 * ld + (size_t)k * T64;
            if (k == 0) {
                stage<false>(sB, Akk, ld);
                __syncthreads();
            } else {
                if (!wait_flag(done + k * T + k - 1, 1, err)) return;
                if (!wait_flag(done + k * T + k, 1, err)) return;
                double* Akm = A + (size_t)k * T64 * ld + (size_t)(k - 1) * T64;
                // L_{k,k-1} = A_{k,k-1} W_{k-1}^T
                stage<true>(sA, Akm, ld);
                stage<true>(sB, work + (size_t)(k - 1) * T64 * T64, T64);
                __syncthreads();
                v4d Lc[2][2];
                acc_zero(Lc);
                mma64(sA, sB, Lc);
                double cv[16];
                acc_load(Akk, ld, cv);
                __syncthreads();  // sA, sB consumed
                acc_for_each(Lc, [&](int r, int c, double v) {
                    Akm[(size_t)r * ld + c] = v;
                    sA[c * SP + r] = v;  // op(A)(r, kk) = L[r][kk]
                });
                publish_flag(done + k * T + k - 1, 2);  // includes the barrier for sA
                if (threadIdx.x == 0)
                    __hip_atomic_fetch_add(colcnt + k - 1, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                // A_kk - L L^T into sB (row-major) for the factorisation
                v4d S[2][2];
                acc_zero(S);
                mma64(sA, sA, S);
                __syncthreads();
                int q = 0;
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
#pragma unroll
                        for (int v = 0; v < 4; ++v, ++q) sB[acc_row(x, v) * SP + acc_col(y)] = cv[q] - S[x][y][v];
                __syncthreads();
            }
            tile_factor_ws(sB, sA, sd, k * T64, info);
            // store_diag: lower L_kk into A, inverse into work
            double* Wk = work + (size_t)k * T64 * T64;
            for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
                const int r = e >> 6, c = e & 63;
                if (c <= r) Akk[(size_t)r * ld + c] = sB[r * SP + c];
                Wk[e] = sA[r * SP + c];
            }
            publish_flag(done + k * T + k, 2);
            __syncthreads();
        }
        return;
    }

    // ---- workers: tiles t = 1 + w, 1 + w + W, ... (row-major lower order)
    const int w = blockIdx.x - 1;
    int kend = -1;
    for (int t = 1 + w; t < ntiles; t += W) {
        int i, j;
        tile_of(t, i, j);
        const int nu = (i == j) ? j - 1 : j;  // panels applied by the owner
        if (nu > kend) kend = nu;
    }
    for (int k = 0; k <= kend; ++k) {
        // phase 1: finish tiles whose updates are complete
        bool have_w = false;
        for (int t = 1 + w; t < ntiles; t += W) {
            int i, j;
            tile_of(t, i, j);
            const int nu = (i == j) ? j - 1 : j;
            if (nu != k) continue;
            if (i <= j + 1) {
                if (k == 0) publish_flag(done + i * T + j, 1);  // nothing to apply
                continue;
            }
            // L_ij = A_ij W_j^T  (A_ij holds every update by now)
            if (!have_w && !wait_flag(done + j * T + j, 2, err)) return;
            have_w = true;
            double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
            stage<true>(sA, Aij, ld);
            stage<true>(sB, work + (size_t)j * T64 * T64, T64);
            __syncthreads();
            v4d out[2][2];
            acc_zero(out);
            mma64(sA, sB, out);
            acc_for_each(out, [&](int r, int c, double v) { Aij[(size_t)r * ld + c] = v; });
            publish_flag(done + i * T + j, 2);
            if (threadIdx.x == 0) __hip_atomic_fetch_add(colcnt + j, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        // phase 2: A_ij -= L_ik L_jk^T on the remaining tiles, in row order,
        // once every L tile of column k is published (one wait per panel)
        bool have_col = false;
        for (int t = 1 + w; t < ntiles; t += W) {
            int i, j;
            tile_of(t, i, j);
            const int nu = (i == j) ? j - 1 : j;
            if (k >= nu) continue;
            if (!have_col && !wait_flag(colcnt + k, T - 1 - k, err)) return;
            have_col = true;
            double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
            double cv[16];
            acc_load(Aij, ld, cv);
            stage<true>(sA, A + (size_t)i * T64 * ld + (size_t)k * T64, ld);
            if (i != j) stage<true>(sB, A + (size_t)j * T64 * ld + (size_t)k * T64, ld);
            __syncthreads();
            v4d S[2][2];
            acc_zero(S);
            mma64(sA, i == j ? sA : sB, S);
            int q = 0;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v, ++q) Aij[(size_t)acc_row(x, v) * ld + acc_col(y)] = cv[q] - S[x][y][v];
            if (k + 1 == nu && i <= j + 1)
                publish_flag(done + i * T + j, 1);  // last update in: hand to workgroup 0
            else
                __syncthreads();  // sA, sB free for the next tile
        }
    }
}

}  // namespace

// Whether a factorisation of order np can run persistently on `cus` compute
// units (at least one worker beside the diagonal workgroup).
bool pbx_persist_fits(int np, int cus) {
    const int T = np / T64;
    return T <= 1 || PER_CU * cus >= 2;
}

// state: ST_FLAGS + T*T ints (the single-launch state block of the work area).
int pbx_potrf_persist(double* A, int np, int ld, double* work, int* info_dev, int* state, int cus, hipStream_t s) {
    const int T = np / T64;
    const int ntiles = T * (T + 1) / 2;
    const int slots = PER_CU * cus;
    const int G = T <= 1 ? 1 : (slots < ntiles ? slots : ntiles);

    hipLaunchKernelGGL(persist_reset_kernel, dim3(8), dim3(256), 0, s, state, T);
    PBX_HIP(hipGetLastError());
    hipLaunchKernelGGL(potrf_persist_kernel, dim3(G), dim3(256), 0, s, A, ld, T, work, info_dev, state);
    PBX_HIP(hipGetLastError());
    hipLaunchKernelGGL(persist_finish_kernel, dim3(1), dim3(64), 0, s, state, info_dev);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// Single-launch tiled Cholesky on gfx950: one persistent kernel walks the
// lower 64 x 64 tiles of the matrix as a dependency graph (left-looking per
// tile), so diagonal factorisations overlap the trailing products instead of
// being separated by ~2T kernel launches.
//
// Task order: tiles are numbered column by column, (j,j), (j+1,j), ...,
// (T-1,j), then column j+1.  Workgroups take task numbers from an atomic
// ticket.  Task (i,j) needs tiles (i,k) and (j,k) for k < j and, for i > j,
// the diagonal tile (j,j): every dependency has a smaller ticket and was taken
// by a workgroup that is already running, so the graph always drains.
//   (i,j): C = A_ij - sum_{k<j} L_ik L_jk^T      (FP64 MFMA, k in order)
//          i == j: factor C in LDS -> L_jj, inv(L_jj)
//          i >  j: L_ij = C inv(L_jj)^T
//          publish done[i][j]
// Completion flags use device-scope release/acquire so tiles written through
// another XCD's L2 are visible.  Every spin is bounded: a wait that exceeds
// its budget raises the error word, after which all remaining waits return
// at once and the kernel drains; the host then reports the failure.
#include "pbx_common.h"
#include "pbx_tile.h"

using namespace pbx_tile;

namespace {

struct DagState {
    int ticket;
    int err;
    int pad[14];
    // followed by T*T done flags
};

// Wait until *flag != 0 (thread 0 polls), bounded by `limit` polls; returns
// false if the wait failed or an error was raised anywhere.
__device__ bool wait_done(const int* flag, int* err, long long limit) {
    __shared__ int ok;
    if (threadIdx.x == 0) {
        long long spins = 0;
        int v = 0, bad = 0;
        while (true) {
            v = __hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            if (v != 0) break;
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                bad = 1;
                break;
            }
            if (++spins > limit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bad = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = !bad;
    }
    __syncthreads();
    return ok != 0;
}

__device__ void publish(int* flag) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Turns an exhausted dependency wait into the factorisation status word.
__global__ void dag_finish_kernel(const int* state, int* info) {
    if (threadIdx.x == 0 && ((const DagState*)state)->err != 0) atomicCAS(info, 0, PBX_INFO_DAG_WAIT);
}

__global__ void dag_reset_kernel(int* state, int T) {
    int n = 16 + T * T;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) state[e] = 0;
}

__global__ __launch_bounds__(256, 2) void chol_dag_kernel(double* A, int ld, int T, double* work, int* info,
                                                       int* state, long long spin_limit) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double sd[T64];
    __shared__ int task;
    DagState* st = (DagState*)state;
    int* done = state + 16;
    const int total = T * (T + 1) / 2;
    while (true) {
        if (threadIdx.x == 0) task = atomicAdd(&st->ticket, 1);
        __syncthreads();
        const int t = task;
        __syncthreads();
        if (t >= total) break;
        // decode column-major lower task index
        int j = 0, base = 0;
        while (base + (T - j) <= t) {
            base += T - j;
            ++j;
        }
        const int i = j + (t - base);
        bool ok = true;
        double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
        double cv[16];
        acc_load(Aij, ld, cv);
        v4d acc[2][2];
        acc_zero(acc);
        for (int k = 0; k < j && ok; ++k) {
            ok = wait_done(done + i * T + k, &st->err, spin_limit);
            if (ok && i != j) ok = wait_done(done + j * T + k, &st->err, spin_limit);
            if (!ok) break;
            stage<true>(sA, A + (size_t)i * T64 * ld + (size_t)k * T64, ld);
            if (i != j) stage<true>(sB, A + (size_t)j * T64 * ld + (size_t)k * T64, ld);
            __syncthreads();
            mma64(sA, i == j ? sA : sB, acc);
            __syncthreads();
        }
        if (!ok) continue;  // error raised: drain the remaining tickets
        if (i == j) {
            int q = 0;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v, ++q) sA[acc_row(x, v) * SP + acc_col(y)] = cv[q] - acc[x][y][v];
            __syncthreads();
            tile_factor(sA, sB, sd, j * T64, info);
            double* Wj = work + (size_t)j * T64 * T64;
            for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
                int r = e >> 6, c = e & 63;
                if (c <= r) Aij[(size_t)r * ld + c] = sA[r * SP + c];
                Wj[e] = sB[r * SP + c];
            }
            publish(done + j * T + j);
        } else {
            // C -> operand s[k][r] = C[r][k]
            int q = 0;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v, ++q) sA[acc_col(y) * SP + acc_row(x, v)] = cv[q] - acc[x][y][v];
            ok = wait_done(done + j * T + j, &st->err, spin_limit);
            if (!ok) continue;
            stage<true>(sB, work + (size_t)j * T64 * T64, T64);  // op(B)(kk,c) = W_j[c][kk]
            __syncthreads();
            v4d out[2][2];
            acc_zero(out);
            mma64(sA, sB, out);
            acc_for_each(out, [&](int r, int c, double v) { Aij[(size_t)r * ld + c] = v; });
            publish(done + i * T + j);
        }
        __syncthreads();
    }
}

}  // namespace

// State block shared by the single-launch factorisations (this file and
// pbx_chol_persist.hip): 16 words, T x T tile flags, T column counters.
size_t pbx_dag_state_ints(int np) {
    int T = np / T64;
    return 16 + (size_t)T * T + T;
}

// In-place lower Cholesky (same contract as pbx_potrf) in one persistent
// launch.  `state` holds pbx_dag_state_ints(np) ints.  Returns PBX_ERR_STATE
// (after the kernel has drained) if a dependency wait ran out of budget.
int pbx_potrf_dag(double* A, int np, int ld, double* work, int* info_dev, int* state, int grid, hipStream_t s) {
    if (np % T64 != 0 || np <= 0) {
        pbx_set_error(PBX_ERR_ARG, "potrf_dag: order %d is not a positive multiple of 64", np);
        return PBX_ERR_ARG;
    }
    const int T = np / T64;
    const int total = T * (T + 1) / 2;
    if (grid <= 0) grid = 512;
    if (grid > total) grid = total;
    hipLaunchKernelGGL(dag_reset_kernel, dim3(8), dim3(256), 0, s, state, T);
    // each poll is an uncached atomic load plus a short sleep (~1 us); 4e6
    // polls bound a failed wait to a few seconds, far beyond any legitimate
    // dependency wait (a whole factorisation takes milliseconds)
    const long long spin_limit = 4000000LL;
    hipLaunchKernelGGL(chol_dag_kernel, dim3(grid), dim3(256), 0, s, A, ld, T, work, info_dev, state, spin_limit);
    hipLaunchKernelGGL(dag_finish_kernel, dim3(1), dim3(64), 0, s, state, info_dev);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

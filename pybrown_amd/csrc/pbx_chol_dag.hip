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

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int T64 = 64;
constexpr int SP = 65;

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void stage_rk(double* s, const double* X, int ld) {
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        int row = e >> 6, col = e & 63;
        s[col * SP + row] = X[(size_t)row * ld + col];
    }
}

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

template <class F>
__device__ __forceinline__ void acc_for_each(v4d acc[2][2], F f) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
            for (int v = 0; v < 4; ++v) f(r0 + 16 * x + lk + 4 * v, c0 + 16 * y + li, acc[x][y][v]);
}

__device__ __forceinline__ double bcast(double x, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    double h = 0.5 * d;
    double r = fma(-h * y, y, 0.5);
    y = fma(y, r, y);
    r = fma(-h * y, y, 0.5);
    y = fma(y, r, y);
    return y;
}

// Factor the 64 x 64 SPD tile in LDS and form inv(L) in sW (lower, zero
// above); 8-column panels on wave 0, MFMA trailing updates, inverse rows on
// waves 1-3 (same algorithm as the launch-based path's v2 tile routine).
__device__ void tile_factor(double* sL, double* sW, double* sd, double (*red)[64], int col0, int* info) {
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
        if (w > 0) {
            const int t = tid - 64;
            const int c = t & 63, part = t >> 6;
            for (int r = c0; r < c0 + 8; ++r) {
                double s = 0.0;
                if (c <= r)
                    for (int k = c + part; k < r; k += 3) s -= sL[r * SP + k] * sW[k * SP + c];
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
            for (int r = 0; r < 16; ++r) {
                __builtin_amdgcn_s_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            }
        }
        if (kb < 7) {
            const int r0 = (w >> 1) * 32, cq = (w & 1) * 32;
            const int li = lane & 15, lk = lane >> 4;
            const int lo = c0 + 8;
            const bool live = (r0 + 31 >= lo) && !(cq > r0 + 31);
            v4d acc[2][2];
            for (int x = 0; x < 2; ++x)
                for (int y = 0; y < 2; ++y) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
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
            if (live)
                acc_for_each(acc, [&](int r, int c, double v) {
                    if (r >= lo && c >= lo && c <= r) sL[r * SP + c] -= v;
                });
            __syncthreads();
        }
    }
    if (tid < 64) sL[tid * SP + tid] = sd[tid];
    __syncthreads();
}

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

__global__ __launch_bounds__(256) void chol_dag_kernel(double* A, int ld, int T, double* work, int* info,
                                                       int* state, long long spin_limit) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double sd[T64];
    __shared__ double red[3][64];
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
        v4d acc[2][2];
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < j && ok; ++k) {
            ok = wait_done(done + i * T + k, &st->err, spin_limit);
            if (ok && i != j) ok = wait_done(done + j * T + k, &st->err, spin_limit);
            if (!ok) break;
            stage_rk(sA, A + (size_t)i * T64 * ld + (size_t)k * T64, ld);
            if (i != j) stage_rk(sB, A + (size_t)j * T64 * ld + (size_t)k * T64, ld);
            __syncthreads();
            mma64(sA, i == j ? sA : sB, acc);
            __syncthreads();
        }
        if (!ok) continue;  // error raised: drain the remaining tickets
        double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
        if (i == j) {
            acc_for_each(acc, [&](int r, int c, double v) { sA[r * SP + c] = Aij[(size_t)r * ld + c] - v; });
            __syncthreads();
            tile_factor(sA, sB, sd, red, j * T64, info);
            double* Wj = work + (size_t)j * T64 * T64;
            for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
                int r = e >> 6, c = e & 63;
                if (c <= r) Aij[(size_t)r * ld + c] = sA[r * SP + c];
                Wj[e] = sB[r * SP + c];
            }
            publish(done + j * T + j);
        } else {
            // C -> operand s[k][r] = C[r][k]
            acc_for_each(acc, [&](int r, int c, double v) { sA[c * SP + r] = Aij[(size_t)r * ld + c] - v; });
            ok = wait_done(done + j * T + j, &st->err, spin_limit);
            if (!ok) continue;
            stage_rk(sB, work + (size_t)j * T64 * T64, T64);  // op(B)(kk,c) = W_j[c][kk]
            __syncthreads();
            v4d out[2][2];
            for (int x = 0; x < 2; ++x)
                for (int y = 0; y < 2; ++y) out[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
            mma64(sA, sB, out);
            acc_for_each(out, [&](int r, int c, double v) { Aij[(size_t)r * ld + c] = v; });
            publish(done + i * T + j);
        }
        __syncthreads();
    }
}

}  // namespace

size_t pbx_dag_state_ints(int np) {
    int T = np / T64;
    return 16 + (size_t)T * T;
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

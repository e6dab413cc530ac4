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
// Variant 4 folds trsm(k+1) into syrk(k): the workgroups of tile
// column k+1 keep their updated tile in registers, wait (bounded) for the
// look-ahead workgroup to publish inv(L_{k+1,k+1}), and solve in place, so
// each panel costs one launch; its diagonal tiles use the wave-specialised
// factorisation (tile_factor_ws).
// Variant 6 (default) factors two tile columns per launch (potrf2_kernel below): one
// rank-128 pass over the trailing matrix per two columns instead of two
// rank-64 passes, and half the launches.
// inv(L_kk) is produced with each diagonal factorisation and kept in `work`
// (T tiles of 64 x 64), where the SPD inverse reuses it.
#include <atomic>
#include <vector>

#include "pbx_common.h"
#include "pbx_tile.h"

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

// Look-ahead state words (variants 4 and 6), in the work area after the
// diagonal inverses (LA_WORDS ints): [LA_FLAG] = index of the last diagonal
// tile published, [LA_ERR] = a dependency wait ran out of budget; variant 6
// adds [LA_UPD] and [LA_JOB + 0..1] (below).
constexpr int LA_FLAG = 2, LA_ERR = 3, LA_WORDS = PBX_LA_INTS;

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

__host__ __device__ inline void tri_decode(int t, int& ii, int& jj) {
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    ii = r;
    jj = t - r * (r + 1) / 2;
}

// Trailing tile (ii, jj), jj <= ii < m, of the r-th trailing workgroup of a
// launch (the first tile, (0, 0), skipped when skip0).  The tiles of one
// launch are independent, so the order only decides locality.  Workgroups
// are dealt to the XCDs round-robin by index, so consecutive trailing
// workgroups cycle through `ns` XCDs (8, or 7 when XCD 0 is left to the
// diagonal chain); residue class r % ns takes one contiguous chunk of the
// tile sequence, and the sequence runs through 16-tile column bands (within
// a band, rows top-down): an XCD then keeps one band's panel rows in its L2
// while the row panels stream past once.
__host__ __device__ inline void trailing_decode(int r, int m, bool skip0, int ns, int& ii, int& jj) {
    const int off = skip0 ? 1 : 0;
    const int Nt = m * (m + 1) / 2 - off;
    const int s = r % ns, k = r / ns;
    const int base = Nt / ns, extra = Nt % ns;
    const int t = s * base + (s < extra ? s : extra) + k + off;
    int c = 0, start = 0;
    for (;;) {  // ends: t < m (m + 1) / 2, the sum of the band counts
        const int w = m - c < 16 ? m - c : 16;
        const int cnt = w * (w + 1) / 2 + (m - c - 16 > 0 ? m - c - 16 : 0) * 16;
        if (t < start + cnt) break;
        start += cnt;
        c += 16;
    }
    const int w = m - c < 16 ? m - c : 16, loc = t - start, tri = w * (w + 1) / 2;
    if (loc < tri) {
        tri_decode(loc, ii, jj);
        ii += c;
        jj += c;
    } else {
        ii = c + 16 + (loc - tri) / 16;
        jj = c + (loc - tri) % 16;
    }
}

// Number of trailing tiles (rows j .. T-1 of each column j) of the
// tile-column pairs q = q0, q0 + K, ... (columns 2q and 2q + 1 below T).
__host__ __device__ inline int far_count(int T, int q0, int K) {
    int n = 0;
    for (int q = q0; 2 * q < T; q += K)
        for (int j = 2 * q; j < 2 * q + 2 && j < T; ++j) n += T - j;
    return n;
}

// Tile (i, j) (absolute indices) of the r-th workgroup of such a slice.
// As in trailing_decode, residue class r % ns takes one contiguous chunk of
// the sequence (columns in order, rows top-down), so an XCD keeps one
// column's panel rows in its L2 while the row panels stream past.
__host__ __device__ inline void far_decode(int r, int T, int q0, int K, int ns, int& i, int& j) {
    const int Nt = far_count(T, q0, K);
    const int s = r % ns, k = r / ns;
    const int base = Nt / ns, extra = Nt % ns;
    int t = s * base + (s < extra ? s : extra) + k;
    for (int q = q0;; q += K)  // ends: t < Nt
        for (int c = 2 * q; c < 2 * q + 2 && c < T; ++c) {
            if (t < T - c) {
                j = c;
                i = c + t;
                return;
            }
            t -= T - c;
        }
}

// Tile (i, j) and lagging panels m0 .. m0+P-1 of trailing workgroup r of
// a two-column launch (potrf2_kernel; the launch's parameters as
// two_col_plan gives them), t0 = c0 + nc.  Without `near`: every trailing
// tile but (t0, t0) when nc = 2, Pt panels.  With `near`: the next launch's
// two tile columns (Pt panels; (t0, t0) belongs to row t0), the pair after
// them (Pc panels), then the far slice (Pf panels; pairs fq0, fq0 + fK, ...).
__host__ __device__ inline void two_col_trailing(int r, int T, int c0, int nc, int near, int Pt, int Pf, int fq0,
                                                 int fK, int Pc, int ns, int& i, int& j, int& m0, int& P) {
    const int m = T - c0 - nc;
    int ii, jj;
    P = Pt;
    if (near) {
        const int qc = c0 / 2 + 2, ncu = Pc > 0 ? far_count(T, qc, T) : 0;
        if (r < 2 * m - 2) {
            ii = 1 + (r < m - 1 ? r : r - (m - 1));
            jj = r < m - 1 ? 0 : 1;
        } else if (r < 2 * m - 2 + ncu) {
            far_decode(r - (2 * m - 2), T, qc, T, ns, ii, jj);
            ii -= c0 + nc;
            jj -= c0 + nc;
            P = Pc;
        } else {
            far_decode(r - (2 * m - 2) - ncu, T, fq0, fK, ns, ii, jj);
            ii -= c0 + nc;
            jj -= c0 + nc;
            P = Pf;
        }
    } else {
        trailing_decode(r, m, nc == 2, ns, ii, jj);
    }
    i = c0 + nc + ii;
    j = c0 + nc + jj;
    m0 = c0 - P;
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

// Every wave of the workgroup acquires at device scope (buffer_inv sc1).  The
// invalidation an acquire performs is ordered only before later loads of the
// wave that issues it; a workgroup barrier does not wait for it, so one
// wave's acquire does not cover the loads the other waves issue after the
// barrier.
__device__ __forceinline__ void acquire_all_waves() {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Wait until *flag >= target (thread 0 polls with device-scope acquire; then
// every wave acquires for itself), bounded; false if the budget ran out or
// another wait already failed.  An exhausted budget sets *err (so the other
// waits stop) and makes the status word PBX_INFO_DAG_WAIT, also over a pivot
// status that the unfinished factor may raise after it or has raised before
// it.
__device__ bool wait_ge_agent(const int* flag, int target, int* err, int* info) {
    __shared__ int ok;
    if (threadIdx.x == 0) {
        int good = 1;
        long long spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > (1 << 22)) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                atomicExch(info, PBX_INFO_DAG_WAIT);
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = good;
    }
    __syncthreads();
    acquire_all_waves();
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
    if (!wait_ge_agent(la + LA_FLAG, k + 1, la + LA_ERR, info)) return;
    stage<true>(sB, work + (size_t)(k + 1) * T64 * T64, T64);  // op(B)(kk, c) = W[c][kk]
    __syncthreads();
    v4d out[2][2];
    acc_zero(out);
    mma64(sA, sB, out);
    acc_for_each(out, [&](int r, int c, double v) { Aij[(size_t)r * ld + c] = v; });
}


// ---------------------------------------------------------------------------
// Variant 6: two tile columns per launch.  Launch (c0, P, nc) applies panels
// c0-P .. c0-1 (P <= 2: one rank-64P update per tile, so the trailing matrix
// is read and written once per two columns) and factors columns c0 .. c0+nc-1
// (nc <= 2).  Workgroups, in dispatch order:
//   0         tile (c0, c0), which arrives fully updated: factor, publish
//             inv(L) ([LA_FLAG] = c0+1); then L_{c0+1,c0} = C W^T, tile
//             (c0+1, c0+1) less L_{c0+1,c0} L_{c0+1,c0}^T, factor, publish
//             ([LA_FLAG] = c0+2).  L tiles are stored after each publish:
//             only the inverses are waited for;
//   1         (nc = 2) job 0: tile (c0+1, c0) through the panels, written to
//             a scratch tile (its place in A receives L_{c0+1,c0} from
//             workgroup 0 while the rows still read the updated value);
//   2         (nc = 2, P > 0) job 1: tile (c0+1, c0+1) through the panels;
//             each job is done by whoever claims it first ([LA_JOB + job]):
//             its own workgroup, or workgroup 0 when it finds the job
//             unclaimed after its first factor; both counted in [LA_UPD];
//   row i     (i >= c0+nc) tiles (i, c0) and (i, c0+1): update, then
//             L_{i,c0} = C W_{c0}^T, its own copy of L_{c0+1,c0},
//             C_{i,c0+1} -= L_{i,c0} L_{c0+1,c0}^T, L_{i,c0+1} = C W_{c0+1}^T.
//             Row c0+2 also brings tile (c0+2, c0+2) fully up to date (the
//             panels, then L_{i,c0} L_{i,c0}^T and L_{i,c0+1} L_{i,c0+1}^T), so
//             the next launch's diagonal chain starts without an update;
//   the rest  trailing tiles (i, j), j >= c0+nc, except (c0+2, c0+2): the
//             panel update.
// Every wait ends, also when other launches on other streams fill the
// compute units: workgroup 0 is the first of its launch on its XCD (lower
// indices are dispatched first there) and waits only for jobs that a
// running workgroup has claimed; row workgroups wait only for workgroup 0
// and for those jobs.  (Waiting for workgroups 1 and 2 themselves could
// stall for good: they may sit undispatched behind other launches' row
// workgroups, which wait on chains that wait on their own helpers.)  Each
// wait is also bounded.
constexpr int LA_UPD = 4, LA_JOB = 5;

__device__ __forceinline__ void la_reset(int* la) {
    if (threadIdx.x >= LA_FLAG && threadIdx.x <= LA_JOB + 1) la[threadIdx.x] = 0;
}

__global__ void la_reset_kernel(int* la) { la_reset(la); }

// Claims job j of launch c0 for this workgroup (true for exactly one
// claimant per launch: the word holds the id of the last launch that
// claimed it, and ids grow from launch to launch).
__device__ __forceinline__ bool claim_job(int* la, int j, int c0) {
    __shared__ int got[2];
    if (threadIdx.x == 0)
        got[j] = __hip_atomic_exchange(la + LA_JOB + j, c0 + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != c0 + 1;
    __syncthreads();
    return got[j] != 0;
}

__device__ __forceinline__ void publish_agent(int* flag, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// S += sum_{m = m0 .. m0+P-1} L_im L_jm^T (i == j: one operand).  Other
// waves may still read sA/sB on return: synchronise before reusing them.
// LEAN: one panel's operands in flight at a time (fewer registers).  PIPE:
// the next panel's operands in flight during each product (the trailing
// tiles of the two-workgroups-per-CU build: 1.5 % at n = 12000, and no
// register spill there; the pairs below suit the one-per-CU build).
template <bool LEAN = false, bool PIPE = false>
__device__ __forceinline__ void panel_update(const double* A, int ld, int i, int j, int m0, int P, double* sA,
                                             double* sB, v4d S[2][2]) {
    const double* Ai = A + (size_t)i * T64 * ld;
    const double* Aj = A + (size_t)j * T64 * ld;
    if (!LEAN && !PIPE && P % 2 == 0) {
        // panels in pairs: both panels' operand tiles are requested before
        // the pair's first product, so their latencies overlap each other and
        // that product
        for (int m = m0; m < m0 + P; m += 2) {
            v2d a0[8], a1[8], b0[8], b1[8];
            stage_fetch(a0, Ai + (size_t)m * T64, ld);
            stage_fetch(a1, Ai + (size_t)(m + 1) * T64, ld);
            if (i != j) {
                stage_fetch(b0, Aj + (size_t)m * T64, ld);
                stage_fetch(b1, Aj + (size_t)(m + 1) * T64, ld);
            }
            if (m > m0) __syncthreads();
            stage_put<true>(sA, a0);
            if (i != j) stage_put<true>(sB, b0);
            __syncthreads();
            mma64(sA, i == j ? sA : sB, S);
            __syncthreads();
            stage_put<true>(sA, a1);
            if (i != j) stage_put<true>(sB, b1);
            __syncthreads();
            mma64(sA, i == j ? sA : sB, S);
        }
        return;
    }
    if (PIPE) {
        // the next panel's operand tiles are requested before this panel's
        // product, so their latency overlaps it
        v2d fa[8], fb[8];
        stage_fetch(fa, Ai + (size_t)m0 * T64, ld);
        if (i != j) stage_fetch(fb, Aj + (size_t)m0 * T64, ld);
        for (int m = m0; m < m0 + P; ++m) {
            if (m > m0) __syncthreads();
            stage_put<true>(sA, fa);
            if (i != j) stage_put<true>(sB, fb);
            __syncthreads();
            if (m + 1 < m0 + P) {
                stage_fetch(fa, Ai + (size_t)(m + 1) * T64, ld);
                if (i != j) stage_fetch(fb, Aj + (size_t)(m + 1) * T64, ld);
            }
            mma64(sA, i == j ? sA : sB, S);
        }
        return;
    }
    for (int m = m0; m < m0 + P; ++m) {
        if (m > m0) __syncthreads();
        stage<true>(sA, Ai + (size_t)m * T64, ld);
        if (i != j) stage<true>(sB, Aj + (size_t)m * T64, ld);
        __syncthreads();
        mma64(sA, i == j ? sA : sB, S);
    }
}

// FP32 bulk: S += sum_m L_im L_jm^T with the operands rounded to FP32 in LDS
// (the FP64 sA/sB buffers hold the FP32 tiles).
template <bool LEAN = false>
__device__ __forceinline__ void panel_update32(const double* A, int ld, int i, int j, int m0, int P, float* sA,
                                               float* sB, v4f S[2][2]) {
    const double* Ai = A + (size_t)i * T64 * ld;
    const double* Aj = A + (size_t)j * T64 * ld;
    if (!LEAN && P % 2 == 0) {  // operand tiles of a pair of panels in flight at once, as in panel_update
        for (int m = m0; m < m0 + P; m += 2) {
            v2d a0[8], a1[8], b0[8], b1[8];
            stage_fetch(a0, Ai + (size_t)m * T64, ld);
            stage_fetch(a1, Ai + (size_t)(m + 1) * T64, ld);
            if (i != j) {
                stage_fetch(b0, Aj + (size_t)m * T64, ld);
                stage_fetch(b1, Aj + (size_t)(m + 1) * T64, ld);
            }
            if (m > m0) __syncthreads();
            stage_put_f32t(sA, a0);
            if (i != j) stage_put_f32t(sB, b0);
            __syncthreads();
            mma64_f32(sA, i == j ? sA : sB, S);
            __syncthreads();
            stage_put_f32t(sA, a1);
            if (i != j) stage_put_f32t(sB, b1);
            __syncthreads();
            mma64_f32(sA, i == j ? sA : sB, S);
        }
        return;
    }
    for (int m = m0; m < m0 + P; ++m) {
        if (m > m0) __syncthreads();
        stage_f32t(sA, Ai + (size_t)m * T64, ld);
        if (i != j) stage_f32t(sB, Aj + (size_t)m * T64, ld);
        __syncthreads();
        mma64_f32(sA, i == j ? sA : sB, S);
    }
}

// Tile (i, j) less panels m0 .. m0+P-1, FP64 or with the FP32 bulk (MIXED),
// at this thread's FP64 accumulator positions (acc_row, acc_col), so stores
// and later loads of the tile stay with the thread that owns each value.
// The FP32 accumulators hold other rows (acc32_row) and are re-laid through
// LDS.  Ends block-synchronised when the FP32 path ran; otherwise other waves
// may still read sA/sB.
struct Updated {
    double v[16];
    __device__ __forceinline__ int row(int q) const { return acc_row(q >> 3, q & 3); }
    __device__ __forceinline__ int col(int q) const { return acc_col((q >> 2) & 1); }
};

template <bool MIXED, bool LEAN = false, bool PIPE = false>
__device__ __forceinline__ void update_tile(const double* A, int ld, int i, int j, int m0, int P, double* sA,
                                            double* sB, Updated& u) {
    const double* Aij = A + (size_t)i * T64 * ld + (size_t)j * T64;
    if (MIXED && P > 0) {
        // the tile in FP64, less FP32 sums over at most two panels (128
        // deep) each, so the accumulation error does not grow with the
        // number of lagging panels a launch applies (the tile is the running
        // FP64 value: no separate copy held through the products)
        v4d cur[2][2];
        {
            double cv[16];
            acc32_load(Aij, ld, cv);
            int q = 0;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v, ++q) cur[x][y][v] = cv[q];
        }
        for (int m = m0; m < m0 + P; m += 2) {
            const int pp = (m0 + P - m) < 2 ? 1 : 2;
            if (m > m0) __syncthreads();
            v4f S[2][2];
            acc32_zero(S);
            panel_update32<LEAN>(A, ld, i, j, m, pp, (float*)sA, (float*)sB, S);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v) cur[x][y][v] -= (double)S[x][y][v];
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v) sA[acc32_row(x, v) * SP + acc_col(y)] = cur[x][y][v];
        __syncthreads();
#pragma unroll
        for (int q2 = 0; q2 < 16; ++q2) u.v[q2] = sA[u.row(q2) * SP + u.col(q2)];
        __syncthreads();
        return;
    }
    acc_load(Aij, ld, u.v);
    v4d S[2][2];
    acc_zero(S);
    if (P > 0) panel_update<LEAN, PIPE>(A, ld, i, j, m0, P, sA, sB, S);
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v, ++q) u.v[q] -= S[x][y][v];
}

// X = the updated tile (lower part only when `lower`)
__device__ __forceinline__ void store_updated(double* X, int ld, const Updated& u, bool lower) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = u.row(q), c = u.col(q);
        if (!lower || c <= r) X[(size_t)r * ld + c] = u.v[q];
    }
}

// s[c][r] = the updated tile (op(A) layout)
__device__ __forceinline__ void put_updated_t(double* s, const Updated& u) {
#pragma unroll
    for (int q = 0; q < 16; ++q) s[u.col(q) * SP + u.row(q)] = u.v[q];
}

// s[c][r] = cv - S at this thread's accumulator positions (op(A) layout)
__device__ __forceinline__ void put_diff_t(double* s, const double cv[16], v4d S[2][2]) {
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v, ++q) s[acc_col(y) * SP + acc_row(x, v)] = cv[q] - S[x][y][v];
}

// s[r][c] = cv - S (row-major tile for the factorisation)
__device__ __forceinline__ void put_diff(double* s, const double cv[16], v4d S[2][2]) {
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v, ++q) s[acc_row(x, v) * SP + acc_col(y)] = cv[q] - S[x][y][v];
}

// s[c][r] = X[r][c] for an accumulator tile (op(A) layout)
__device__ __forceinline__ void put_t(double* s, v4d X[2][2]) {
    acc_for_each(X, [&](int r, int c, double v) { s[c * SP + r] = v; });
}

// X = cv - S in global memory (lower part only when `lower`)
__device__ __forceinline__ void store_diff(double* X, int ld, const double cv[16], v4d S[2][2], bool lower) {
    int q = 0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v, ++q) {
                const int r = acc_row(x, v), c = acc_col(y);
                if (!lower || c <= r) X[(size_t)r * ld + c] = cv[q] - S[x][y][v];
            }
}

// Lower part of the diagonal tile X -= op(A) op(A)^T, op(A) in sA (block
// synchronised by the caller); the load of X overlaps the product.
__device__ __forceinline__ void syrk_store(double* X, int ld, const double* sA) {
    double cv[16];
    acc_load(X, ld, cv);
    v4d S[2][2];
    acc_zero(S);
    mma64(sA, sA, S);
    store_diff(X, ld, cv, S, true);
}

// Device-scope (write-through) store: the diagonal chain's stores then do
// not sit dirty in its XCD's L2 for its next release to write back.
__device__ __forceinline__ void st_dev(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_w_dev(const double* sW, double* Wk) {
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) st_dev(Wk + e, sW[(e >> 6) * SP + (e & 63)]);
}
__device__ __forceinline__ void store_l_dev(const double* sL, double* Akk, int ld) {
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        const int r = e >> 6, c = e & 63;
        if (c <= r) st_dev(Akk + (size_t)r * ld + c, sL[r * SP + c]);
    }
}

__device__ __forceinline__ void store_w(const double* sW, double* Wk) {
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) Wk[e] = sW[(e >> 6) * SP + (e & 63)];
}

__device__ __forceinline__ void store_l(const double* sL, double* Akk, int ld) {
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        const int r = e >> 6, c = e & 63;
        if (c <= r) Akk[(size_t)r * ld + c] = sL[r * SP + c];
    }
}

#ifdef PBX_CHAIN_STAMPS
// Diagnostic build only: wall-clock stamps of the two-column launches
// ([launch][event]), read back with pbx_debug_chain_stamps.
__device__ long long g_chain_ts[128][12];
#define CHAIN_TS(c0, i) \
    do { \
        if (threadIdx.x == 0) g_chain_ts[(c0) / 2][i] = wall_clock64(); \
    } while (0)
#define CHAIN_TV(c0, i, v) \
    do { \
        if (threadIdx.x == 0) g_chain_ts[(c0) / 2][i] = (v); \
    } while (0)
extern "C" int pbx_debug_chain_stamps(long long* out) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_chain_ts), sizeof(g_chain_ts)) == hipSuccess ? 0 : -1;
}
#else
#define CHAIN_TS(c0, i) \
    do { \
    } while (0)
#define CHAIN_TV(c0, i, v) \
    do { \
    } while (0)
#endif

// Job j of launch (c0, P): tile (c0+1, c0+j) through the panels, job 0 to
// the scratch tile, job 1 in place; then counted in [LA_UPD].  sA and sB are
// free again on return.
template <bool MIXED, bool LEAN>
__device__ __forceinline__ void job(double* A, int ld, int c0, int P, int j, double* scr, double* sA, double* sB,
                                    int* la) {
    Updated u;
    update_tile<MIXED, LEAN>(A, ld, c0 + 1, c0 + j, c0 - P, P, sA, sB, u);
    if (j == 0)
        store_updated(scr, T64, u, false);
    else
        store_updated(A + (size_t)(c0 + 1) * T64 * ld + (size_t)(c0 + 1) * T64, ld, u, true);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(la + LA_UPD, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    CHAIN_TS(c0, 8 + j);
}

// Products with a lower-triangular factor and lower halves, one 16 x 16
// block per MFMA chain in mma64's order (k ascending in steps of 4), so each
// block equals mma64's to the bit (the terms it skips are exact zeros), at
// 40 (of 64) and 48 (of 64) MFMAs per wave.  Operands are loaded before the
// chains, so no LDS latency sits inside them.
//
// The ten lower 16 x 16 blocks (X, Y) of a 64 x 64 tile, 3/3/2/2 per wave:
// 0: (0,0) (3,0) (3,3); 1: (1,0) (1,1) (3,1); 2: (2,0) (2,1); 3: (2,2) (3,2).
struct LowerBlocks {
    int n, X[3], Y[3];
};
__device__ __forceinline__ LowerBlocks lower_blocks(int w) {
    LowerBlocks lb;
    lb.n = w < 2 ? 3 : 2;
    lb.X[0] = w < 2 ? w : 2;
    lb.Y[0] = w == 3 ? 2 : 0;
    lb.X[1] = (w == 0 || w == 3) ? 3 : w;
    lb.Y[1] = w == 0 ? 0 : (w == 3 ? 2 : 1);
    lb.X[2] = 3;
    lb.Y[2] = w == 0 ? 3 : 1;
    return lb;
}

// acc[b] = block (w, b) of X B^T, B lower triangular (block b sums over
// k = 0 .. 16b + 15): sXop holds X in operand layout (s[k * SP + r] =
// X[r][k]), sB holds B row-major.  Register v of acc[b] is element
// (16w + lk + 4v, 16b + li).
__device__ __forceinline__ void tri_rowblock(const double* sXop, const double* sB, v4d acc[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    double a[16], bw[40];
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) a[ks] = sXop[(4 * ks + lk) * SP + 16 * w + li];
#pragma unroll
    for (int b = 0, i = 0; b < 4; ++b)
#pragma unroll
        for (int ks = 0; ks < 4 * b + 4; ++ks, ++i) bw[i] = sB[(16 * b + li) * SP + 4 * ks + lk];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (ks < 4 * b + 4) acc[b] = mfma(a[ks], bw[2 * b * (b + 1) + ks], acc[b]);
}

// s[k * SP + r] = X[r][k] for the row block tri_rowblock produced
__device__ __forceinline__ void put_rowblock_t(double* s, const v4d acc[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) s[(16 * b + li) * SP + 16 * w + lk + 4 * q] = acc[b][q];
}

// S[s] = block (lb.X[s], lb.Y[s]) of X X^T over k = 0 .. 63, sXop = X in
// operand layout
__device__ __forceinline__ void lower_syrk(const double* sXop, const LowerBlocks& lb, v4d S[3]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    double pa[3][16], pb[3][16];
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
        const double* pk = sXop + (4 * ks + lk) * SP;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            pa[s][ks] = (s < lb.n) ? pk[16 * lb.X[s] + li] : 0.0;
            pb[s][ks] = (s < lb.n) ? pk[16 * lb.Y[s] + li] : 0.0;
        }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < 3; ++s) S[s] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks)
#pragma unroll
        for (int s = 0; s < 3; ++s)
            if (s < lb.n) S[s] = mfma(pa[s][ks], pb[s][ks], S[s]);
}

// Between the two tile factorisations of a two-column launch, after the
// fused factorisation of tile (c0, c0) (W = inv(L) in sW, row-major; C =
// tile (c0+1, c0) staged in sC in operand layout and T = tile (c0+1, c0+1)
// in sT row-major, all through the lagging panels):
//   Lo = C W^T  (= L_{c0+1,c0}), tri_rowblock;
//   W of tile (c0, c0) released to the row workgroups (flag = v): wave 1
//               wrote it during the factorisation;
//   Lo also to sLo (row-major, may be sW): the caller stores it to A after
//               its second factorisation (no workgroup of the launch reads
//               it);
//   sC = T - Lo Lo^T over the ten lower blocks, row-major (C is dead by
//               then): the next tile to factor.  Lo passes through sT.
// Lo and the new tile equal the sequential path's (mma64, put_diff) bits.
__device__ __forceinline__ void fused_between(const double* sW, double* sC, double* sT, double* sLo, int* flag,
                                              int v) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    const LowerBlocks lb = lower_blocks(w);
    double cv[3][4];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cv[s][q] = (s < lb.n) ? sT[(16 * lb.X[s] + lk + 4 * q) * SP + 16 * lb.Y[s] + li] : 0.0;
    __syncthreads();
    v4d acc[4];
    tri_rowblock(sC, sW, acc);
    put_rowblock_t(sT, acc);
    __syncthreads();
    // every wave has read its W operands: sLo (which may be sW) takes Lo
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) sLo[(16 * w + lk + 4 * q) * SP + 16 * b + li] = acc[b][q];
    v4d S[3];
    lower_syrk(sT, lb, S);
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        if (s >= lb.n) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) sC[(16 * lb.X[s] + lk + 4 * q) * SP + 16 * lb.Y[s] + li] = cv[s][q] - S[s][q];
    }
    // W was written by wave 1 as it formed it: release it
    publish_agent(flag, v);
}

template <int V, int OCC, bool MIXED>
__global__ __launch_bounds__(256, OCC) void potrf2_kernel(double* A, int ld, int T, int c0, int P, int nc,
                                                        int upd_target, double* work, double* scr, int* info,
                                                        int* la, int chain_jobs, int Pt, int near, int Pf, int fq0,
                                                        int fK, int Pc) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double sd[T64];
    const int nh = nc == 2 ? (P > 0 ? 2 : 1) : 0;
    const int nw = T - c0 - nc;
    // One-workgroup-per-CU build: workgroups are dealt to the XCDs
    // round-robin, and those that would share XCD 0 with workgroup 0 leave at
    // once (the others take roles 1, 2, ... in order), so the diagonal
    // chain's agent-scope releases (which write back its XCD's L2) and
    // acquires (which invalidate it) meet its own data only.
    int b = blockIdx.x;
    if (OCC == 1 && b != 0) {
        if ((b & 7) == 0) return;
        b -= b >> 3;
    }
    auto tile = [&](int i, int j) { return A + (size_t)i * T64 * ld + (size_t)j * T64; };
    int* err = la + LA_ERR;
    double* W0 = work + (size_t)c0 * T64 * T64;
    double* W1 = W0 + T64 * T64;

    if (b == 0) {
        CHAIN_TS(c0, 0);
        stage<false>(sA, tile(c0, c0), ld);
        __syncthreads();
        if (OCC == 1 && V == 5 && nc == 2) {
            // One-workgroup-per-CU build: waves 2 and 3 stage tiles (c0+1, c0)
            // and (c0+1, c0+1) during the factorisation once the helper jobs
            // have published them, and the products for the second column
            // follow at once (fused_between).  If the jobs were late, the
            // sequential path below runs instead, to the same bits.
            extern __shared__ double fuse_lds[];
            double* sC = fuse_lds;
            double* sT = fuse_lds + T64 * SP;
            const FuseCtx fx{scr, T64, tile(c0 + 1, c0 + 1), ld, la + LA_UPD, upd_target, sC, sT};
#ifdef PBX_NO_PROGRESSIVE_W
            double* const W0p = nullptr;
            double* const W1p = nullptr;
#else
            double* const W0p = W0;
            double* const W1p = W1;
#endif
            bool fused = tile_factor_ws<true>(sA, sB, sd, c0 * T64, info, &fx, W0p);
            CHAIN_TS(c0, 1);
            CHAIN_TV(c0, 2, fused ? 1 : 0);
            if (!fused) {
                // staged late: now, if the jobs are done (else the
                // sequential path, which also does unclaimed jobs)
                __shared__ int jobs_done;
                if (threadIdx.x == 0)
                    jobs_done = __hip_atomic_load(la + LA_UPD, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= upd_target;
                __syncthreads();
                if (jobs_done) {
                    acquire_all_waves();
                    stage<true>(sC, scr, T64);
                    stage<false>(sT, tile(c0 + 1, c0 + 1), ld);
                    __syncthreads();
                    fused = true;
                }
            }
            if (fused) {
                if (!W0p) store_w_dev(sB, W0);
                // W0 is in global memory once released: sB then holds Lo
                fused_between(sB, sC, sT, sB, la + LA_FLAG, c0 + 1);
                CHAIN_TS(c0, 3);
                tile_factor_ws(sC, sT, sd, (c0 + 1) * T64, info, nullptr, W1p);
                CHAIN_TS(c0, 4);
                if (!W1p) store_w_dev(sT, W1);
                publish_agent(la + LA_FLAG, c0 + 2);
                CHAIN_TS(c0, 5);
                // the factors themselves: read by no workgroup of this launch
                store_l_dev(sA, tile(c0, c0), ld);
                for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x)
                    st_dev(tile(c0 + 1, c0) + (size_t)(e >> 6) * ld + (e & 63), sB[(e >> 6) * SP + (e & 63)]);
                store_l_dev(sC, tile(c0 + 1, c0 + 1), ld);
                return;
            }
        } else {
            factor_tile<V>(sA, sB, sd, c0 * T64, info);
        }
        store_w_dev(sB, W0);
        publish_agent(la + LA_FLAG, c0 + 1);
        store_l_dev(sA, tile(c0, c0), ld);
        if (nc == 1) return;
        // Usually the jobs are done by now (one acquire load sees it);
        // otherwise the unclaimed ones (their workgroups not yet running) are
        // done here, one panel at a time (a rare path; registers are kept for
        // the chain), and the claimed ones waited for.
        __shared__ int pending;
        if (threadIdx.x == 0) {
            pending = __hip_atomic_load(la + LA_UPD, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < upd_target;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        acquire_all_waves();
        bool helped = false;
        // (marked unlikely: the compiler then keeps this code out of the
        // chain's straight-line path, which is measurably faster)
        if (__builtin_expect(pending != 0, 0)) {
            for (int j = 0; j < nh; ++j)
                if (claim_job(la, j, c0)) {
                    job<MIXED, true>(A, ld, c0, P, j, scr, sA, sB, la);
                    helped = true;
                }
            if (!wait_ge_agent(la + LA_UPD, upd_target, err, info)) return;
        }
        if (helped) {
            stage<true>(sB, W0, T64);  // op(B)(kk, c) = W[c][kk]
        } else {
            // W^T in sB (op(B)(kk, c) = W[c][kk]); W is lower triangular
            for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
                const int r = e >> 6, c = e & 63;
                if (c > r) {
                    const double t = sB[r * SP + c];
                    sB[r * SP + c] = sB[c * SP + r];
                    sB[c * SP + r] = t;
                }
            }
        }
        double* Aoff = tile(c0 + 1, c0);
        double cv[16];
        acc_load(tile(c0 + 1, c0 + 1), ld, cv);
        stage<true>(sA, scr, T64);
        __syncthreads();
        v4d Lo[2][2];
        acc_zero(Lo);
        mma64(sA, sB, Lo);
        acc_for_each(Lo, [&](int r, int c, double v) { st_dev(Aoff + (size_t)r * ld + c, v); });
        __syncthreads();
        put_t(sA, Lo);
        __syncthreads();
        v4d S[2][2];
        acc_zero(S);
        mma64(sA, sA, S);
        put_diff(sB, cv, S);
        __syncthreads();
        CHAIN_TS(c0, 3);
        factor_tile<V>(sB, sA, sd, (c0 + 1) * T64, info);
        CHAIN_TS(c0, 4);
        store_w_dev(sA, W1);
        publish_agent(la + LA_FLAG, c0 + 2);
        CHAIN_TS(c0, 5);
        store_l_dev(sB, tile(c0 + 1, c0 + 1), ld);
        return;
    }
    if (b <= nh) {
        if (!chain_jobs && claim_job(la, b - 1, c0)) job<MIXED, false>(A, ld, c0, P, b - 1, scr, sA, sB, la);
        return;
    }
    if (b <= nh + nw) {
        const int i = c0 + nc + (b - 1 - nh);
        double* A0 = tile(i, c0);
        if (nc == 1) {
            Updated u;
            update_tile<MIXED, false, OCC == 2>(A, ld, i, c0, c0 - P, P, sA, sB, u);
            if (P > 0) __syncthreads();
            put_updated_t(sA, u);
            if (!wait_ge_agent(la + LA_FLAG, c0 + 1, err, info)) return;
            stage<true>(sB, W0, T64);
            __syncthreads();
            v4d out[2][2];
            acc_zero(out);
            mma64(sA, sB, out);
            acc_for_each(out, [&](int r, int c, double v) { st_dev(A0 + (size_t)r * ld + c, v); });
            return;
        }
        double* A1 = tile(i, c0 + 1);
        double* Aii = tile(i, i);
        const bool diag = (i == c0 + 2);  // also finishes tile (i, i) for the next launch
        if (diag && Pt > 0) {  // (a trailing tile: it lags Pt panels)
            Updated u;
            update_tile<MIXED, false, OCC == 2>(A, ld, i, i, c0 - Pt, Pt, sA, sB, u);
            store_updated(Aii, ld, u, true);
            // other threads read these values back (lower-block layout)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        // tile (i, c0+1) through the panels, back in place (each thread later
        // reads back the values it wrote)
        if (P > 0) {
            Updated u;
            update_tile<MIXED, false, OCC == 2>(A, ld, i, c0 + 1, c0 - P, P, sA, sB, u);
            store_updated(A1, ld, u, false);
            __syncthreads();
        }
        // tile (i, c0): update, then L_{i,c0} = C W_{c0}^T
        v4d L0[2][2];
        {
            Updated u;
            update_tile<MIXED, false, OCC == 2>(A, ld, i, c0, c0 - P, P, sA, sB, u);
            if (P > 0) __syncthreads();
            put_updated_t(sA, u);
        }
        // (products with the triangular W through tri_rowblock: the same
        // bits as mma64 with fewer MFMAs; rows then in row-block layout)
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
        auto store_rowblock = [&](double* X, const v4d R[4]) {
#pragma unroll
            for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                for (int q = 0; q < 4; ++q) st_dev(X + (size_t)(16 * w + lk + 4 * q) * ld + 16 * bb + li, R[bb][q]);
        };
        if (!wait_ge_agent(la + LA_FLAG, c0 + 1, err, info)) return;
        stage<false>(sB, W0, T64);
        __syncthreads();
        v4d L0b[4];
        tri_rowblock(sA, sB, L0b);
        store_rowblock(A0, L0b);
        // L_{c0+1,c0} = C_{c0+1,c0} W_{c0}^T from workgroup 1's scratch tile
        if (!wait_ge_agent(la + LA_UPD, upd_target, err, info)) return;
        stage<true>(sA, scr, T64);
        __syncthreads();
        v4d Lob[4];
        tri_rowblock(sA, sB, Lob);
        __syncthreads();
        // C_{i,c0+1} -= L_{i,c0} L_{c0+1,c0}^T
        put_rowblock_t(sA, L0b);
        put_rowblock_t(sB, Lob);
        __syncthreads();
        // diag: the lower blocks of tile (i, i) less L_{i,c0} L_{i,c0}^T,
        // kept in registers (lower_blocks layout) until L_{i,c0+1} is known
        const LowerBlocks lb = lower_blocks(w);
        double dq[3][4], dqq[16];
        {
            double c1[16];
            acc_load(A1, ld, c1);
            v4d S[2][2];
            acc_zero(S);
            mma64(sA, sB, S);
            if (diag && OCC == 2) {
                // two workgroups per CU: no room for lower_syrk's operands
                double cvq[16];
                acc_load(Aii, ld, cvq);
                v4d S0[2][2];
                acc_zero(S0);
                mma64(sA, sA, S0);
                int q = 0;
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
#pragma unroll
                        for (int v = 0; v < 4; ++v, ++q) dqq[q] = cvq[q] - S0[x][y][v];
            } else if (diag) {
                double cvd[3][4];
#pragma unroll
                for (int t = 0; t < 3; ++t)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        cvd[t][q] = (t < lb.n) ? Aii[(size_t)(16 * lb.X[t] + lk + 4 * q) * ld + 16 * lb.Y[t] + li] : 0.0;
                v4d S0[3];
                lower_syrk(sA, lb, S0);
#pragma unroll
                for (int t = 0; t < 3; ++t)
#pragma unroll
                    for (int q = 0; q < 4; ++q) dq[t][q] = cvd[t][q] - S0[t][q];
            }
            __syncthreads();
            put_diff_t(sA, c1, S);
        }
        if (!wait_ge_agent(la + LA_FLAG, c0 + 2, err, info)) return;
        if (diag) CHAIN_TS(c0, 6);
        stage<false>(sB, W1, T64);
        __syncthreads();
        if (diag) CHAIN_TS(c0, 7);
        v4d L1b[4];
        tri_rowblock(sA, sB, L1b);
        store_rowblock(A1, L1b);
        if (diag) CHAIN_TS(c0, 11);
        if (diag && OCC == 2) {
            __syncthreads();
            put_rowblock_t(sA, L1b);
            __syncthreads();
            v4d S1[2][2];
            acc_zero(S1);
            mma64(sA, sA, S1);
            store_diff(Aii, ld, dqq, S1, true);
        } else if (diag) {
            __syncthreads();
            put_rowblock_t(sA, L1b);
            __syncthreads();
            v4d S1[3];
            lower_syrk(sA, lb, S1);
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t >= lb.n) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 16 * lb.X[t] + lk + 4 * q, c = 16 * lb.Y[t] + li;
                    if (c <= r) st_dev(Aii + (size_t)r * ld + c, dq[t][q] - S1[t][q]);
                }
            }
            CHAIN_TS(c0, 10);
        }
        return;
    }
    // trailing tile (i, j), j >= c0 + nc, less its lagging panels
    int i, j, m0u, Pu;
    two_col_trailing(b - 1 - nh - nw, T, c0, nc, near, Pt, Pf, fq0, fK, Pc, OCC == 1 ? 7 : 8, i, j, m0u, Pu);
    Updated u;
    update_tile<MIXED, false, OCC == 2>(A, ld, i, j, m0u, Pu, sA, sB, u);
    double* X = tile(i, j);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = u.row(q), c = u.col(q);
        if (i != j || c <= r) st_dev(X + (size_t)r * ld + c, u.v[q]);
    }
}

// ---------------------------------------------------------------------------
// Triangular inverse X = L^{-1} (lower) from a factored A and its diagonal
// inverses W_i, by recursive doubling over aligned tile blocks:
//   level b = 1, 2, 4, ...: for each block pair B1 = [s, s+b),
//   B2 = [s+b, min(s+2b, T)) (s a multiple of 2b) whose halves are already
//   inverted,  X_21 = -X_22 (L_21 X_11).
// Two launches per level: T1 = L_21 X_11 (tile (p, q) sums over k = q..b-1,
// X_11 being lower triangular), stored in X's unused upper triangle at tile
// (s+q, s+b+p); then X_21 = -X_22 T1 (k = 0..p).  About 2T tile products on
// the critical path instead of T^2/2 for a row-by-row substitution.

__global__ void copy_diag_inv_kernel(double* X, int ldx, const double* work) {
    const int i = blockIdx.x;
    double* Xii = X + (size_t)i * T64 * ldx + (size_t)i * T64;
    const double* Wi = work + (size_t)i * T64 * T64;
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        int r = e >> 6, c = e & 63;
        Xii[(size_t)r * ldx + c] = Wi[e];
    }
}

// acc += sum over k of op(A)_k op(B)_k, with the next pair of operand tiles
// fetched into registers while the current product runs.  RKA/RKB select
// the staging layout: true = the tile itself as that operand of the product
// for A (op(A) = tile), false = the tile itself for B (op(B) = tile).
template <class FA, class FB>
__device__ __forceinline__ void chain_product(int k0, int k1, FA tileA, FB tileB, int lda, int ldb, double* sA,
                                              double* sB, v4d acc[2][2]) {
    if (k0 > k1) return;
    v2d fa[8], fb[8];
    stage_fetch(fa, tileA(k0), lda);
    stage_fetch(fb, tileB(k0), ldb);
    for (int k = k0; k <= k1; ++k) {
        if (k > k0) __syncthreads();
        stage_put<true>(sA, fa);   // op(A)(r, kk) = tileA[r][kk]
        stage_put<false>(sB, fb);  // op(B)(kk, c) = tileB[kk][c]
        __syncthreads();
        if (k < k1) {
            stage_fetch(fa, tileA(k + 1), lda);
            stage_fetch(fb, tileB(k + 1), ldb);
        }
        mma64(sA, sB, acc);
    }
}

// The same on the FP32 matrix cores (mixed-precision inverse): operands
// rounded to FP32 as they are staged, each 64-deep slab accumulated in FP32
// and added into the FP64 sum (FP32 accumulator layout).
template <class FA, class FB>
__device__ __forceinline__ void chain_product_f32(int k0, int k1, FA tileA, FB tileB, int lda, int ldb, float* sA,
                                                  float* sB, v4d sum[2][2]) {
    if (k0 > k1) return;
    v2d fa[8], fb[8];
    stage_fetch(fa, tileA(k0), lda);
    stage_fetch(fb, tileB(k0), ldb);
    for (int k = k0; k <= k1; ++k) {
        if (k > k0) __syncthreads();
        stage_put_f32t(sA, fa);  // op(A) = tileA
        stage_put_f32n(sB, fb);  // op(B) = tileB
        __syncthreads();
        if (k < k1) {
            stage_fetch(fa, tileA(k + 1), lda);
            stage_fetch(fb, tileB(k + 1), ldb);
        }
        v4f acc[2][2];
        acc32_zero(acc);
        mma64_f32(sA, sB, acc);
        acc32_add_to(acc, sum);
    }
}

// One level's tile products, in FP64 or (MIXED) on the FP32 matrix cores;
// results are stored in FP64 either way.
template <bool MIXED, class FA, class FB, class FO>
__device__ __forceinline__ void level_product(int k0, int k1, FA tileA, FB tileB, int lda, int ldb, FO out) {
    if constexpr (MIXED) {
        __shared__ float fA[T64 * SPF];
        __shared__ float fB[T64 * SPF];
        v4d sum[2][2];
        acc_zero(sum);
        chain_product_f32(k0, k1, tileA, tileB, lda, ldb, fA, fB, sum);
        acc32d_for_each(sum, out);
    } else {
        __shared__ double sA[T64 * SP];
        __shared__ double sB[T64 * SP];
        v4d acc[2][2];
        acc_zero(acc);
        chain_product(k0, k1, tileA, tileB, lda, ldb, sA, sB, acc);
        acc_for_each(acc, out);
    }
}

template <int PHASE, bool MIXED>
__global__ __launch_bounds__(256) void trtri_level_kernel(const double* A, int ld, double* X, int ldx, int T, int b) {
    // Workgroups in order of decreasing chain length, so the short chains
    // fill the launch's tail: phase 0's tile (p, q) sums b - q products,
    // phase 1's p + 1.  Index = (length rank, pair, other coordinate).
    const int npairs = (T + b - 1) / (2 * b);
    const int rank = blockIdx.x / (npairs * b), rem2 = blockIdx.x - rank * npairs * b;
    const int pair = rem2 / b, o = rem2 - pair * b;
    const int p = PHASE == 0 ? o : b - 1 - rank;
    const int q = PHASE == 0 ? rank : o;
    const int s = pair * 2 * b;
    const int h2 = (s + 2 * b < T ? s + 2 * b : T) - (s + b);
    if (p >= h2) return;  // whole workgroup: B2 is shorter than b
    auto Lt = [&](int i, int j) { return A + (size_t)i * T64 * ld + (size_t)j * T64; };
    auto Xt = [&](int i, int j) { return X + (size_t)i * T64 * ldx + (size_t)j * T64; };
    if (PHASE == 0) {
        // T1(p, q) = sum_{k=q}^{b-1} L(s+b+p, s+k) X(s+k, s+q)
        double* out = Xt(s + q, s + b + p);
        level_product<MIXED>(q, b - 1, [&](int k) { return Lt(s + b + p, s + k); },
                             [&](int k) { return (const double*)Xt(s + k, s + q); }, ld, ldx,
                             [&](int r, int c, double v) { out[(size_t)r * ldx + c] = v; });
    } else {
        // X(s+b+p, s+q) = -sum_{k=0}^{p} X(s+b+p, s+b+k) T1(k, q)
        double* out = Xt(s + b + p, s + q);
        level_product<MIXED>(0, p, [&](int k) { return (const double*)Xt(s + b + p, s + b + k); },
                             [&](int k) { return (const double*)Xt(s + q, s + b + k); }, ldx, ldx,
                             [&](int r, int c, double v) { out[(size_t)r * ldx + c] = -v; });
    }
}

// The same level in FP64 over 128 x 128 blocks of tile pairs (b even):
// block (P, Q) holds the tiles p = 2P, 2P+1 and q = 2Q, 2Q+1 of one pair's
// product (rows p >= h2 absent).  32-deep slices of both operands pass
// through LDS with the next slice's loads in flight, and wave w holds the
// 64 x 64 quadrant (w/2, w%2) as 4 x 4 MFMA blocks (half the LDS reads per
// product of the 64 x 64 tiles).  A block sums over the k range of its
// longest tile; the terms outside a tile's own range have a zero operand
// there (phase 0: X(s+2Q, s+2Q+1), above the diagonal, for q = 2Q+1;
// phase 1: k = 2P+1 for p = 2P), so each tile sums its products in
// trtri_level_kernel's order: the same values.
template <int PHASE>
__global__ __launch_bounds__(256, 2) void trtri128_kernel(const double* A, int ld, double* X, int ldx, int T, int b) {
    constexpr int KC = 32, SPX = 129, NV = 128 * KC / 2 / 256;
    __shared__ double sA[KC * SPX];
    __shared__ double sB[KC * SPX];
    const int hb = b / 2, npairs = (T + b - 1) / (2 * b);
    // decreasing chain length: phase 0's block (P, Q) sums b - 2Q tiles,
    // phase 1's up to 2P + 2
    const int rank = blockIdx.x / (npairs * hb), rem2 = blockIdx.x - rank * npairs * hb;
    const int pair = rem2 / hb, o = rem2 - pair * hb;
    const int P = PHASE == 0 ? o : hb - 1 - rank;
    const int Q = PHASE == 0 ? rank : o;
    const int s = pair * 2 * b;
    const int h2 = (s + 2 * b < T ? s + 2 * b : T) - (s + b);
    if (2 * P >= h2) return;
    const int rows = 2 * P + 1 < h2 ? 128 : 64;  // rows of tile p = 2P+1 exist
    const int kt0 = PHASE == 0 ? 2 * Q : 0;
    const int kt1 = PHASE == 0 ? b - 1 : (2 * P + 1 < h2 ? 2 * P + 1 : 2 * P);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 64, c0 = (w & 1) * 64;
    v2d fa[NV], fb[NV];
    // this thread's staging coordinates: A rows rA + 16 it (the block's two
    // tile rows lie consecutively), columns kc, kc + 1 of the slice; B slice
    // rows kB + 4 it, block columns cB, cB + 1
    const int rA = threadIdx.x / (KC / 2), kc = (threadIdx.x % (KC / 2)) * 2;
    const int kB = threadIdx.x / 64, cB = (threadIdx.x % 64) * 2;
    const double* Arow = PHASE == 0 ? A + (size_t)(T64 * (s + b + 2 * P) + rA) * ld
                                    : X + (size_t)(T64 * (s + b + 2 * P) + rA) * ldx;
    const int lda = PHASE == 0 ? ld : ldx;
    const double* Bcol = PHASE == 0 ? X + (size_t)kB * ldx + T64 * (s + 2 * Q) + cB
                                    : X + (size_t)(T64 * (s + 2 * Q + (cB >> 6)) + kB) * ldx + (cB & 63);
    const bool bHigh = cB >= 64;
    // slice kk0 .. kk0+31 of k tile kt
    auto fetch = [&](int kt, int kk0) {
        const double* pa = Arow + (PHASE == 0 ? T64 * (s + kt) : T64 * (s + b + kt)) + kk0 + kc;
        const double* pb = PHASE == 0 ? Bcol + (size_t)(T64 * (s + kt) + kk0) * ldx
                                      : Bcol + (size_t)kk0 * ldx + T64 * (s + b + kt);
        const bool bLive = PHASE == 0 ? !(kt == 2 * Q && bHigh) : true;
        const int aLo = (PHASE == 1 && kt == 2 * P + 1) ? 64 : 0;  // phase 1: tile row 2P takes no k = 2P+1
#pragma unroll
        for (int it = 0; it < NV; ++it) {
            const int r = rA + 16 * it;
            fa[it] = (r < rows && r >= aLo) ? *(const v2d*)(pa + (size_t)(16 * it) * lda) : (v2d){0.0, 0.0};
            fb[it] = bLive ? *(const v2d*)(pb + (size_t)(4 * it) * ldx) : (v2d){0.0, 0.0};
        }
    };
    v4d acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int nsl = (kt1 - kt0 + 1) * (T64 / KC);  // slices, two per k tile
    fetch(kt0, 0);
    for (int t = 0; t < nsl; ++t) {
            if (t > 0) __syncthreads();
#pragma unroll
            for (int it = 0; it < NV; ++it) {
                const int r = rA + 16 * it, kk = kB + 4 * it;
                sA[kc * SPX + r] = fa[it].x;
                sA[(kc + 1) * SPX + r] = fa[it].y;
                sB[kk * SPX + cB] = fb[it].x;
                sB[kk * SPX + cB + 1] = fb[it].y;
            }
            __syncthreads();
            if (t + 1 < nsl) fetch(kt0 + (t + 1) / 2, ((t + 1) & 1) * KC);
#pragma unroll 2
            for (int k = 0; k < KC; k += 4) {
                const double* pa = sA + (k + lk) * SPX;
                const double* pb = sB + (k + lk) * SPX;
                double a[4], bb[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) a[x] = pa[r0 + 16 * x + li];
#pragma unroll
                for (int y = 0; y < 4; ++y) bb[y] = pb[c0 + 16 * y + li];
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[x][y] = mfma(a[x], bb[y], acc[x][y]);
            }
        }
    const int p = 2 * P + (r0 >> 6), q = 2 * Q + (c0 >> 6);
    if (p >= h2) return;
    double* out = PHASE == 0 ? X + (size_t)T64 * (s + q) * ldx + (size_t)T64 * (s + b + p)
                             : X + (size_t)T64 * (s + b + p) * ldx + (size_t)T64 * (s + q);
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * x + lk + 4 * v, c = 16 * y + li;
                out[(size_t)r * ldx + c] = PHASE == 0 ? acc[x][y][v] : -acc[x][y][v];
            }
}

// Factorisation residual: acc[0] += ||D - B B^T||^2, acc[1] += ||D||^2 over
// the leading n x n block (B lower; each strictly lower tile counts twice).
// One workgroup per lower tile (i, j): E_ij = D_ij - sum_{k <= j} B_ik B_jk^T.
// Diagonal tiles of B hold stale values above the diagonal (the factor is
// written over the lower triangle only), so those are zeroed once staged.
__device__ __forceinline__ void zero_upper_t(double* s) {
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        const int r = e >> 6, c = e & 63;
        if (c > r) s[c * SP + r] = 0.0;  // op layout: s[col][row]
    }
}

__global__ __launch_bounds__(256) void residual_kernel(const double* D, const double* B, int ld, int n,
                                                       double* acc) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    __shared__ double red[2][4];
    int i, j;
    tri_decode(blockIdx.x, i, j);
    v4d S[2][2];
    acc_zero(S);
    for (int k = 0; k <= j; ++k) {
        if (k > 0) __syncthreads();
        stage<true>(sA, B + (size_t)i * T64 * ld + (size_t)k * T64, ld);  // op(A) = B_ik
        stage<true>(sB, B + (size_t)j * T64 * ld + (size_t)k * T64, ld);  // op(B)(kk, c) = B_jk[c][kk]
        if (k == i || k == j) {
            __syncthreads();
            if (k == i) zero_upper_t(sA);
            if (k == j) zero_upper_t(sB);
        }
        __syncthreads();
        mma64(sA, sB, S);
    }
    const double* Dij = D + (size_t)i * T64 * ld + (size_t)j * T64;
    double e2 = 0.0, d2 = 0.0;
    acc_for_each(S, [&](int r, int c, double v) {
        if (i * T64 + r < n && j * T64 + c < n) {
            const double d = Dij[(size_t)r * ld + c];
            e2 += (d - v) * (d - v);
            d2 += d * d;
        }
    });
    const double w = (i == j) ? 1.0 : 2.0;
    for (int off = 32; off > 0; off >>= 1) {
        e2 += __shfl_down(e2, off);
        d2 += __shfl_down(d2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = e2;
        red[1][threadIdx.x >> 6] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(acc, w * (red[0][0] + red[0][1] + red[0][2] + red[0][3]));
        atomicAdd(acc + 1, w * (red[1][0] + red[1][1] + red[1][2] + red[1][3]));
    }
}

// Y = scale * X^T X, lower tiles computed, both triangles written:
// Y_ij = sum_{m >= i} X_mi^T X_mj for i >= j.  MIXED: products on the FP32
// matrix cores (operands rounded to FP32, FP32 accumulation within each
// 64-deep slab, FP64 across slabs).
// The same in FP64 over 128 x 128 blocks (tile rows 2P, 2P+1 by tile
// columns 2Q, 2Q+1, P >= Q; half blocks at an odd T's edge): Y_ij =
// scale sum_{m >= i} X_mi^T X_mj for the block's lower tiles.  Both
// operands are slabs of X's rows (X[64m + k][128P + r] and [128Q + c]),
// staged 32 rows at a time as they lie, with the next slab's loads in
// flight during each product; wave w holds the 64 x 64 quadrant (w/2, w%2)
// as 4 x 4 MFMA blocks (half the LDS reads per product of the 64 x 64
// tiles' 2 x 2).  Row m = 2P adds to tile row 2P only (X_{2P,2P+1} is
// above the diagonal: zero here), so every tile sums its products over m
// and k in the order lauum_kernel does: the same values.
__global__ __launch_bounds__(256, 2) void lauum128_kernel(const double* X, int ldx, double* Y, int ld, int T,
                                                          double scale) {
    constexpr int KC = 32, SPX = 129, NV = 128 * KC / 2 / 256;
    __shared__ double sA[KC * SPX];
    __shared__ double sB[KC * SPX];
    int P, Q;
    tri_decode(blockIdx.x, P, Q);  // P ascending: the longest sums first
    const int wA = 64 * T - 128 * P < 128 ? 64 : 128, wB = 64 * T - 128 * Q < 128 ? 64 : 128;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 64, c0 = (w & 1) * 64;
    const double* XA = X + 128 * P;
    const double* XB = X + 128 * Q;
    v2d fa[NV], fb[NV];
    // slab rows k0 .. k0+31 of the matrix (row 64m + kk): columns beyond the
    // block's width, and on row tile 2P the columns of tile 2P+1, are zero
    auto fetch = [&](int k0) {
        const int lim = k0 < 128 * P + 64 ? 64 : 128;  // row tile 2P: tile 2P+1 of A (and B when P == Q) is zero
#pragma unroll
        for (int it = 0; it < NV; ++it) {
            const int e = threadIdx.x + it * 256, kk = e / 64, col = (e % 64) * 2;
            const size_t off = (size_t)(k0 + kk) * ldx + col;
            fa[it] = (col < wA && col < lim) ? *(const v2d*)(XA + off) : (v2d){0.0, 0.0};
            fb[it] = (col < wB && (P != Q || col < lim)) ? *(const v2d*)(XB + off) : (v2d){0.0, 0.0};
        }
    };
    v4d acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int kbeg = 128 * P, kend = 64 * T;
    fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += KC) {
        if (k0 > kbeg) __syncthreads();
#pragma unroll
        for (int it = 0; it < NV; ++it) {
            const int e = threadIdx.x + it * 256, kk = e / 64, col = (e % 64) * 2;
            sA[kk * SPX + col] = fa[it].x;
            sA[kk * SPX + col + 1] = fa[it].y;
            sB[kk * SPX + col] = fb[it].x;
            sB[kk * SPX + col + 1] = fb[it].y;
        }
        __syncthreads();
        if (k0 + KC < kend) fetch(k0 + KC);
#pragma unroll 2
        for (int k = 0; k < KC; k += 4) {
            const double* pa = sA + (k + lk) * SPX;
            const double* pb = sB + (k + lk) * SPX;
            double a[4], b[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) a[x] = pa[r0 + 16 * x + li];
#pragma unroll
            for (int y = 0; y < 4; ++y) b[y] = pb[c0 + 16 * y + li];
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] = mfma(a[x], b[y], acc[x][y]);
        }
    }
    // tile (i, j) of this quadrant; above the diagonal or beyond the edge: none
    const int i = 2 * P + (r0 >> 6), j = 2 * Q + (c0 >> 6);
    if (i >= T || j >= T || j > i) return;
    double* Yij = Y + (size_t)i * T64 * ld + (size_t)j * T64;
    double* Yji = Y + (size_t)j * T64 * ld + (size_t)i * T64;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * x + lk + 4 * v, c = 16 * y + li;
                const double sv = acc[x][y][v] * scale;
                if (i != j || c <= r) {
                    Yij[(size_t)r * ld + c] = sv;
                    Yji[(size_t)c * ld + r] = sv;
                }
            }
}

template <bool MIXED>
__global__ __launch_bounds__(256) void lauum_kernel(const double* X, int ldx, double* Y, int ld, int T,
                                                    double scale) {
    int i, j;
    tri_decode(blockIdx.x, i, j);
    double* Yij = Y + (size_t)i * T64 * ld + (size_t)j * T64;
    double* Yji = Y + (size_t)j * T64 * ld + (size_t)i * T64;
    auto put = [&](int r, int c, double v) {
        double s = v * scale;
        if (i != j || c <= r) {
            Yij[(size_t)r * ld + c] = s;
            Yji[(size_t)c * ld + r] = s;
        }
    };
    if constexpr (MIXED) {
        __shared__ float fA[T64 * SPF];
        __shared__ float fB[T64 * SPF];
        v4d sum[2][2];
        acc_zero(sum);
        for (int m = i; m < T; ++m) {
            v2d va[8], vb[8];
            stage_fetch(va, X + (size_t)m * T64 * ldx + (size_t)i * T64, ldx);
            stage_fetch(vb, X + (size_t)m * T64 * ldx + (size_t)j * T64, ldx);
            stage_put_f32n(fA, va);  // op(A)(r,k) = X_mi[k][r]
            stage_put_f32n(fB, vb);  // op(B)(k,c) = X_mj[k][c]
            __syncthreads();
            v4f acc[2][2];
            acc32_zero(acc);
            mma64_f32(fA, fB, acc);
            acc32_add_to(acc, sum);
            __syncthreads();
        }
        acc32d_for_each(sum, put);
    } else {
        __shared__ double sA[T64 * SP];
        __shared__ double sB[T64 * SP];
        v4d acc[2][2];
        acc_zero(acc);
        for (int m = i; m < T; ++m) {
            stage<false>(sA, X + (size_t)m * T64 * ldx + (size_t)i * T64, ldx);  // op(A)(r,k) = X_mi[k][r]
            stage<false>(sB, X + (size_t)m * T64 * ldx + (size_t)j * T64, ldx);  // op(B)(k,c) = X_mj[k][c]
            __syncthreads();
            mma64(sA, sB, acc);
            __syncthreads();
        }
        acc_for_each(acc, put);
    }
}

// Newton-Schulz refinement of an SPD inverse, X1 = X0 + X0 (I - A X0), in
// two launches over 64 x 64 tiles (A and X0 symmetric, both triangles held):
//   ns_residual: E = I - A X0 in FP64 (the residual needs the full
//                precision: its entries are of the size of X0's error);
//   ns_update:   X1_ij = scale (X0_ij + sum_k X0_ik E_kj) for the lower
//                tiles, both triangles written (X1 is neither X0 nor E).  E is small, so its product with X0 runs
//                on the FP32 matrix cores when MIXED: the correction carries
//                an FP32-relative error, times |E|.
__global__ __launch_bounds__(256) void ns_residual_kernel(const double* A, const double* X0, int ld, int T,
                                                          double* E) {
    __shared__ double sA[T64 * SP];
    __shared__ double sB[T64 * SP];
    const int i = blockIdx.x / T, j = blockIdx.x - i * T;
    v4d acc[2][2];
    acc_zero(acc);
    // (A X0)_ij = sum_k A_ik X0_kj
    chain_product(0, T - 1, [&](int k) { return A + (size_t)i * T64 * ld + (size_t)k * T64; },
                  [&](int k) { return X0 + (size_t)k * T64 * ld + (size_t)j * T64; }, ld, ld, sA, sB, acc);
    double* Eij = E + (size_t)i * T64 * ld + (size_t)j * T64;
    acc_for_each(acc, [&](int r, int c, double v) { Eij[(size_t)r * ld + c] = ((i == j && r == c) ? 1.0 : 0.0) - v; });
}

template <bool MIXED>
__global__ __launch_bounds__(256) void ns_update_kernel(const double* X0, const double* E, int ld, int T,
                                                        double* X1, double scale) {
    int i, j;
    tri_decode(blockIdx.x, i, j);
    const double* X0ij = X0 + (size_t)i * T64 * ld + (size_t)j * T64;
    double* Xij = X1 + (size_t)i * T64 * ld + (size_t)j * T64;
    double* Xji = X1 + (size_t)j * T64 * ld + (size_t)i * T64;
    level_product<MIXED>(0, T - 1, [&](int k) { return X0 + (size_t)i * T64 * ld + (size_t)k * T64; },
                         [&](int k) { return E + (size_t)k * T64 * ld + (size_t)j * T64; }, ld, ld,
                         [&](int r, int c, double v) {
                             const double x = (X0ij[(size_t)r * ld + c] + v) * scale;
                             if (i != j || c <= r) {
                                 Xij[(size_t)r * ld + c] = x;
                                 Xji[(size_t)c * ld + r] = x;
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
// (la: look-ahead words of the factorisation that follows, reset here)
__global__ void copy_lower_kernel(const double* src, double* dst, int np, int ld, int* la) {
    if (la && blockIdx.x == 0) la_reset(la);
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

// Reversal of an np x np matrix, J S J with J the order-np reversal:
// dst[i][j] = scale * src[np-1-i][np-1-j] inside the image of the source's
// n x n block (leading when src_low, trailing otherwise), and pad_diag on
// the diagonal (zero elsewhere) in the image of the pad.  Both triangles are
// written.
__global__ void reverse_sym_kernel(const double* src, double* dst, int n, int np, int ld, double scale,
                                   double pad_diag, int src_low) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)np * np) return;
    const int i = (int)(t / np), j = (int)(t % np);
    const int ri = np - 1 - i, rj = np - 1 - j;
    const int lo = src_low ? 0 : np - n;
    double v;
    if (ri >= lo && ri < lo + n && rj >= lo && rj < lo + n)
        v = scale * src[(size_t)ri * ld + rj];
    else
        v = (i == j) ? pad_diag : 0.0;
    dst[(size_t)i * ld + j] = v;
}

// B = scale * J X^T J for lower-triangular X (B lower; zero above):
// B[i][j] = scale * X[np-1-j][np-1-i].
__global__ void reverse_transpose_lower_kernel(const double* X, double* B, int np, int ld, double scale) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)np * np) return;
    const int i = (int)(t / np), j = (int)(t % np);
    B[(size_t)i * ld + j] = (j <= i) ? scale * X[(size_t)(np - 1 - j) * ld + (np - 1 - i)] : 0.0;
}

// Inverses of B's diagonal tiles for B = c J X^T J, X = L^{-1}: tile k of B
// is c revT(X_{T-1-k}) (revT(M)[r][q] = M[63-q][63-r]), so its inverse is
// revT(L_{T-1-k}) / c, read from L's diagonal tiles (lower parts).
__global__ void reverse_diag_inverse_kernel(const double* L, double* W, int T, int ld, double inv_c) {
    const int k = blockIdx.x;
    const double* Lt = L + (size_t)(T - 1 - k) * T64 * ld + (size_t)(T - 1 - k) * T64;
    double* Wk = W + (size_t)k * T64 * T64;
    for (int e = threadIdx.x; e < T64 * T64; e += blockDim.x) {
        const int r = e >> 6, q = e & 63;
        Wk[e] = (q <= r) ? inv_c * Lt[(size_t)(63 - q) * ld + (63 - r)] : 0.0;
    }
}

// A well-conditioned SPD matrix of order n in an np x np buffer (pad =
// identity) for timing the factorisation: A_ij = 1 / (1 + |i - j|) off the
// diagonal, n + 1 on it (strictly diagonally dominant).
__global__ void spd_fill_kernel(double* A, int n, int np, int ld) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)np * np) return;
    const int r = (int)(t / np), c = (int)(t % np);
    double v;
    if (r < n && c < n)
        v = (r == c) ? (double)n + 1.0 : 1.0 / (1.0 + (double)abs(r - c));
    else
        v = (r == c) ? 1.0 : 0.0;
    A[(size_t)r * ld + c] = v;
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

// work = [T inverse tiles of the diagonal factor][LA_WORDS look-ahead ints]
//        [one scratch tile]
static size_t dense_scratch_offset(int np) { return (size_t)np * T64 + LA_WORDS / 2; }
size_t pbx_dense_work_doubles(int np) { return dense_scratch_offset(np) + (size_t)T64 * T64; }

// Cholesky variant: 1 = launch per panel, column-by-column diagonal tiles;
// 2 = launch per panel, 8-column panel diagonal tiles; 4 = one launch per
// panel with the solve folded in and wave-specialised diagonal tiles; 6
// (default) = two tile columns per launch (rank-128 trailing updates); 0
// selects the default.  (3 and 5 were persistent single-launch variants,
// slower than 6; they are no longer built.)
constexpr int kDefaultPotrf = 6;
static int g_potrf_variant = kDefaultPotrf;
static int g_own_cu_tiles = 64;
static std::atomic<int> g_shape_generation{0};  // launch-shape changes (keys captured graphs)
static std::atomic<int> g_far_period{4};
// smallest level (block size in tiles) of the triangular inverse that runs
// over 128 x 128 blocks: below it the levels have too few blocks to fill
// the GPU with the larger ones (n = 7232: level 32 0.44 against 0.38 ms,
// level 64 1.18 + 0.79 against 1.43 + 0.89 ms)
constexpr int g_trtri128_min_b = 64;  // K of the large orders' whole-trailing updates
static int g_chain_jobs = 0;
extern "C" int pbx_set_potrf_chain_jobs(int on) {
    g_chain_jobs = on != 0;
    return PBX_OK;
}
// Largest tile-column count factored with one workgroup per compute unit
// (default 64; larger orders keep two per CU).
extern "C" int pbx_set_potrf_own_cu_tiles(int t) {
    if (t < 0) return PBX_ERR_ARG;
    g_own_cu_tiles = t;
    ++g_shape_generation;
    return PBX_OK;
}
extern "C" int pbx_get_potrf_own_cu_tiles(void) { return g_own_cu_tiles; }
// K: the large orders' whole trailing matrix is updated every K launches
// (1 = every launch; default 4).
extern "C" int pbx_set_potrf_far_period(int k) {
    if (k < 1 || k > 8) return PBX_ERR_ARG;
    g_far_period = k;
    ++g_shape_generation;
    return PBX_OK;
}

// Per-order launch shapes, as chosen by a tuner (pybrown_amd.autotune):
// code[T] for a factorisation of T tile columns is 0 (the process-wide
// variant and own-CU setting), 4 (variant 4), 61 (variant 6, one workgroup
// per CU) or 62 (variant 6, two per CU).  The MIXED factor exists in the
// two-column launches only and keeps variant 6.  Every setter bumps the
// generation that keys captured graphs.
constexpr int kShapeTableMax = 1024;
static std::atomic<int> g_shape_table[kShapeTableMax];
int pbx_potrf_config_generation() { return g_shape_generation.load(); }

struct PotrfShape {
    int variant;
    bool own_cu;
};
static PotrfShape potrf_shape(int T, int mixed) {
    PotrfShape sh{g_potrf_variant, T <= g_own_cu_tiles};
    const int code = T < kShapeTableMax ? g_shape_table[T].load(std::memory_order_relaxed) : 0;
    if (code == 4) sh.variant = 4;
    if (code == 61 || code == 62) {
        sh.variant = 6;
        sh.own_cu = code == 61;
    }
    if (mixed) sh.variant = 6;
    return sh;
}

extern "C" int pbx_set_potrf_table(const int* codes, int len) {
    if (len < 0 || len > kShapeTableMax || (len > 0 && !codes)) {
        pbx_set_error(PBX_ERR_ARG, "set_potrf_table: length %d outside [0, %d]", len, kShapeTableMax);
        return PBX_ERR_ARG;
    }
    for (int t = 0; t < len; ++t)
        if (codes[t] != 0 && codes[t] != 4 && codes[t] != 61 && codes[t] != 62) {
            pbx_set_error(PBX_ERR_ARG, "set_potrf_table: code %d at %d (0, 4, 61 or 62)", codes[t], t);
            return PBX_ERR_ARG;
        }
    for (int t = 0; t < kShapeTableMax; ++t) g_shape_table[t].store(t < len ? codes[t] : 0);
    ++g_shape_generation;
    return PBX_OK;
}

extern "C" int pbx_get_potrf_shape(int T, int* variant, int* own_cu) {
    if (T <= 0 || !variant || !own_cu) return PBX_ERR_ARG;
    const PotrfShape sh = potrf_shape(T, 0);
    *variant = sh.variant;
    *own_cu = sh.own_cu;
    return PBX_OK;
}

extern "C" int pbx_set_potrf_variant(int v) {
    if (v < 0 || v > 6 || v == 3 || v == 5) return PBX_ERR_ARG;
    g_potrf_variant = v ? v : kDefaultPotrf;
    ++g_shape_generation;
    return PBX_OK;
}
extern "C" int pbx_get_potrf_variant(void) { return g_potrf_variant; }

// Test hook (pbx_debug_inject_wait_failure): counts down the launch
// sequences that report through a status word -- factorisations and the
// midpoint's B0^T solve -- and appends to the k-th one a kernel that sets its
// status word to PBX_INFO_DAG_WAIT, as an exhausted bounded wait would.
static std::atomic<int> g_inject{0};

__global__ void inject_status_kernel(int* info) {
    if (threadIdx.x == 0) atomicExch(info, PBX_INFO_DAG_WAIT);
}

void pbx_status_sequence_queued(int* info_dev, hipStream_t s) {
    int k = g_inject.load();
    while (k > 0 && !g_inject.compare_exchange_weak(k, k - 1)) {
    }
    if (k == 1) PBX_LAUNCH(inject_status_kernel, dim3(1), dim3(64), 0, s, info_dev);
}

extern "C" int pbx_debug_inject_wait_failure(int k) {
    if (k < 0) return PBX_ERR_ARG;
    g_inject.store(k);
    return PBX_OK;
}

int pbx_fill_pad(double* A, int n, int np, int ld, hipStream_t s) {
    if (np == n) return PBX_OK;
    size_t total = (size_t)np * np;
    PBX_LAUNCH(fill_pad_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, A, n, np, ld);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_copy_lower(const double* src, double* dst, int np, int ld, hipStream_t s) {
    int T = np / T64;
    PBX_LAUNCH(copy_lower_kernel, dim3(T * (T + 1) / 2), dim3(256), 0, s, src, dst, np, ld, nullptr);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// One two-column launch of a factorisation: its columns (c0, nc), the
// previous launch's column count P, its workgroup counts (nh helper jobs,
// nw rows, ntr trailing tiles) and the lags of its trailing tiles
// (two_col_trailing).
struct TwoColLaunch {
    int c0, P, nc, nh, nw, Pt, near, Pf, fq0, K, Pc, ntr;
};

// The launch sequence of a T-tile factorisation (K: the far period; one
// workgroup per CU: K = 1 and no rolling).
// Far period of the one-workgroup-per-CU orders' rolling schedule (0: every
// trailing tile in every launch).  C2 (n = 3000): 0.974 ms with every tile
// in every launch; rolling 0.967 (period 2), 0.971 (3), 0.957 (4),
// 0.976 (5), 0.981 (6), 1.056 ms (8).
static std::atomic<int> g_own_cu_roll{4};
extern "C" int pbx_set_potrf_own_cu_roll(int k) {
    if (k < 0 || k > 8) return PBX_ERR_ARG;
    g_own_cu_roll = k;
    ++g_shape_generation;
    return PBX_OK;
}

static void two_col_plan(int T, bool own_cu, int K, std::vector<TwoColLaunch>& plan) {
    if (own_cu && g_own_cu_roll.load() > 0) {
        // the rolling schedule (below) in the one-per-CU build
        K = g_own_cu_roll.load();
        own_cu = false;
    }
    // rolling far updates up to 128 tile columns; above, every K-th launch
    // updates the whole trailing matrix (n = 12000: 15.0 ms either way, but
    // 32 GB of HBM traffic against 49 GB rolling, whose slices re-read each
    // row panel from HBM once per slice)
    const bool roll = !own_cu && T <= 128;
    int c_far = 0;
    std::vector<int> upd(roll ? (T + 1) / 2 + 2 : 0, 0);  // per pair: panels applied, 0 .. upd-1
    plan.clear();
    for (int c0 = 0, P = 0; c0 < T;) {
        const int nc = T - c0 < 2 ? T - c0 : 2;
        const int nh = nc == 2 ? (P > 0 ? 2 : 1) : 0;
        const int nw = T - c0 - nc;
        int Pt = c0 - c_far;
        int near = (Pt > 0 && Pt < 2 * K) ? 1 : 0;
        int Pf = 0, fq0 = 0, nfar = 0, Pc = 0;
        if (!roll) {
            if (Pt > 0 && !near) c_far = c0;
        } else if (nc < 2 || nw == 0) {
            Pt = near = 0;  // no tile beyond the launch's own columns
        } else {
            const int L = c0 / 2, qn = L + 1, qc = L + 2;
            Pt = c0 - upd[qn];
            upd[qn] = c0;
            near = Pt > 0 ? 1 : 0;
            if (2 * qc < T) {  // catch-up of the next launch's near pair
                Pc = c0 - upd[qc];
                upd[qc] = c0;
                nfar += Pc > 0 ? far_count(T, qc, T) : 0;
            }
            fq0 = qc + 1;
            while (fq0 % K != L % K) ++fq0;
            if (2 * fq0 < T) {
                Pf = c0 - upd[fq0];
                for (int q = fq0; 2 * q < T; q += K) upd[q] = c0;
                nfar += Pf > 0 ? far_count(T, fq0, K) : 0;
            }
        }
        int ntr = 0;
        if (roll)
            ntr = near ? 2 * nw - 2 + nfar : 0;
        else if (Pt > 0 && nw > 0)
            ntr = near ? 2 * nw - 2 : nw * (nw + 1) / 2 - (nc == 2 ? 1 : 0);
        plan.push_back({c0, P, nc, nh, nw, Pt, near, Pf, fq0, roll ? K : 1, Pc, ntr});
        P = nc;
        c0 += nc;
    }
}

// Every tile update a T-tile factorisation's launches make, as the kernel's
// roles decode them: records (launch, i, j, m0, P) of tile (i, j) less
// panels m0 .. m0+P-1, five ints each, into rec (up to cap records).
// Returns the number of records (possibly more than cap).  The rows also
// subtract their launch's own two columns from tile (c0+2, c0+2) (recorded
// as panels c0, c0+1); the diagonal chain's own tiles are not recorded.
extern "C" int pbx_debug_two_col_updates(int T, int own_cu, int K, int* rec, int cap) {
    if (T <= 0 || K <= 0) return PBX_ERR_ARG;
    std::vector<TwoColLaunch> plan;
    two_col_plan(T, own_cu != 0, own_cu ? 1 : K, plan);
    int n = 0;
    auto put = [&](int L, int i, int j, int m0, int P) {
        if (P > 0 && n < cap && rec) {
            int* r = rec + 5 * n;
            r[0] = L, r[1] = i, r[2] = j, r[3] = m0, r[4] = P;
        }
        if (P > 0) ++n;
    };
    for (int L = 0; L < (int)plan.size(); ++L) {
        const TwoColLaunch& l = plan[L];
        if (l.nc == 2) {  // helper jobs
            put(L, l.c0 + 1, l.c0, l.c0 - l.P, l.P);
            put(L, l.c0 + 1, l.c0 + 1, l.c0 - l.P, l.P);
        }
        for (int i = l.c0 + l.nc; i < T; ++i) {  // rows
            put(L, i, l.c0, l.c0 - l.P, l.P);
            if (l.nc == 2) put(L, i, l.c0 + 1, l.c0 - l.P, l.P);
            if (l.nc == 2 && i == l.c0 + 2) {
                put(L, i, i, l.c0 - l.Pt, l.Pt);
                put(L, i, i, l.c0, 2);
            }
        }
        for (int r = 0; r < l.ntr; ++r) {
            int i, j, m0, P;
            two_col_trailing(r, T, l.c0, l.nc, l.near, l.Pt, l.Pf, l.fq0, l.K, l.Pc, own_cu ? 7 : 8, i, j, m0, P);
            put(L, i, j, m0, P);
        }
    }
    return n;
}

// Variant 6 launch sequence: launch (c0, P, nc) for c0 = 0, 2, 4, ...
// la_ready: the look-ahead words are already zero (reset by the kernel
// before, on the same stream).
template <bool MIXED>
static void potrf_two_col_launch(double* A, int np, int ld, double* work, int* info_dev, hipStream_t s,
                                 bool la_ready, bool own_cu) {
    const int T = np / T64;
    int* la = (int*)(work + (size_t)np * T64);
    // Up to 64 tile columns the trailing updates of a launch finish well
    // within its diagonal chain, so each workgroup gets a compute unit of
    // its own (static plus dynamic LDS beyond half the CU's 160 KB): nothing
    // shares the chain's CU.  The dynamic part holds the two tiles the
    // diagonal chain stages for its fused second column (2 x 64 x SP
    // doubles).  Larger orders are throughput-bound and keep two workgroups
    // per CU.
    const size_t dyn_lds = own_cu ? 2 * T64 * SP * sizeof(double) : 0;
    if (!la_ready) PBX_LAUNCH(la_reset_kernel, dim3(1), dim3(64), 0, s, la);
    int upd_target = 0;
    // Large orders (two workgroups per CU) update each far trailing tile
    // once per K launches (2K panels at once, a rank-128K product), so each
    // far tile is read and written once per 2K columns (C4, n = 12000:
    // 41 GB of HBM traffic per factorisation with K = 1, 28.5 GB with K = 2;
    // K = 4, the default, was the fastest of 1, 2, 3, 4 and 6 there).  The
    // far updates roll: tile-column pair q (columns 2q, 2q+1) takes its
    // lagging panels in the launches L = c0/2 with L = q mod K, and the rest
    // in launch q-1, the one before it becomes a row pair (its `near`
    // columns).  Each launch then carries 1/K of the far work beside its
    // diagonal chain (n = 7232: 4.71 -> 4.54 ms against all of it in every
    // K-th launch).  Launch L also brings pair L+2 up to date (lag Pc), so
    // the next launch's near columns and its diagonal row lag two panels
    // only (n = 7232: 4.54 -> 4.40 ms).  The pairs of a launch's slice
    // (q >= L+3) were last updated together (launch L-K, or never), so one
    // lag Pf serves them all.  Orders up to the own-CU limit are bound
    // by the diagonal chain and take K = 1, every trailing tile in every
    // launch (K = 2 cost 3 % at n = 3000).  The row and helper tiles of a
    // launch always lag two panels.
    std::vector<TwoColLaunch> plan;
    two_col_plan(T, own_cu, own_cu ? 1 : g_far_period.load(), plan);
    for (const TwoColLaunch& l : plan) {
        upd_target += l.nh;
        // own-CU build: roles 1 .. W sit at the indices that are not
        // multiples of 8 (the smallest grid holding W of them; role
        // b - floor(b / 8) of index b runs over 1 .. W exactly once)
        if (own_cu) {
            const int W = l.nh + l.nw + l.ntr;
            int grid = 1;
            while ((grid - 1) - (grid - 1) / 8 < W) ++grid;
            PBX_LAUNCH((potrf2_kernel<5, 1, MIXED>), dim3(grid), dim3(256), dyn_lds, s, A, ld, T, l.c0, l.P, l.nc,
                       upd_target, work, work + dense_scratch_offset(np), info_dev, la, g_chain_jobs, l.Pt, l.near,
                       l.Pf, l.fq0, l.K, l.Pc);
        } else {
            PBX_LAUNCH((potrf2_kernel<5, 2, MIXED>), dim3(1 + l.nh + l.nw + l.ntr), dim3(256), 0, s, A, ld, T, l.c0,
                       l.P, l.nc, upd_target, work, work + dense_scratch_offset(np), info_dev, la, g_chain_jobs,
                       l.Pt, l.near, l.Pf, l.fq0, l.K, l.Pc);
        }
    }
}

static void potrf_two_col(double* A, int np, int ld, double* work, int* info_dev, hipStream_t s, bool mixed,
                          bool la_ready = false) {
    const bool own_cu = potrf_shape(np / T64, mixed).own_cu;
    if (mixed)
        potrf_two_col_launch<true>(A, np, ld, work, info_dev, s, la_ready, own_cu);
    else
        potrf_two_col_launch<false>(A, np, ld, work, info_dev, s, la_ready, own_cu);
}

// The two-column factorisation (variant 6, FP64) of A in place, its
// look-ahead words already zero (cleared by the kernel before it on the same
// stream).  For callers that stage A themselves (pbx_noalloc.hip).
int pbx_potrf_two_col_ready(double* A, int np, int ld, double* work, int* info_dev, hipStream_t s) {
    if (np % T64 != 0 || np <= 0) {
        pbx_set_error(PBX_ERR_ARG, "potrf: order %d is not a positive multiple of 64", np);
        return PBX_ERR_ARG;
    }
    potrf_two_col(A, np, ld, work, info_dev, s, false, true);
    pbx_status_sequence_queued(info_dev, s);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// In-place lower Cholesky of the np x np matrix A; inverses of the diagonal
// factor tiles are left in work (pbx_dense_work_doubles(np) doubles).
// *info_dev must be zero on entry; it receives column+1 of the first
// non-positive pivot.  mixed: FP32 trailing updates (always the two-column
// launches).
int pbx_potrf(double* A, int np, int ld, double* work, int* info_dev, hipStream_t s, int mixed) {
    if (np % T64 != 0 || np <= 0) {
        pbx_set_error(PBX_ERR_ARG, "potrf: order %d is not a positive multiple of 64", np);
        return PBX_ERR_ARG;
    }
    const int T = np / T64;
    const int variant = potrf_shape(T, mixed).variant;
    if (variant == 6) {
        potrf_two_col(A, np, ld, work, info_dev, s, mixed != 0);
        pbx_status_sequence_queued(info_dev, s);
        PBX_HIP(hipGetLastError());
        return PBX_OK;
    }
    if (variant >= 4) {
        int* la = (int*)(work + (size_t)np * T64);
        PBX_LAUNCH(potrf_diag_kernel<5>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev, la);
        if (T > 1) PBX_LAUNCH(trsm_kernel, dim3(T - 1), dim3(256), 0, s, A, ld, 0, work);
        for (int k = 0; k + 1 < T; ++k) {
            int cnt = (T - 1 - k) * (T - k) / 2;
            PBX_LAUNCH(syrk_trsm_kernel<5>, dim3(cnt), dim3(256), 0, s, A, ld, k, work, info_dev, la);
        }
        pbx_status_sequence_queued(info_dev, s);
        PBX_HIP(hipGetLastError());
        return PBX_OK;
    }
    const bool v2 = variant == 2;
    if (v2)
        PBX_LAUNCH(potrf_diag_kernel<2>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev, nullptr);
    else
        PBX_LAUNCH(potrf_diag_kernel<1>, dim3(1), dim3(256), 0, s, A, ld, 0, work, info_dev, nullptr);
    for (int k = 0; k + 1 < T; ++k) {
        PBX_LAUNCH(trsm_kernel, dim3(T - 1 - k), dim3(256), 0, s, A, ld, k, work);
        int cnt = (T - 1 - k) * (T - k) / 2;
        if (v2)
            PBX_LAUNCH(syrk_kernel<2>, dim3(cnt), dim3(256), 0, s, A, ld, k, work, info_dev);
        else
            PBX_LAUNCH(syrk_kernel<1>, dim3(cnt), dim3(256), 0, s, A, ld, k, work, info_dev);
    }
    pbx_status_sequence_queued(info_dev, s);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// B = the lower part of D, then B = its Cholesky factor as pbx_potrf.  With
// the two-column launches the copy also resets their look-ahead words, one
// launch fewer per factorisation.
int pbx_potrf_copy(const double* D, double* B, int np, int ld, double* work, int* info_dev, hipStream_t s,
                   int mixed) {
    if (np % T64 != 0 || np <= 0) {
        pbx_set_error(PBX_ERR_ARG, "potrf: order %d is not a positive multiple of 64", np);
        return PBX_ERR_ARG;
    }
    const int T = np / T64;
    const bool two_col = potrf_shape(T, mixed).variant == 6;
    int* la = two_col ? (int*)(work + (size_t)np * T64) : nullptr;
    PBX_LAUNCH(copy_lower_kernel, dim3(T * (T + 1) / 2), dim3(256), 0, s, D, B, np, ld, la);
    if (!two_col) return pbx_potrf(B, np, ld, work, info_dev, s, mixed);
    potrf_two_col(B, np, ld, work, info_dev, s, mixed != 0, true);
    pbx_status_sequence_queued(info_dev, s);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// X = inv(L) (lower tiles; the upper tiles hold the recursion's partial
// products) from the factor L in A and its diagonal-tile inverses in work.
template <bool MIXED>
static void trtri(const double* A, int T, int ld, const double* work, double* X, hipStream_t s) {
    PBX_LAUNCH(copy_diag_inv_kernel, dim3(T), dim3(256), 0, s, X, ld, work);
    for (int b = 1; b < T; b *= 2) {
        const int npairs = (T - b + 2 * b - 1) / (2 * b);
        if (!MIXED && b >= g_trtri128_min_b) {
            const int hb = b / 2;
            PBX_LAUNCH(trtri128_kernel<0>, dim3(npairs * hb * hb), dim3(256), 0, s, A, ld, X, ld, T, b);
            PBX_LAUNCH(trtri128_kernel<1>, dim3(npairs * hb * hb), dim3(256), 0, s, A, ld, X, ld, T, b);
            continue;
        }
        PBX_LAUNCH((trtri_level_kernel<0, MIXED>), dim3(npairs * b * b), dim3(256), 0, s, A, ld, X, ld, T, b);
        PBX_LAUNCH((trtri_level_kernel<1, MIXED>), dim3(npairs * b * b), dim3(256), 0, s, A, ld, X, ld, T, b);
    }
}

template <bool MIXED>
static void trtri_lauum(double* A, int T, int ld, double* work, double* X, hipStream_t s, double scale) {
    // every lower tile of X is written below before lauum reads it
    trtri<MIXED>(A, T, ld, work, X, s);
    if (MIXED) {
        PBX_LAUNCH(lauum_kernel<MIXED>, dim3(T * (T + 1) / 2), dim3(256), 0, s, X, ld, A, ld, T, scale);
    } else {
        const int nb = (T + 1) / 2;
        PBX_LAUNCH(lauum128_kernel, dim3(nb * (nb + 1) / 2), dim3(256), 0, s, X, ld, A, ld, T, scale);
    }
}

// A := scale * inv(A) for SPD A (both triangles written).  X is an np x np
// scratch matrix (leading dimension ld).  precision 0: FP64 throughout;
// 1: the factorisation's trailing updates and the products of the triangular
// inverse and of X^T X on the FP32 matrix cores (diagonal tiles, their
// inverses and every stored value FP64); 2: as 1, then one Newton-Schulz step
// with an FP64 residual (C: a further np x np matrix, overwritten).
int pbx_spd_inverse(double* A, int np, int ld, double* work, double* X, int* info_dev, hipStream_t s,
                    double scale, int precision, double* C) {
    if (precision < 0 || precision > 2 || (precision == 2 && !C)) {
        pbx_set_error(PBX_ERR_ARG, "spd_inverse: precision %d (2 needs a second scratch matrix)", precision);
        return PBX_ERR_ARG;
    }
    const int T = np / T64;
    if (precision == 2) {
        // C = A; C := inv(C) in the mixed precision; X := I - A C;
        // A := scale (C + C X)
        PBX_HIP(hipMemcpyAsync(C, A, sizeof(double) * (size_t)np * ld, hipMemcpyDeviceToDevice, s));
        const int rc = pbx_potrf(C, np, ld, work, info_dev, s, 1);
        if (rc) return rc;
        trtri_lauum<true>(C, T, ld, work, X, s, 1.0);
        PBX_LAUNCH(ns_residual_kernel, dim3(T * T), dim3(256), 0, s, A, C, ld, T, X);
        PBX_LAUNCH(ns_update_kernel<true>, dim3(T * (T + 1) / 2), dim3(256), 0, s, C, X, ld, T, A, scale);
        PBX_HIP(hipGetLastError());
        return PBX_OK;
    }
    const int rc = pbx_potrf(A, np, ld, work, info_dev, s, precision);
    if (rc) return rc;
    if (precision)
        trtri_lauum<true>(A, T, ld, work, X, s, scale);
    else
        trtri_lauum<false>(A, T, ld, work, X, s, scale);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// D = kBT inv(R) and its lower Cholesky factor B in one pass (box.py:537 and
// :546-556 for a lubrication step that rebuilds D and factors it).  With J
// the order-np reversal, J R J = L L^T gives R = (J L J)(J L J)^T with J L J
// upper triangular, so inv(R) = (J L^{-T} J)(J L^{-T} J)^T and J L^{-T} J is
// lower triangular with a positive diagonal: the Cholesky factor of inv(R).
// Hence B = sqrt(kBT) J X^T J and D = kBT J X^T X J with X = inv(L): one
// factorisation, one triangular inverse and one X^T X, where inverting R and
// then factoring D take two factorisations.  The reversed pad carries kBT on
// its diagonal, so D and B come out with identity pads.  Wc receives the
// inverses of B's diagonal tiles (the factorisation's work area, as a
// factorisation of D leaves them); R is overwritten with D; X and Wi are
// scratch (np x np, and a factorisation work area).  form_D 0 leaves D
// unformed: R then holds X = inv(L) (lower tiles), from which
// pbx_inverse_and_factor_finish forms D later with the same operations.
int pbx_inverse_and_factor(double* R, double* X, double* B, double* Wc, double* Wi, int n, int np, int ld,
                           double kBT, int* info_dev, hipStream_t s, int form_D) {
    if (np % T64 != 0 || np <= 0 || n <= 0 || n > np || !(kBT > 0.0)) {
        pbx_set_error(PBX_ERR_ARG, "inverse_and_factor: n %d np %d kBT %g", n, np, kBT);
        return PBX_ERR_ARG;
    }
    const int T = np / T64;
    const double c = sqrt(kBT);
    const unsigned g = (unsigned)(((size_t)np * np + 255) / 256);
    PBX_LAUNCH(reverse_sym_kernel, dim3(g), dim3(256), 0, s, R, X, n, np, ld, 1.0, kBT, 1);
    const int rc = pbx_potrf(X, np, ld, Wi, info_dev, s, 0);  // X := L
    if (rc) return rc;
    trtri<false>(X, T, ld, Wi, R, s);                                 // R := inv(L) (lower tiles)
    PBX_LAUNCH(reverse_transpose_lower_kernel, dim3(g), dim3(256), 0, s, R, B, np, ld, c);
    PBX_LAUNCH(reverse_diag_inverse_kernel, dim3(T), dim3(256), 0, s, X, Wc, T, ld, 1.0 / c);
    if (form_D) return pbx_inverse_and_factor_finish(R, X, n, np, ld, kBT, s);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// The factorisation half of pbx_inverse_and_factor: L (np x np) := the
// Cholesky factor of J R J (pad diagonal kBT), the inverses of its diagonal
// tiles in W (a factorisation work area).  R is left unchanged.  With
// B = sqrt(kBT) J L^{-T} J, B xi and B^{-T} xi take one triangular solve and
// one triangular product with L; pbx_factor_reversed_finish forms B itself.
int pbx_factor_reversed(const double* R, double* L, double* W, int n, int np, int ld, double kBT, int* info_dev,
                        hipStream_t s) {
    if (np % T64 != 0 || np <= 0 || n <= 0 || n > np || !(kBT > 0.0)) {
        pbx_set_error(PBX_ERR_ARG, "factor_reversed: n %d np %d kBT %g", n, np, kBT);
        return PBX_ERR_ARG;
    }
    const unsigned g = (unsigned)(((size_t)np * np + 255) / 256);
    PBX_LAUNCH(reverse_sym_kernel, dim3(g), dim3(256), 0, s, R, L, n, np, ld, 1.0, kBT, 1);
    return pbx_potrf(L, np, ld, W, info_dev, s, 0);
}

// B (in place of L) and the inverses of its diagonal tiles (in place of W)
// from pbx_factor_reversed's L and W, with X = inv(L) left in Xout (lower
// tiles) for pbx_inverse_and_factor_finish: the same operations as
// pbx_inverse_and_factor, so the same values.
int pbx_factor_reversed_finish(double* L, double* W, double* Xout, int np, int ld, double kBT, hipStream_t s) {
    const int T = np / T64;
    const double c = sqrt(kBT);
    const unsigned g = (unsigned)(((size_t)np * np + 255) / 256);
    trtri<false>(L, T, ld, W, Xout, s);
    PBX_LAUNCH(reverse_diag_inverse_kernel, dim3(T), dim3(256), 0, s, L, W, T, ld, 1.0 / c);
    PBX_LAUNCH(reverse_transpose_lower_kernel, dim3(g), dim3(256), 0, s, Xout, L, np, ld, c);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// R := D = kBT J X^T X J from R = X = inv(L) as pbx_inverse_and_factor left
// it with form_D 0 (X: np x np scratch).
int pbx_inverse_and_factor_finish(double* R, double* X, int n, int np, int ld, double kBT, hipStream_t s) {
    const int T = np / T64;
    const unsigned g = (unsigned)(((size_t)np * np + 255) / 256);
    const int nb = (T + 1) / 2;
    PBX_LAUNCH(lauum128_kernel, dim3(nb * (nb + 1) / 2), dim3(256), 0, s, R, ld, X, ld, T, kBT);  // X := kBT X^T X
    PBX_LAUNCH(reverse_sym_kernel, dim3(g), dim3(256), 0, s, X, R, n, np, ld, 1.0, 1.0, 0);        // R := D
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// ||D - B B^T||_F / ||D||_F over the leading n x n block (acc: 2 doubles of
// device scratch, cleared here).
int pbx_factor_residual(const double* D, const double* B, int n, int np, int ld, double* acc, double* rel,
                        hipStream_t s) {
    const int T = np / T64;
    PBX_HIP(hipMemsetAsync(acc, 0, 2 * sizeof(double), s));
    PBX_LAUNCH(residual_kernel, dim3(T * (T + 1) / 2), dim3(256), 0, s, D, B, ld, n, acc);
    PBX_HIP(hipGetLastError());
    double h[2];
    PBX_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
    PBX_HIP(hipStreamSynchronize(s));
    *rel = h[1] > 0 ? sqrt(h[0] / h[1]) : 0.0;
    return PBX_OK;
}

int pbx_launch_spd_fill(double* A, int n, int np, int ld, hipStream_t s) {
    const size_t total = (size_t)np * np;
    PBX_LAUNCH(spd_fill_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, A, n, np, ld);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_gemm_nt_test(const double* A, const double* B, double* C, int M, int N, int K, hipStream_t s) {
    if (M % 64 || N % 64 || K % 64) {
        pbx_set_error(PBX_ERR_ARG, "gemm_nt: dimensions must be multiples of 64");
        return PBX_ERR_ARG;
    }
    PBX_LAUNCH(gemm_nt_kernel, dim3(N / 64, M / 64), dim3(256), 0, s, A, B, C, M, N, K);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

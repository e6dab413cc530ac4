// Propagation kernels on gfx950: Ermak-McCammon / midpoint sub-steps with the
// signed plane-crossing count and the periodic wrap (reference:
// src/bdsim/pyBrown/box.py:353-389, :614-617; bead.py:92-132, :170-183).
//
// Matrix-vector products read the row-major matrices one wave per row:
// lanes stride the row (64 x 8 B = 512 B per load), a butterfly reduces the
// partial sums.  They are HBM-bound; the lower-triangular product with B
// reads only c <= r.
#include "pbx_common.h"

namespace {

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// dot of a row segment [0, len) with x
__device__ __forceinline__ double row_dot(const double* __restrict__ row, const double* __restrict__ x, int len) {
    const int lane = threadIdx.x & 63;
    double s0 = 0.0, s1 = 0.0;
    int c = lane;
    for (; c + 64 < len; c += 128) {
        s0 += row[c] * x[c];
        s1 += row[c + 64] * x[c + 64];
    }
    if (c < len) s0 += row[c] * x[c];
    return wave_sum(s0 + s1);
}

struct FluxPlane {
    double n[3], p[3];
    int on;
};

// sign of n.(r - p) > 0 with plain left-to-right arithmetic
__device__ __forceinline__ int plane_side(const FluxPlane& f, double x, double y, double z) {
#pragma clang fp contract(off)
    double d = f.n[0] * (x - f.p[0]) + f.n[1] * (y - f.p[1]) + f.n[2] * (z - f.p[2]);
    return d > 0.0 ? 1 : 0;
}

// r_b += v_b (bead b), with the per-label signed crossing count
__device__ __forceinline__ void move_bead(double* pos, const double* v, int b, const FluxPlane& f,
                                          const int* label, long long* flux) {
    double x0 = pos[3 * b], y0 = pos[3 * b + 1], z0 = pos[3 * b + 2];
    double x1 = x0 + v[0], y1 = y0 + v[1], z1 = z0 + v[2];
    pos[3 * b] = x1;
    pos[3 * b + 1] = y1;
    pos[3 * b + 2] = z1;
    if (f.on) {
        int s0 = plane_side(f, x0, y0, z0), s1 = plane_side(f, x1, y1, z1);
        if (s0 != s1) atomicAdd((unsigned long long*)&flux[label[b]], (unsigned long long)(long long)(s1 - s0));
    }
}

// Stochastic sub-step (box.py:366-375): BX = ((B xi) * sqrt(2 dt)) * mult,
// B lower-triangular.  One 192-thread workgroup per bead (a wave per row).
// A replayed step's noise vector and flux slot (cursor non-null).
__device__ __forceinline__ void cursor_pointers(const PbxStepCursor* cur, const double*& xi, long long*& flux) {
    if (!cur) return;
    const long long st = cur->step;
    xi = cur->xi_dev + ((cur->ring_base + st) % cur->ring_att) * cur->n;
    flux = cur->flux_log + st * cur->nl;
}

__global__ __launch_bounds__(192) void stoch_kernel(const double* __restrict__ B, int ld, const double* xi,
                                                    double sq2dt, double mult, double* pos, double* disp,
                                                    FluxPlane f, const int* label, long long* flux,
                                                    const PbxStepCursor* cur) {
    cursor_pointers(cur, xi, flux);
    __shared__ double v[3];
    const int b = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int r = 3 * b + w;
    double s = row_dot(B + (size_t)r * ld, xi, r + 1);
    if ((threadIdx.x & 63) == 0) {
        double d = (s * sq2dt) * mult;
        v[w] = d;
        disp[r] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) move_bead(pos, v, b, f, label, flux);
}

// nohi variant: B is the element-wise sqrt of the diagonal D (box.py:369)
__global__ void stoch_diag_kernel(const double* __restrict__ Bd, const double* xi, double sq2dt,
                                  double mult, double* pos, double* disp, int N, FluxPlane f, const int* label,
                                  long long* flux, const PbxStepCursor* cur) {
    cursor_pointers(cur, xi, flux);
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    double v[3];
    for (int c = 0; c < 3; ++c) {
        v[c] = ((Bd[3 * b + c] * xi[3 * b + c]) * sq2dt) * mult;
        disp[3 * b + c] = v[c];
    }
    move_bead(pos, v, b, f, label, flux);
}

// Deterministic sub-step (box.py:353-362): FX = (((dt/kBT) D) F) * mult with
// D full symmetric; one wave per row.
__global__ __launch_bounds__(192) void det_kernel(const double* __restrict__ D, int ld, const double* __restrict__ F,
                                                  double k, double mult, int n, double* pos, FluxPlane f,
                                                  const int* label, long long* flux, const PbxStepCursor* cur) {
    const double* unused_xi = nullptr;
    cursor_pointers(cur, unused_xi, flux);
    __shared__ double v[3];
    const int b = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int r = 3 * b + w;
    const int lane = threadIdx.x & 63;
    const double* row = D + (size_t)r * ld;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += (k * row[c]) * F[c];
    s = wave_sum(s);
    if (lane == 0) v[w] = s * mult;
    __syncthreads();
    if (threadIdx.x == 0) move_bead(pos, v, b, f, label, flux);
}

__global__ void det_diag_kernel(const double* __restrict__ Dd, const double* __restrict__ F, double k, double mult,
                                int N, double* pos, FluxPlane f, const int* label, long long* flux,
                                const PbxStepCursor* cur) {
    const double* unused_xi = nullptr;
    cursor_pointers(cur, unused_xi, flux);
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    double v[3];
    for (int c = 0; c < 3; ++c) v[c] = ((k * Dd[3 * b + c]) * F[3 * b + c]) * mult;
    move_bead(pos, v, b, f, label, flux);
}

// r += sign * v for all beads (undo of a rejected sub-step, midpoint drift)
__global__ void translate_kernel(const double* __restrict__ vin, double sign, int N, double* pos, FluxPlane f,
                                 const int* label, long long* flux) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    double v[3] = {sign * vin[3 * b], sign * vin[3 * b + 1], sign * vin[3 * b + 2]};
    move_bead(pos, v, b, f, label, flux);
}

// bead.py:170-183 keep_in_box: wrap into [-L/2, L/2)
// With sel non-null the wrap is part of a batch of K move attempts queued
// before the host knows its outcome: it applies only when the batch accepted
// a move (*sel < K); a rejected batch leaves the translated-back positions.
// The outcome, the status word, the flux counters, the energy and the
// wrapped positions are written to the host-mapped buffers of `o`.
__global__ void wrap_kernel(double* pos, int n3, double L, const int* __restrict__ sel, int K, PbxWrapOut o) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        if (o.status && sel) o.status[0] = *sel;
        if (o.status && o.info) o.status[1] = *o.info;
        if (o.E) *o.E = *o.E_dev;
    }
    if (o.flux && t < o.nl) {
        o.flux[t] = o.flux_dev[t];
        if (o.flux_clear) o.flux_clear[t] = 0;
    }
    if (t >= n3 || (sel && *sel >= K)) return;
    double x = pos[t];
    if (isfinite(x)) {
        if (fabs(x) > 64.0 * L) x = fmod(x, L);
        while (x < -L / 2) x += L;
        while (x >= L / 2) x -= L;
        pos[t] = x;
    }
    if (o.pos) o.pos[t] = x;
}

// y = (L x) * scale for lower-triangular L, one wave per row
__global__ __launch_bounds__(256) void lower_gemv_kernel(const double* __restrict__ Lm, int n, int ld,
                                                         const double* __restrict__ x, double* y, double scale) {
    int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    double s = row_dot(Lm + (size_t)r * ld, x, r + 1);
    if ((threadIdx.x & 63) == 0) y[r] = s * scale;
}

// y = alpha * (D x - Lo z): full D, lower-triangular Lo (midpoint stochastic
// drift, box.py:308: m/2 sqrt(2dt) (D_mid B0^{-T} - B0) xi with x = B0^{-T} xi)
__global__ __launch_bounds__(256) void drift_kernel(const double* __restrict__ D, const double* __restrict__ Lo,
                                                    int n, int ld, const double* __restrict__ x,
                                                    const double* __restrict__ z, double alpha, double* y) {
    int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    double s1 = row_dot(D + (size_t)r * ld, x, n);
    double s2 = row_dot(Lo + (size_t)r * ld, z, r + 1);
    if ((threadIdx.x & 63) == 0) y[r] = alpha * (s1 - s2);
}

// y = alpha * (D1 - D0) x over full rows: the midpoint's stochastic drift in
// the mixed-precision engine, alpha (D_mid B0^{-T} - B0) xi rewritten with
// B0 xi = B0 B0^T B0^{-T} xi = D0 x (x = B0^{-T} xi).  Equal in exact
// arithmetic; without the cancellation of two nearly equal products, which
// would multiply the FP32-bulk factor's residual B0 B0^T - D0 by alpha.
__global__ __launch_bounds__(256) void drift_diff_kernel(const double* __restrict__ D1, const double* __restrict__ D0,
                                                         int n, int ld, const double* __restrict__ x, double alpha,
                                                         double* y) {
    int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += (D1[(size_t)r * ld + c] - D0[(size_t)r * ld + c]) * x[c];
    s = wave_sum(s);
    if (lane == 0) y[r] = alpha * s;
}

// y += beta * (D1 - D0) F over full rows (midpoint deterministic drift)
__global__ __launch_bounds__(256) void det_drift_kernel(const double* __restrict__ D1, const double* __restrict__ D0,
                                                        int n, int ld, const double* __restrict__ F, double beta,
                                                        double* y) {
    int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += (beta * (D1[(size_t)r * ld + c] - D0[(size_t)r * ld + c])) * F[c];
    s = wave_sum(s);
    if (lane == 0) y[r] += s;
}

// y = alpha (c v - B x) for lower-triangular B: the midpoint's stochastic
// drift when D_mid w arrives as c v (c = kBT, v = R_mid^{-1} w from two
// triangular solves) instead of from an explicit D_mid.
__global__ __launch_bounds__(256) void drift_solved_kernel(const double* __restrict__ Bm, int n, int ld,
                                                           const double* __restrict__ v,
                                                           const double* __restrict__ x, double c, double alpha,
                                                           double* y) {
    int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    double s = row_dot(Bm + (size_t)r * ld, x, r + 1);
    if ((threadIdx.x & 63) == 0) y[r] = alpha * (c * v[r] - s);
}

// y = alpha (c v - bx): the same drift with B0 xi already formed (bx)
__global__ void drift_vec_kernel(const double* __restrict__ v, const double* __restrict__ bx, int n, double c,
                                 double alpha, double* y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) y[r] = alpha * (c * v[r] - bx[r]);
}

// y = scale J L J x over the whole order np (J the reversal, L lower
// triangular): y[i] = scale sum_{c <= r} L[r][c] x[np-1-c], r = np-1-i.
// With L the factor of J R J this is B^{-T} x for the noise factor
// B = sqrt(kBT) J L^{-T} J of D = kBT inv(R) (scale 1 / sqrt(kBT)).
__global__ __launch_bounds__(256) void rev_lower_gemv_kernel(const double* __restrict__ Lm, int np, int ld,
                                                             const double* __restrict__ x, double scale,
                                                             double* y) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= np) return;
    const int r = np - 1 - i, lane = threadIdx.x & 63;
    const double* row = Lm + (size_t)r * ld;
    double s0 = 0.0, s1 = 0.0;
    int c = lane;
    for (; c + 64 <= r; c += 128) {
        s0 += row[c] * x[np - 1 - c];
        s1 += row[c + 64] * x[np - 1 - (c + 64)];
    }
    if (c <= r) s0 += row[c] * x[np - 1 - c];
    const double t = wave_sum(s0 + s1);
    if (lane == 0) y[i] = scale * t;
}

// Stochastic sub-step from B xi already formed (bx): r += (bx sqrt(2 dt)) mult
__global__ void stoch_vec_kernel(const double* __restrict__ bx, double sq2dt, double mult, double* pos, double* disp,
                                 int N, FluxPlane f, const int* label, long long* flux) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    double v[3];
    for (int c = 0; c < 3; ++c) {
        v[c] = (bx[3 * b + c] * sq2dt) * mult;
        disp[3 * b + c] = v[c];
    }
    move_bead(pos, v, b, f, label, flux);
}

// ---------------------------------------------------------------------------
// Triangular solves with lower-triangular L (order np = 64 T) given the
// inverses W_k of its diagonal tiles:
//   TR:  L^T y = x,  y_k = W_k^T (x_k - sum_{m>k} L_mk^T y_m)  (rows from the last up)
//   !TR: L y = x,    y_k = W_k   (x_k - sum_{m<k} L_km y_m)    (rows from the first down)
// One workgroup per tile row.  Rows are handed out in the order the
// workgroups start (a ticket counter that grows by T per call), so each
// waits only on rows held by workgroups already running, whatever the
// dispatch order and whatever else occupies the compute units.
//
// Rows pass their y_k to one another as self-validating words: each double
// is stored as two 64-bit words, one 32-bit half beside the call's epoch in
// each (relaxed device-scope atomic stores and loads, each word written and
// read whole).  A reader that finds the epoch in both words holds the value,
// so a hand-off is one L2 round trip, with no flag and no fence.  The epoch
// differs between calls; words from an earlier call never match.
//
// In a row's workgroup, lane o of every wave holds output o.  Wave 0 holds
// the tile of the row published last (the one on the critical path) and W_k
// in registers from the start; waves 1-3 take the row's other tiles in
// publication order, round robin, and pass their partial sums to wave 0
// through LDS.  Products take y_m from the lanes by readlane, one entry at a
// time, in a fixed order.  The critical path per row is one wait, two
// 64-long products and the stores.  Waits are bounded: an exhausted wait sets
// *info to PBX_INFO_DAG_WAIT, and every other wait then gives up.

__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// acc + sum_i a[i] v_i, v_i held by lane i
__device__ __forceinline__ double lane_dot(const double (&a)[64], double v, double acc) {
#pragma unroll
    for (int i = 0; i < 64; ++i) acc = fma(a[i], lane_value(v, i), acc);
    return acc;
}

__device__ __forceinline__ void chain_put(unsigned long long* p, double v, unsigned ep) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long e = (unsigned long long)ep << 32;
    __hip_atomic_store(p, (b & 0xffffffffull) | e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, (b >> 32) | e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every lane of the wave waits for its value (lane o: p[2 o], p[2 o + 1]);
// false when the wait was given up.
__device__ __forceinline__ bool chain_get(const unsigned long long* p, unsigned ep, int* info, double& v) {
    const int o = threadIdx.x & 63;
    int spins = 0;
    while (true) {
        const unsigned long long w0 = __hip_atomic_load(p + 2 * o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long w1 = __hip_atomic_load(p + 2 * o + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = (unsigned)(w0 >> 32) == ep && (unsigned)(w1 >> 32) == ep;
        if (__all(ok)) {
            v = __longlong_as_double((long long)((w0 & 0xffffffffull) | (w1 << 32)));
            return true;
        }
        if ((++spins & 255) == 0) {
            if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
            if (spins >= (1 << 22)) {
                if (o == 0) atomicCAS(info, 0, PBX_INFO_DAG_WAIT);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// a[i] = A_km[o][i]: L_mk^T (TR: lane o reads column o of L_mk, coalesced)
// or L_km (lane o reads row o)
template <bool TR>
__device__ __forceinline__ void chain_tile(const double* __restrict__ Lm, int ld, int k, int m, int o,
                                           double (&a)[64]) {
    if (TR) {
        const double* p = Lm + (size_t)(64 * m) * ld + 64 * k + o;
#pragma unroll
        for (int i = 0; i < 64; ++i) a[i] = p[(size_t)i * ld];
    } else {
        const double2* p = (const double2*)(Lm + (size_t)(64 * k + o) * ld + 64 * m);
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const double2 t = p[j];
            a[2 * j] = t.x;
            a[2 * j + 1] = t.y;
        }
    }
}

// rev: x and y in reversed order (x'[i] = x[np-1-i]), y scaled by yscale
// (the hand-off words carry the unscaled solution)
template <bool TR>
__global__ __launch_bounds__(256) void trsv_chain_kernel(const double* __restrict__ Lm, int np, int ld,
                                                         const double* __restrict__ W,
                                                         const double* __restrict__ x, double* __restrict__ y,
                                                         unsigned long long* words, int* ticket, unsigned ep,
                                                         int* info, int rev, double yscale) {
    __shared__ double part[3][64];
    __shared__ int done;
    __shared__ int row;
    const int T = np / 64;
    if (threadIdx.x == 0) {
        const unsigned t =
            __hip_atomic_fetch_add((unsigned*)ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) % (unsigned)T;
        row = TR ? T - 1 - (int)t : (int)t;
        done = 0;
    }
    __syncthreads();
    const int k = row;
    const int o = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ndep = TR ? T - 1 - k : k;  // tiles this row waits on; dependency q is published q-th
    if (wv == 0) {
        double a[64], wk[64];
        const int mc = TR ? k + 1 : k - 1;  // the row published last
        if (ndep > 0) chain_tile<TR>(Lm, ld, k, mc, o, a);
        const double* Wk = W + (size_t)k * 64 * 64;
        if (TR) {
#pragma unroll
            for (int i = 0; i < 64; ++i) {
                const double t = Wk[i * 64 + o];  // W_k[i][o], nonzero for i >= o
                wk[i] = i >= o ? t : 0.0;
            }
        } else {
            const double2* p = (const double2*)(Wk + o * 64);  // W_k[o][i], nonzero for i <= o
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                const double2 t = p[j];
                wk[2 * j] = 2 * j <= o ? t.x : 0.0;
                wk[2 * j + 1] = 2 * j + 1 <= o ? t.y : 0.0;
            }
        }
        const int idx = rev ? np - 1 - (64 * k + o) : 64 * k + o;
        double r = x[idx];
        if (ndep > 1) {
            int it = 0;
            while (__hip_atomic_load(&done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < 3) {
                if (++it >= (1 << 24)) {
                    if (o == 0) atomicCAS(info, 0, PBX_INFO_DAG_WAIT);
                    return;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            r -= (part[0][o] + part[1][o]) + part[2][o];
        }
        if (ndep > 0) {
            double ym;
            if (!chain_get(words + (size_t)2 * 64 * mc, ep, info, ym)) return;
            r -= lane_dot(a, ym, 0.0);
        }
        const double yk = lane_dot(wk, r, 0.0);
        chain_put(words + (size_t)2 * (64 * k + o), yk, ep);
        y[idx] = yscale * yk;
    } else {
        double acc = 0.0;
        for (int q = wv - 1; q < ndep - 1; q += 3) {
            const int m = TR ? T - 1 - q : q;
            double a[64];
            chain_tile<TR>(Lm, ld, k, m, o, a);
            double ym;
            if (!chain_get(words + (size_t)2 * 64 * m, ep, info, ym)) break;
            acc = lane_dot(a, ym, acc);
        }
        part[wv - 1][o] = acc;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (o == 0) __hip_atomic_fetch_add(&done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// ---------------------------------------------------------------------------
// Batched move attempts (box.py:202-250 loop, evaluated K attempts at a time).
//
// A rejected Ermak attempt is undone by translating back, so attempt k starts
// from base_k = (base_{k-1} + d_{k-1}) - d_{k-1}; the candidate positions of
// K consecutive attempts follow from K displacement vectors d_k = B xi_k
// sqrt(2 dt), which one pass over B yields for all k.  The first candidate
// without an overlap is the accepted move, exactly as in the attempt-by-
// attempt loop.  Net plane crossings over a chain of translations telescope
// to side(end) - side(start) per bead.

constexpr int BT = 64;        // B tile edge
constexpr int BSP = 65;       // LDS stride
constexpr int KC_MAX = 32;    // candidates per batch
constexpr int SLAB_MAX = 12;  // column slabs of the partial products (pbx_attempt_slab_width)

// Partial products P[s][r][k] = sum_{c in slab s, c <= r} B[r][c] xi_k[c] for
// row block i = blockIdx.x, column slab s = blockIdx.y (S tiles wide) and
// candidates KC * blockIdx.z .. + KC - 1 (P holds KCP candidates per row).
// Wave w computes rows 16w..16w+15 of the block for all KC candidates with
// v_mfma_f64_16x16x4_f64 (A = B rows from LDS, B operand = xi tile from LDS).
template <int KC>
__global__ __launch_bounds__(256) void stoch_batch_kernel(const double* __restrict__ B, int ld, int np, int n,
                                                          const double* __restrict__ ring, int ring_pos, int ring_att,
                                                          int K, int S, int KCP, double* __restrict__ P) {
    typedef double v4d __attribute__((ext_vector_type(4)));
    typedef double v2d __attribute__((ext_vector_type(2)));
    __shared__ double sB[BT * BSP];
    __shared__ double sX[KC * BSP];
    const int i = blockIdx.x, sl = blockIdx.y;
    const int kz = blockIdx.z * KC;  // first candidate of this workgroup
    const int j0 = sl * S;
    if (j0 > i) return;
    const int j1 = min(j0 + S, i + 1);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    v4d acc[KC / 16];
#pragma unroll
    for (int cb = 0; cb < KC / 16; ++cb) acc[cb] = (v4d){0.0, 0.0, 0.0, 0.0};
    const double* Bblk = B + (size_t)i * BT * ld;
    // global loads of tile j into registers (B tile and its xi block)
    v2d vb[8];
    double vx[KC / 4];
    auto fetch = [&](int j) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int e = tid + it * 256;
            vb[it] = *(const v2d*)(Bblk + (size_t)(e >> 5) * ld + (size_t)j * BT + (e & 31) * 2);
        }
#pragma unroll
        for (int it = 0; it < KC / 4; ++it) {
            const int e = tid + it * 256;
            const int k = kz + (e >> 6), c = j * BT + (e & 63);
            int slot = ring_pos + k;
            if (slot >= ring_att) slot -= ring_att;
            vx[it] = (k < K && c < n) ? ring[(size_t)slot * n + c] : 0.0;
        }
    };
    fetch(j0);
    for (int j = j0; j < j1; ++j) {
        __syncthreads();  // previous tile fully consumed
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int e = tid + it * 256;
            const int r = e >> 5, c = (e & 31) * 2;
            // diagonal tile: only c <= r belongs to the lower-triangular B
            sB[r * BSP + c] = (j == i && c > r) ? 0.0 : vb[it].x;
            sB[r * BSP + c + 1] = (j == i && c + 1 > r) ? 0.0 : vb[it].y;
        }
#pragma unroll
        for (int it = 0; it < KC / 4; ++it) {
            const int e = tid + it * 256;
            sX[(e >> 6) * BSP + (e & 63)] = vx[it];
        }
        __syncthreads();
        if (j + 1 < j1) fetch(j + 1);  // in flight during this tile's products
#pragma unroll 4
        for (int c = 0; c < BT; c += 4) {
            const double a = sB[(16 * w + li) * BSP + c + lk];
#pragma unroll
            for (int cb = 0; cb < KC / 16; ++cb)
                acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sX[(16 * cb + li) * BSP + c + lk], acc[cb], 0, 0, 0);
        }
    }
    // accumulator register v of lane group lk holds row lk + 4v, column li
    double* Ps = P + (size_t)sl * np * KCP + kz;
#pragma unroll
    for (int cb = 0; cb < KC / 16; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = i * BT + 16 * w + lk + 4 * v;
            Ps[(size_t)r * KCP + 16 * cb + li] = acc[cb][v];
        }
}

// Candidate positions of K attempts from the current positions.  A
// workgroup takes 4 beads; thread (bead, k) forms d_k for the bead's three
// coordinates: the slab partials summed in slab order, times sqrt(2 dt)
// (B xi, box.py:371), or ((Bd xi) sqrt(2 dt)) for the diagonal nohi factor
// (box.py:369).  Then per coordinate the chain cand[k] = base + d_k,
// base <- cand[k] - d_k; row K of cand keeps the chain end.  delta[b] bounds
// the distance of every candidate position of bead b from its current one
// (the overlap kernel's pair filter).  Block 0 also clears the overlap mask.
constexpr int CB = 4;  // beads per workgroup (32 threads each)
__global__ __launch_bounds__(128) void cand_kernel(const double* __restrict__ pos, const double* __restrict__ P, int np,
                                                   int KC, int S, int N, int K, double sq2dt,
                                                   const double* __restrict__ Bd, const double* __restrict__ ring,
                                                   int ring_pos, int ring_att, double* __restrict__ cand,
                                                   double* __restrict__ delta, int* __restrict__ cflag) {
    // d_k is rounded on its own before it is added, as the reference forms
    // the displacement array first (no fused multiply-add into base + d)
#pragma clang fp contract(off)
    __shared__ double sd[CB][3][KC_MAX + 1];
    __shared__ double sm[CB][3];
    const int tid = threadIdx.x;
    const int k = tid & 31, bl = tid >> 5;
    const int b = blockIdx.x * CB + bl;
    const int n = 3 * N;
    if (blockIdx.x == 0 && tid == 0) cflag[0] = 0;
    if (b < N && k < K) {
        int slot = ring_pos + k;
        if (slot >= ring_att) slot -= ring_att;
        if (Bd) {
            for (int c = 0; c < 3; ++c) {
                const int r = 3 * b + c;
                sd[bl][c][k] = ((Bd[r] * ring[(size_t)slot * n + r]) * sq2dt) * 1.0;
            }
        } else {
            // every slab partial requested before the first is added (the
            // sums keep slab order)
            double v[3][SLAB_MAX];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int r = 3 * b + c, ns = (r / BT) / S + 1;
#pragma unroll
                for (int s = 0; s < SLAB_MAX; ++s) v[c][s] = s < ns ? P[((size_t)s * np + r) * KC + k] : 0.0;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int r = 3 * b + c, ns = (r / BT) / S + 1;
                double acc = 0.0;
#pragma unroll
                for (int s = 0; s < SLAB_MAX; ++s)
                    if (s < ns) acc += v[c][s];
                sd[bl][c][k] = (acc * sq2dt) * 1.0;
            }
        }
    }
    __syncthreads();
    // chain: thread (bead, coordinate) walks the K candidates
    if (k < 3 && b < N) {
        const int r = 3 * b + k;
        const double p0 = pos[r];
        double base = p0, far = 0.0;
        for (int q = 0; q < K; ++q) {
            const double d = sd[bl][k][q];
            const double c = base + d;
            sd[bl][k][q] = c;
            far = fmax(far, fabs(c - p0));
            base = c - d;
        }
        sd[bl][k][K] = base;
        sm[bl][k] = fmax(far, fabs(base - p0));
    }
    __syncthreads();
    if (k == 0 && b < N) {
        const double m0 = sm[bl][0], m1 = sm[bl][1], m2 = sm[bl][2];
        delta[b] = sqrt(m0 * m0 + m1 * m1 + m2 * m2) * (1.0 + 1e-12);
    }
    // coalesced write-out: for each candidate row, the 3 * CB contiguous
    // coordinates of this workgroup's beads
    const int nb = min(CB, N - blockIdx.x * CB);
    for (int e = tid; e < (K + 1) * 3 * CB; e += 128) {
        const int q = e / (3 * CB), o = e % (3 * CB);
        const int bb = o / 3, c = o % 3;
        if (bb < nb) cand[(size_t)q * n + 3 * (blockIdx.x * CB) + o] = sd[bb][c][q];
    }
}

// Overlap flags of K candidates: 16 x 16 bead tiles (pbx_pair_tile), one
// pair per thread; bit k of cflag[0] is set when candidate k overlaps; pairs
// marked in `mask` (Nt x Nt, may be null) are exempt, as in bead.c:180.
// Beads N..Nt-1 are immobile: the same position in every candidate, no
// displacement, and no pairs among themselves.  A pair can overlap in some
// candidate only if its current minimum-image distance is at most
// S + thr + delta_i + delta_j (the torus distance moves by at most the two
// displacements, and the reference test flags only pairs within S + thr);
// the workgroup lists those pairs in LDS and its threads then take the
// (pair, candidate) tests together, every candidate position requested at
// once.
__global__ __launch_bounds__(256) void overlaps_batch_kernel(const double* __restrict__ cand,
                                                             const double* __restrict__ pos,
                                                             const double* __restrict__ delta,
                                                             const double* __restrict__ a, int N, int Nt, int K,
                                                             double L, double thr, const int* __restrict__ mask,
                                                             int* cflag) {
    __shared__ int s_np;
    __shared__ unsigned s_hit;
    __shared__ int s_i[256], s_j[256];
    int bi, bj;
    pbx_pair_tile(blockIdx.x, (N + 15) / 16, bi, bj);
    const int i = bi * 16 + (threadIdx.x >> 4), j = bj * 16 + (threadIdx.x & 15);
    if (threadIdx.x == 0) {
        s_np = 0;
        s_hit = 0;
    }
    __syncthreads();
    // j < i and j < N: j moves; i moves when i < N.  mask: the reference's
    // connection matrix (bonded pairs are not tested)
    if (i < Nt && j < i && j < N && !(mask && mask[(size_t)j + (size_t)Nt * i])) {
        double s2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double d = pbx_min_image(pos[3 * i + c] - pos[3 * j + c], L);
            s2 += d * d;
        }
        const double reach = a[i] + a[j] + thr + (i < N ? delta[i] : 0.0) + delta[j];
        // a relative margin far above the rounding of these few operations
        if (!(sqrt(s2) > reach * (1.0 + 1e-9) + 1e-9 * L)) {
            const int q = atomicAdd(&s_np, 1);
            s_i[q] = i;
            s_j[q] = j;
        }
    }
    __syncthreads();
    const int np_ = s_np;
    if (np_ == 0) return;
    const int n = 3 * N;
    unsigned m = 0;  // bit k: some listed pair overlaps in candidate k
    for (int it = threadIdx.x; it < np_ * K; it += 256) {
        const int q = it / K, k = it - q * K;
        const int pi = s_i[q], pj = s_j[q];
        const double* C = cand + (size_t)k * n;
        const double* Ri = pi >= N ? pos + 3 * pi : C + 3 * pi;
        if (pbx_pair_overlaps(Ri[0] - C[3 * pj], Ri[1] - C[3 * pj + 1], Ri[2] - C[3 * pj + 2], a[pi], a[pj], L, thr))
            m |= 1u << k;
    }
    if (m) atomicOr(&s_hit, m);
    __syncthreads();
    if (threadIdx.x == 0 && s_hit) atomicOr((unsigned*)cflag, s_hit);
}

// Apply the first candidate without an overlap (or the chain end when all K
// overlapped): positions, per-label net crossings, and the index in *sel.
__global__ void select_kernel(const double* __restrict__ cand, int N, int K, const int* __restrict__ cflag, double* pos,
                              FluxPlane f, const int* __restrict__ label, long long* flux, int* sel) {
    const unsigned all = K >= 32 ? 0xffffffffu : ((1u << K) - 1u);
    const unsigned clean = ~(unsigned)cflag[0] & all;
    const int acc = clean ? __builtin_ctz(clean) : K;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) *sel = acc;
    if (b >= N) return;
    const double* src = cand + (size_t)acc * 3 * N;
    const double x0 = pos[3 * b], y0 = pos[3 * b + 1], z0 = pos[3 * b + 2];
    const double x1 = src[3 * b], y1 = src[3 * b + 1], z1 = src[3 * b + 2];
    pos[3 * b] = x1;
    pos[3 * b + 1] = y1;
    pos[3 * b + 2] = z1;
    if (f.on) {
        const int s0 = plane_side(f, x0, y0, z0), s1 = plane_side(f, x1, y1, z1);
        if (s0 != s1) atomicAdd((unsigned long long*)&flux[label[b]], (unsigned long long)(long long)(s1 - s0));
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers used by the simulation driver (pbx_api.hip)

struct PbxFlux {
    FluxPlane f;
};

int pbx_launch_stoch(const double* B, int ld, int diag_only, const double* xi, double sq2dt, double mult,
                     double* pos, double* disp, int N, const double* plane_n, const double* plane_p, int flux_on,
                     const int* label, long long* flux, hipStream_t s, const PbxStepCursor* cur) {
    FluxPlane f;
    for (int c = 0; c < 3; ++c) {
        f.n[c] = plane_n ? plane_n[c] : 0.0;
        f.p[c] = plane_p ? plane_p[c] : 0.0;
    }
    f.on = flux_on;
    if (diag_only)
        PBX_LAUNCH(stoch_diag_kernel, dim3((N + 255) / 256), dim3(256), 0, s, B, xi, sq2dt, mult, pos, disp,
                           N, f, label, flux, cur);
    else
        PBX_LAUNCH(stoch_kernel, dim3(N), dim3(192), 0, s, B, ld, xi, sq2dt, mult, pos, disp, f, label,
                           flux, cur);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_det(const double* D, int ld, int diag_only, const double* F, double k, double mult, double* pos,
                   int N, const double* plane_n, const double* plane_p, int flux_on, const int* label,
                   long long* flux, hipStream_t s, const PbxStepCursor* cur) {
    FluxPlane f;
    for (int c = 0; c < 3; ++c) {
        f.n[c] = plane_n ? plane_n[c] : 0.0;
        f.p[c] = plane_p ? plane_p[c] : 0.0;
    }
    f.on = flux_on;
    if (diag_only)
        PBX_LAUNCH(det_diag_kernel, dim3((N + 255) / 256), dim3(256), 0, s, D, F, k, mult, N, pos, f, label,
                           flux, cur);
    else
        PBX_LAUNCH(det_kernel, dim3(N), dim3(192), 0, s, D, ld, F, k, mult, 3 * N, pos, f, label, flux, cur);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

__global__ void cursor_advance_kernel(PbxStepCursor* cur) {
    if (threadIdx.x == 0) cur->step += 1;
}

int pbx_launch_cursor_advance(PbxStepCursor* cur, hipStream_t s) {
    PBX_LAUNCH(cursor_advance_kernel, dim3(1), dim3(64), 0, s, cur);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// ---- divergence term by finite differences (box.py:560-610) -------------

__global__ void div_shift_kernel(double* pos, int idx, double delta) {
    if (threadIdx.x == 0) pos[idx] += delta;
}

// div[r] = -sum_c D[r][c] (one wave per row)
__global__ void div_init_kernel(const double* D, int n, int ld, double* div) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= n) return;
    double acc = 0.0;
    for (int c = lane; c < n; c += 64) acc += D[(size_t)r * ld + c];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane == 0) div[r] = -acc;
}

// div[c] += D[row][c]
__global__ void div_add_row_kernel(const double* D, int row, int n, int ld, double* div) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) div[c] += D[(size_t)row * ld + c];
}

// out[c] = (div[c] / eps) * dt
__global__ void div_finish_kernel(const double* div, int n, double eps, double dt, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) out[c] = (div[c] / eps) * dt;
}

int pbx_launch_div_shift(double* pos, int idx, double delta, hipStream_t s) {
    PBX_LAUNCH(div_shift_kernel, dim3(1), dim3(64), 0, s, pos, idx, delta);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_div_init(const double* D, int n, int ld, double* div, hipStream_t s) {
    PBX_LAUNCH(div_init_kernel, dim3((n + 3) / 4), dim3(256), 0, s, D, n, ld, div);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_div_add_row(const double* D, int row, int n, int ld, double* div, hipStream_t s) {
    PBX_LAUNCH(div_add_row_kernel, dim3((n + 255) / 256), dim3(256), 0, s, D, row, n, ld, div);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_div_finish(const double* div, int n, double eps, double dt, double* out, hipStream_t s) {
    PBX_LAUNCH(div_finish_kernel, dim3((n + 255) / 256), dim3(256), 0, s, div, n, eps, dt, out);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_translate(const double* v, double sign, double* pos, int N, const double* plane_n,
                         const double* plane_p, int flux_on, const int* label, long long* flux, hipStream_t s) {
    FluxPlane f;
    for (int c = 0; c < 3; ++c) {
        f.n[c] = plane_n ? plane_n[c] : 0.0;
        f.p[c] = plane_p ? plane_p[c] : 0.0;
    }
    f.on = flux_on;
    PBX_LAUNCH(translate_kernel, dim3((N + 255) / 256), dim3(256), 0, s, v, sign, N, pos, f, label, flux);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_wrap(double* pos, int N, double L, hipStream_t s) {
    PBX_CHECK_BOX(L);
    int n3 = 3 * N;
    PBX_LAUNCH(wrap_kernel, dim3((n3 + 255) / 256), dim3(256), 0, s, pos, n3, L, nullptr, 0, PbxWrapOut{});
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_wrap_publish(double* pos, int N, double L, const int* sel, int K, const PbxWrapOut& o,
                            hipStream_t s) {
    PBX_CHECK_BOX(L);
    const int n3 = 3 * N, nt = o.flux && o.nl > n3 ? o.nl : n3;
    PBX_LAUNCH(wrap_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, pos, n3, L, sel, K, o);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_lower_gemv(const double* Lm, int n, int ld, const double* x, double* y, double scale, hipStream_t s) {
    PBX_LAUNCH(lower_gemv_kernel, dim3((n + 3) / 4), dim3(256), 0, s, Lm, n, ld, x, y, scale);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_drift(const double* D, const double* Lo, int n, int ld, const double* x, const double* z, double alpha,
                     double* y, hipStream_t s) {
    PBX_LAUNCH(drift_kernel, dim3((n + 3) / 4), dim3(256), 0, s, D, Lo, n, ld, x, z, alpha, y);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_drift_diff(const double* D1, const double* D0, int n, int ld, const double* x, double alpha, double* y,
                          hipStream_t s) {
    PBX_LAUNCH(drift_diff_kernel, dim3((n + 3) / 4), dim3(256), 0, s, D1, D0, n, ld, x, alpha, y);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_det_drift(const double* D1, const double* D0, int n, int ld, const double* F, double beta, double* y,
                         hipStream_t s) {
    PBX_LAUNCH(det_drift_kernel, dim3((n + 3) / 4), dim3(256), 0, s, D1, D0, n, ld, F, beta, y);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

void pbx_status_sequence_queued(int* info_dev, hipStream_t s);

int pbx_launch_trsv_chain(const double* Lm, int np, int ld, const double* W, const double* x, double* y,
                          unsigned long long* words, int* ticket, unsigned epoch, int trans, int* info,
                          hipStream_t s, int rev, double yscale) {
    if (np <= 0 || np % 64 || ld < np || ld % 2 || epoch == 0) {
        pbx_set_error(PBX_ERR_ARG, "trsv_chain: np=%d ld=%d epoch=%u", np, ld, epoch);
        return PBX_ERR_ARG;
    }
    if (trans)
        PBX_LAUNCH(trsv_chain_kernel<true>, dim3(np / 64), dim3(256), 0, s, Lm, np, ld, W, x, y, words, ticket, epoch,
                   info, rev, yscale);
    else
        PBX_LAUNCH(trsv_chain_kernel<false>, dim3(np / 64), dim3(256), 0, s, Lm, np, ld, W, x, y, words, ticket,
                   epoch, info, rev, yscale);
    pbx_status_sequence_queued(info, s);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_drift_vec(const double* v, const double* bx, int n, double c, double alpha, double* y, hipStream_t s) {
    PBX_LAUNCH(drift_vec_kernel, dim3((n + 255) / 256), dim3(256), 0, s, v, bx, n, c, alpha, y);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_rev_lower_gemv(const double* Lm, int np, int ld, const double* x, double scale, double* y,
                              hipStream_t s) {
    PBX_LAUNCH(rev_lower_gemv_kernel, dim3((np + 3) / 4), dim3(256), 0, s, Lm, np, ld, x, scale, y);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_stoch_vec(const double* bx, double sq2dt, double mult, double* pos, double* disp, int N,
                         const double* plane_n, const double* plane_p, int flux_on, const int* label,
                         long long* flux, hipStream_t s) {
    FluxPlane f;
    for (int c = 0; c < 3; ++c) {
        f.n[c] = plane_n ? plane_n[c] : 0.0;
        f.p[c] = plane_p ? plane_p[c] : 0.0;
    }
    f.on = flux_on;
    PBX_LAUNCH(stoch_vec_kernel, dim3((N + 127) / 128), dim3(128), 0, s, bx, sq2dt, mult, pos, disp, N, f, label,
               flux);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_drift_solved(const double* Bm, int n, int ld, const double* v, const double* x, double c, double alpha,
                            double* y, hipStream_t s) {
    PBX_LAUNCH(drift_solved_kernel, dim3((n + 3) / 4), dim3(256), 0, s, Bm, n, ld, v, x, c, alpha, y);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// ---------------------------------------------------------------------------
// batched attempts

int pbx_attempt_slab_width(int np) {
    const int T = np / BT;
    int S = (T + 11) / 12;
    return S < 2 ? 2 : S;
}

size_t pbx_attempt_partial_doubles(int np) {
    const int T = np / BT;
    const int S = pbx_attempt_slab_width(np);
    const int nslab = (T + S - 1) / S;
    return (size_t)nslab * np * KC_MAX;
}

int pbx_attempt_kmax() { return KC_MAX; }

int pbx_launch_attempt_batch(const double* B, const double* Bd, int ld, int np, int N, int Nt, const double* pos,
                             const double* ring, int ring_pos, int ring_att, int K, double sq2dt, double* P,
                             double* cand, int* cflag, const double* a, double L, double thr, const int* mask,
                             hipStream_t s) {
    if (K < 1 || K > KC_MAX || ring_pos < 0 || ring_pos >= ring_att || Nt < N) {
        pbx_set_error(PBX_ERR_ARG, "attempt batch: K=%d ring_pos=%d ring_att=%d N=%d Nt=%d", K, ring_pos, ring_att, N,
                      Nt);
        return PBX_ERR_ARG;
    }
    PBX_CHECK_BOX(L);
    const int n = 3 * N;
    const int S = pbx_attempt_slab_width(np);
    const int KC = K <= 16 ? 16 : 32;
    if (!Bd) {
        if (np % BT || n > np) {
            pbx_set_error(PBX_ERR_ARG, "attempt batch: bad order np=%d n=%d", np, n);
            return PBX_ERR_ARG;
        }
        const int T = np / BT;
        if ((T + S - 1) / S > SLAB_MAX) {
            pbx_set_error(PBX_ERR_ARG, "attempt batch: %d column slabs exceed %d", (T + S - 1) / S, SLAB_MAX);
            return PBX_ERR_ARG;
        }
        // 16 candidates per workgroup: a batch of 32 runs as two workgroup
        // sets (twice the workgroups in flight over the same B tiles)
        dim3 grid(T, (T + S - 1) / S, KC / 16);
        PBX_LAUNCH(stoch_batch_kernel<16>, grid, dim3(256), 0, s, B, ld, np, n, ring, ring_pos, ring_att, K, S, KC,
                   P);
        PBX_HIP(hipGetLastError());
    }
    double* delta = cand + (size_t)(K + 1) * n;
    PBX_LAUNCH(cand_kernel, dim3((N + CB - 1) / CB), dim3(32 * CB), 0, s, pos, P, np, KC, S, N, K, sq2dt, Bd, ring,
                       ring_pos, ring_att, cand, delta, cflag);
    PBX_HIP(hipGetLastError());
    if (Nt > 1) {
        PBX_LAUNCH(overlaps_batch_kernel, dim3(pbx_pair_tiles(N, Nt)), dim3(256), 0, s, cand, pos, delta, a, N,
                           Nt, K, L, thr, mask, cflag);
        PBX_HIP(hipGetLastError());
    }
    return PBX_OK;
}

int pbx_launch_attempt_select(const double* cand, int N, int K, const int* cflag, double* pos, const double* plane_n,
                              const double* plane_p, int flux_on, const int* label, long long* flux, int* sel,
                              hipStream_t s) {
    FluxPlane f;
    for (int c = 0; c < 3; ++c) {
        f.n[c] = plane_n ? plane_n[c] : 0.0;
        f.p[c] = plane_p ? plane_p[c] : 0.0;
    }
    f.on = flux_on;
    PBX_LAUNCH(select_kernel, dim3((N + 255) / 256), dim3(256), 0, s, cand, N, K, cflag, pos, f, label, flux,
                       sel);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// Shared definitions for the MI355X (gfx950) hydrodynamic-step kernels.
//
// Storage conventions used by every kernel in this library:
//  * bead data: xyz as N x 3 doubles (bead-major), radii as N doubles;
//  * dense 3N x 3N matrices are row-major with leading dimension `ld`, padded
//    to `np = round_up(3N, 64)` rows/columns; the pad holds the identity
//    (zero off-diagonal, one on the diagonal) so blocked factorisations run
//    on whole 64 x 64 tiles with no bounds checks and leave the pad intact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define PBX_TILE 64

// status codes returned by the C ABI (0 = success)
#define PBX_OK 0
#define PBX_ERR_ARG -1
#define PBX_ERR_HIP -2
#define PBX_ERR_OOM -3
#define PBX_ERR_MOVES -4     // max_move_attempts exceeded
#define PBX_ERR_OVERLAP -5   // starting configuration overlaps
#define PBX_ERR_STATE -6
// k > 0: the Cholesky factorisation met a non-positive pivot at column k-1
// (except PBX_INFO_DAG_WAIT: a bounded wait between workgroups of a
// factorisation or of the B0^T solve ran out of budget; the kernels drained
// and their result is not valid)
#define PBX_INFO_DAG_WAIT 0x40000000

static inline size_t pbx_round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

extern "C" void pbx_set_error(int code, const char* fmt, ...);
extern "C" void pbx_clear_error(void);

// Every kernel launch of the library goes through PBX_LAUNCH, which also
// counts it (pbx_kernel_launches): launches per step are what bounds the
// step rate of small simulations.
void pbx_count_launch();
#define PBX_LAUNCH(...)              \
    do {                             \
        pbx_count_launch();          \
        hipLaunchKernelGGL(__VA_ARGS__); \
    } while (0)

#define PBX_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            pbx_set_error(PBX_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,          \
                          hipGetErrorString(e_));                                          \
            return PBX_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

// ---------------------------------------------------------------------------
// device helpers

// Minimum-image reduction with the reference's loop semantics
// (src/bdsim/clib/bead.c:44-50); the result lies in (-L/2, L/2].  Callers
// guarantee L > 0; non-finite d is returned unchanged and |d| > 64 L is first
// reduced with fmod, so every loop ends after at most 65 iterations.
__device__ __forceinline__ double pbx_min_image(double d, double L) {
    if (!isfinite(d)) return d;
    if (fabs(d) > 64.0 * L) d = fmod(d, L);
    const double h = L / 2.0;
    while (d >= h) d -= L;
    while (d <= -h) d += L;
    return d;
}

// clib/bead.c:169-260 pair test: overlap iff no raw axis difference lies in
// the open skip band (S, L-S) / (-(L-S), -S) and the minimum-image distance
// is at most S + threshold (S = ai + aj).  d = r_i - r_j.
__device__ __forceinline__ bool pbx_pair_overlaps(double dx, double dy, double dz, double ai, double aj, double L,
                                                  double thr) {
#pragma clang fp contract(off)
    double d[3] = {dx, dy, dz};
    const double S = ai + aj;
    const double Sp = L - S;
    for (int c = 0; c < 3; ++c)
        if ((d[c] > S && d[c] < Sp) || (d[c] < -S && d[c] > -Sp)) return false;
    for (int c = 0; c < 3; ++c) d[c] = pbx_min_image(d[c], L);
    const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    return dist <= ai + aj + thr;
}

// Tiles of 16 x 16 bead pairs for an overlap test over Nt beads of which the
// first Nm move (the rest are immobile, box.py:649-668): the lower triangle
// of the Tm = ceil(Nm / 16) mobile tile rows, then Tt - Tm further tile rows
// of Tm tiles each.  Every pair (i, j) with j < i and j < Nm lies in exactly
// one tile; pairs of two immobile beads, which the reference's connection
// matrix exempts (bead.py:401-413), lie in none except the few sharing a
// tile with mobile ones (callers test j < Nm).
__device__ __forceinline__ void pbx_pair_tile(int t, int Tm, int& bi, int& bj) {
    const int tri = Tm * (Tm + 1) / 2;
    if (t >= tri) {
        t -= tri;
        bi = Tm + t / Tm;
        bj = t % Tm;
        return;
    }
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    bi = r;
    bj = t - r * (r + 1) / 2;
}
static inline int pbx_pair_tiles(int Nm, int Nt) {
    const int Tm = (Nm + 15) / 16, Tt = (Nt + 15) / 16;
    return Tm * (Tm + 1) / 2 + (Tt - Tm) * Tm;
}

// Host-side check shared by every launcher that reduces by a box length.
static inline bool pbx_box_ok(double L) { return L > 0.0 && L < 1e300; }
#define PBX_CHECK_BOX(L)                                                                   \
    do {                                                                                   \
        if (!pbx_box_ok(L)) {                                                              \
            pbx_set_error(PBX_ERR_ARG, "box_length must be positive and finite (%g)", (L)); \
            return PBX_ERR_ARG;                                                            \
        }                                                                                  \
    } while (0)

// ---------------------------------------------------------------------------
// launchers shared across translation units

// pair assembly (pbx_pairs.hip)
// sel/selK: skip the launch on the device when *sel >= selK (null: never)
int pbx_launch_rpy(const double* xyz, const double* a, int N, double L, double scale, double* M, int ld,
                   int write_pad, hipStream_t s, const int* sel = nullptr, int selK = 0);
int pbx_launch_smith(const double* xyz, const double* a, int N, double L, double alpha, int m_max, int n_max,
                     double scale, double* M, int ld, int write_pad, hipStream_t s, const int* sel = nullptr,
                     int selK = 0);
int pbx_launch_rpy_packed(const double* a, const double* P, int N, double* M, int ld, hipStream_t s);
int pbx_launch_smith_packed(const double* a, const double* P, int N, double L, double alpha, int m_max, int n_max,
                            double* M, int ld, hipStream_t s);
int pbx_launch_pointers(const double* xyz, int N, double L, double* P3, hipStream_t s);
int pbx_launch_immobile_pointers(const double* xm, const double* xi, int nm, int ni, double L, double* P3,
                                 hipStream_t s);
// overlap test over Nt beads whose first Nm move; mask (Nt x Nt) or null
int pbx_launch_overlaps(const double* xyz, const double* a, int Nt, int Nm, double L, double thr, const int* mask,
                        int* flag, hipStream_t s);

// lubrication (pbx_lub.hip): C = ((dR * mul) / div) + C
int pbx_launch_jo_correction(const double* xyz, const double* P, const double* a, int N, double L, double cutoff,
                             int cichocki, double mul, double div, double* C, int ld, hipStream_t s);
int pbx_launch_jo_2b(double ai, double aj, double rx, double ry, double rz, double* out18, hipStream_t s);

// dense linear algebra (pbx_dense.hip); np must be a multiple of 64
// mixed != 0: the trailing updates on FP32 matrix cores (operands rounded to
// FP32, FP32 accumulation, added into the FP64 matrix); diagonal tiles and
// solves in FP64.
int pbx_potrf(double* A, int np, int ld, double* work, int* info_dev, hipStream_t s, int mixed = 0);
// Force field of the step (pbx_forces.hip): per-bead incidence lists hold
// entries 2*bond+role (role 0: first bead) and 4*angle+role (roles 0..2).
// Bead indices >= N are immobile (no force row; box.py:649-668).
struct PbxForceArgs {
    int N;   // mobile beads
    int Nt;  // mobile + immobile
    double L;
    double fex[3];
    int reg_on[3];
    double reg_lo[3], reg_hi[3];
    int lj;  // 0 none, 1 LJ-6, 2 LJ-12, 3 LJ-6-12
    double alpha;
    const double* hc;
    const double* eps;
    const int* bond_ptr;
    const int* bond_ent;
    const int* bond_ij;
    const double* bond_par;
    const int* ang_ptr;
    const int* ang_ent;
    const int* ang_ijk;
    const double* ang_par;
    double unit;
};
int pbx_launch_forces(const PbxForceArgs& a, const double* pos, double* F, double* e_part, double* E,
                      hipStream_t s);

int pbx_factor_residual(const double* D, const double* B, int n, int np, int ld, double* acc, double* rel,
                        hipStream_t s);
// A := scale * inv(A); precision 0 FP64, 1 FP32 matrix-core bulk, 2 as 1 plus
// one Newton-Schulz step with an FP64 residual (C: np x np scratch, else null)
int pbx_spd_inverse(double* A, int np, int ld, double* work, double* X, int* info_dev, hipStream_t s,
                    double scale, int precision = 0, double* C = nullptr);
int pbx_fill_pad(double* A, int n, int np, int ld, hipStream_t s);
// D = kBT inv(R) (into R) and its Cholesky factor B (with the inverses of
// its diagonal tiles in Wc) from one factorisation of the reversed R;
// form_D 0: B only, R left holding inv(L) for pbx_inverse_and_factor_finish
int pbx_inverse_and_factor(double* R, double* X, double* B, double* Wc, double* Wi, int n, int np, int ld,
                           double kBT, int* info_dev, hipStream_t s, int form_D = 1);
int pbx_inverse_and_factor_finish(double* R, double* X, int n, int np, int ld, double kBT, hipStream_t s);
// L := chol(J R J) (pad diagonal kBT), tile inverses in W; then, when B
// itself is wanted, B := sqrt(kBT) J inv(L)^T J in place of L, W := B's tile
// inverses, Xout := inv(L)
int pbx_factor_reversed(const double* R, double* L, double* W, int n, int np, int ld, double kBT, int* info_dev,
                        hipStream_t s);
int pbx_factor_reversed_finish(double* L, double* W, double* Xout, int np, int ld, double kBT, hipStream_t s);
int pbx_copy_lower(const double* src, double* dst, int np, int ld, hipStream_t s);
int pbx_potrf_copy(const double* D, double* B, int np, int ld, double* work, int* info_dev, hipStream_t s,
                   int mixed);
size_t pbx_dense_work_doubles(int np);
// The factorisation's work area: np x 64 doubles of diagonal-tile inverses,
// then PBX_LA_INTS look-ahead ints (zero before a two-column factorisation),
// then one 64 x 64 scratch tile.
constexpr int PBX_LA_INTS = 16;
int pbx_potrf_two_col_ready(double* A, int np, int ld, double* work, int* info_dev, hipStream_t s);
// Bumped by every change of the factorisation's launch shapes (variant,
// own-CU threshold, per-order table): captured launch sequences are keyed by it.
int pbx_potrf_config_generation();
// a diagonally dominant SPD matrix of order n (pad = identity), for timing
int pbx_launch_spd_fill(double* A, int n, int np, int ld, hipStream_t s);

// Per-step indices of a step replayed from a captured graph: the kernels
// that take the step's noise vector or flux slot read them here, and the
// step's last node advances `step`.  Device memory, one per simulation.
struct PbxStepCursor {
    long long step;       // steps since the run started
    long long ring_base;  // noise ring slot of the run's first step (one attempt per step)
    long long ring_att;   // slots in the noise ring
    long long n;          // draws per slot (3N)
    long long nl;         // flux counters per step
    const double* xi_dev;
    long long* flux_log;
};
int pbx_launch_cursor_advance(PbxStepCursor* cur, hipStream_t s);

// End-of-step values the wrap kernel writes straight into mapped pinned host
// memory (device-side aliases of hipHostMalloc buffers), in place of one
// copy per value.  Any pointer may be null.
struct PbxWrapOut {
    double* pos;                 // the wrapped positions
    int* status;                 // [0] the batch's accepted index, [1] the factorisation status word
    const int* info;             // device status word
    long long* flux;             // per-label crossing counters
    const long long* flux_dev;
    long long* flux_clear;       // device counters to clear for the next step (distinct from flux_dev)
    int nl;
    double* E;                   // force-field energy
    const double* E_dev;
};

// batched move attempts (pbx_step.hip): candidates for K <= pbx_attempt_kmax()
// consecutive attempts (xi of attempt k in ring slot (ring_pos + k) mod
// ring_att), their overlap flags, then the selection of the first accepted
int pbx_attempt_kmax();
int pbx_attempt_slab_width(int np);
size_t pbx_attempt_partial_doubles(int np);
// (pos and a hold the N mobile beads, then Nt - N immobile ones)
int pbx_launch_attempt_batch(const double* B, const double* Bd, int ld, int np, int N, int Nt, const double* pos,
                             const double* ring, int ring_pos, int ring_att, int K, double sq2dt, double* P,
                             double* cand, int* cflag, const double* a, double L, double thr, const int* mask,
                             hipStream_t s);
int pbx_launch_attempt_select(const double* cand, int N, int K, const int* cflag, double* pos, const double* plane_n,
                              const double* plane_p, int flux_on, const int* label, long long* flux, int* sel,
                              hipStream_t s);
int pbx_gemm_nt_test(const double* A, const double* B, double* C, int M, int N, int K, hipStream_t s);

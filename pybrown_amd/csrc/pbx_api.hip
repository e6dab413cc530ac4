// C ABI of libpybrown_amd.so (declared in include/pybrown_amd.h).
//
// Two layers:
//  1. Stateless entry points with the reference's exact names, signatures and
//     semantics (host pointers in and out; src/bdsim/clib/diff_tensor.c and
//     bead.c), so pyBrown's ctypes bindings can load this library unchanged.
//  2. A device-resident simulation handle (pbx_sim_*) that keeps positions,
//     matrices and factors in HBM and performs Box.propagate
//     (src/bdsim/pyBrown/box.py:150-349) with HIP kernels.  The N(0,1) draws
//     come from the native RandomState-equivalent generator (pbx_noise.hip,
//     run ahead on a producer thread) or from a host callback, and are
//     consumed attempt by attempt in the order box.py consumes them.
#include "pbx_common.h"
#include "pbx_noise.h"
#include "../../include/pybrown_amd.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <atomic>
#include <mutex>
#include <thread>
#include <shared_mutex>
#include <vector>

// launchers from pbx_step.hip
int pbx_launch_stoch(const double* B, int ld, int diag_only, const double* xi, double sq2dt, double mult,
                     double* pos, double* disp, int N, const double* plane_n, const double* plane_p, int flux_on,
                     const int* label, long long* flux, hipStream_t s, const PbxStepCursor* cur = nullptr);
int pbx_launch_det(const double* D, int ld, int diag_only, const double* F, double k, double mult, double* pos,
                   int N, const double* plane_n, const double* plane_p, int flux_on, const int* label,
                   long long* flux, hipStream_t s, const PbxStepCursor* cur = nullptr);
int pbx_launch_translate(const double* v, double sign, double* pos, int N, const double* plane_n,
                         const double* plane_p, int flux_on, const int* label, long long* flux, hipStream_t s);
int pbx_launch_wrap(double* pos, int N, double L, hipStream_t s);
int pbx_launch_wrap_publish(double* pos, int N, double L, const int* sel, int K, const PbxWrapOut& o,
                            hipStream_t s);
int pbx_launch_div_shift(double* pos, int idx, double delta, hipStream_t s);
int pbx_launch_div_init(const double* D, int n, int ld, double* div, hipStream_t s);
int pbx_launch_div_add_row(const double* D, int row, int n, int ld, double* div, hipStream_t s);
int pbx_launch_div_finish(const double* div, int n, double eps, double dt, double* out, hipStream_t s);
int pbx_launch_drift(const double* D, const double* Lo, int n, int ld, const double* x, const double* z, double alpha,
                     double* y, hipStream_t s);
int pbx_launch_det_drift(const double* D1, const double* D0, int n, int ld, const double* F, double beta, double* y,
                         hipStream_t s);
int pbx_launch_drift_diff(const double* D1, const double* D0, int n, int ld, const double* x, double alpha, double* y,
                          hipStream_t s);
// trans 1: L^T y = x, 0: L y = x.  words: 2 np zeroed 64-bit words before
// the first call; ticket: one zeroed int; epoch: 1, 2, ... per call
// rev: x and y reversed (x'[i] = x[np-1-i]); y scaled by yscale
int pbx_launch_trsv_chain(const double* Lm, int np, int ld, const double* W, const double* x, double* y,
                          unsigned long long* words, int* ticket, unsigned epoch, int trans, int* info,
                          hipStream_t s, int rev, double yscale);
int pbx_launch_drift_vec(const double* v, const double* bx, int n, double c, double alpha, double* y, hipStream_t s);
int pbx_launch_rev_lower_gemv(const double* Lm, int np, int ld, const double* x, double scale, double* y,
                              hipStream_t s);
int pbx_launch_stoch_vec(const double* bx, double sq2dt, double mult, double* pos, double* disp, int N,
                         const double* plane_n, const double* plane_p, int flux_on, const int* label,
                         long long* flux, hipStream_t s);
int pbx_launch_drift_solved(const double* Bm, int n, int ld, const double* v, const double* x, double c, double alpha,
                            double* y, hipStream_t s);
int pbx_launch_jo_correction(const double* xyz, const double* P, const double* a, int N, double L, double cutoff,
                             int cich, double mul, double div, double* C, int ld, hipStream_t s);

// ---------------------------------------------------------------------------
// error state

static thread_local int g_err_code = 0;
static thread_local char g_err_msg[512] = "";

extern "C" void pbx_set_error(int code, const char* fmt, ...) {
    g_err_code = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err_msg, sizeof(g_err_msg), fmt, ap);
    va_end(ap);
}

extern "C" const char* pbx_last_error(void) { return g_err_msg; }
extern "C" int pbx_last_error_code(void) { return g_err_code; }
extern "C" void pbx_clear_error(void) {
    g_err_code = 0;
    g_err_msg[0] = 0;
}
extern "C" const char* pbx_version(void) { return "pybrown_amd 0.1 gfx950"; }

static std::atomic<long long> g_launches{0};
void pbx_count_launch() { g_launches.fetch_add(1, std::memory_order_relaxed); }
extern "C" long long pbx_kernel_launches(void) { return g_launches.load(std::memory_order_relaxed); }

// ---------------------------------------------------------------------------
// small RAII helpers for the stateless layer

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (bytes == 0) bytes = 8;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            pbx_set_error(PBX_ERR_OOM, "hipMalloc(%zu) failed", bytes);
            p = nullptr;
            return PBX_ERR_OOM;
        }
        return PBX_OK;
    }
    template <class T>
    T* as() {
        return (T*)p;
    }
};

#define TRY(x)                   \
    do {                         \
        int rc_ = (x);           \
        if (rc_) return rc_;     \
    } while (0)

int up(DevBuf& b, const void* h, size_t bytes) {
    TRY(b.alloc(bytes));
    if (bytes) PBX_HIP(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
    return PBX_OK;
}

void report(int rc, const char* where) {
    if (rc) fprintf(stderr, "pybrown_amd: %s failed (%d): %s\n", where, rc, g_err_msg);
}

int rpy_matrix_impl(const double* as, const double* P, int N, double* M, int smith, double L, double alpha, int m,
                    int nn) {
    if (N <= 0) return PBX_OK;
    size_t n = 3 * (size_t)N;
    size_t npairs = (size_t)N * (N - 1) / 2;
    DevBuf da, dP, dM;
    TRY(up(da, as, sizeof(double) * N));
    TRY(up(dP, P, sizeof(double) * 3 * npairs));
    TRY(dM.alloc(sizeof(double) * n * n));
    int rc = smith ? pbx_launch_smith_packed(da.as<double>(), dP.as<double>(), N, L, alpha, m, nn, dM.as<double>(),
                                             (int)n, 0)
                   : pbx_launch_rpy_packed(da.as<double>(), dP.as<double>(), N, dM.as<double>(), (int)n, 0);
    TRY(rc);
    PBX_HIP(hipMemcpy(M, dM.p, sizeof(double) * n * n, hipMemcpyDeviceToHost));
    return PBX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// 1. reference-compatible stateless entry points

// diff_tensor.c:73 RPY_M_matrix
extern "C" void RPY_M_matrix(double* as, double* pointers, int number_of_beads, double* M_matrix) {
    pbx_clear_error();
    report(rpy_matrix_impl(as, pointers, number_of_beads, M_matrix, 0, 0, 0, 0, 0), "RPY_M_matrix");
}

// diff_tensor.c:170 RPY_Smith_M_matrix
extern "C" void RPY_Smith_M_matrix(double* as, double* pointers, double box_length, double alpha, int m_max,
                                   int n_max, int number_of_beads, double* M_matrix) {
    pbx_clear_error();
    report(rpy_matrix_impl(as, pointers, number_of_beads, M_matrix, 1, box_length, alpha, m_max, n_max),
           "RPY_Smith_M_matrix");
}

// diff_tensor.c:270 JO_2B_R_matrix (writes 18 values)
extern "C" void JO_2B_R_matrix(double ai, double aj, double rx, double ry, double rz, double* R_matrix) {
    pbx_clear_error();
    DevBuf d;
    int rc = d.alloc(sizeof(double) * 18);
    if (!rc) rc = pbx_launch_jo_2b(ai, aj, rx, ry, rz, d.as<double>(), 0);
    if (!rc && hipMemcpy(R_matrix, d.p, sizeof(double) * 18, hipMemcpyDeviceToHost) != hipSuccess) {
        pbx_set_error(PBX_ERR_HIP, "JO_2B_R_matrix: copy-out failed");
        rc = PBX_ERR_HIP;
    }
    report(rc, "JO_2B_R_matrix");
}

// diff_tensor.c:346 JO_R_lubrication_correction_matrix (accumulates into C)
extern "C" void JO_R_lubrication_correction_matrix(double* as, double* pointers, int number_of_beads,
                                                   double cutoff_distance, int cichocki_correction,
                                                   double* correction_matrix) {
    pbx_clear_error();
    auto run = [&]() -> int {
        int N = number_of_beads;
        if (N <= 1) return PBX_OK;
        size_t n = 3 * (size_t)N, npairs = (size_t)N * (N - 1) / 2;
        DevBuf da, dP, dC;
        TRY(up(da, as, sizeof(double) * N));
        TRY(up(dP, pointers, sizeof(double) * 3 * npairs));
        TRY(up(dC, correction_matrix, sizeof(double) * n * n));
        TRY(pbx_launch_jo_correction(nullptr, dP.as<double>(), da.as<double>(), N, 0.0, cutoff_distance,
                                     cichocki_correction, 1.0, 1.0, dC.as<double>(), (int)n, 0));
        PBX_HIP(hipMemcpy(correction_matrix, dC.p, sizeof(double) * n * n, hipMemcpyDeviceToHost));
        return PBX_OK;
    };
    report(run(), "JO_R_lubrication_correction_matrix");
}

// bead.c:27 pointer_pbc_matrix (diagonal entries keep the caller's values)
extern "C" void pointer_pbc_matrix(double* positions, int number_of_beads, double box_length, double* pointers) {
    pbx_clear_error();
    auto run = [&]() -> int {
        int N = number_of_beads;
        if (N <= 1) return PBX_OK;
        DevBuf dx, dP;
        TRY(up(dx, positions, sizeof(double) * 3 * N));
        TRY(up(dP, pointers, sizeof(double) * 3 * (size_t)N * N));
        TRY(pbx_launch_pointers(dx.as<double>(), N, box_length, dP.as<double>(), 0));
        PBX_HIP(hipMemcpy(pointers, dP.p, sizeof(double) * 3 * (size_t)N * N, hipMemcpyDeviceToHost));
        return PBX_OK;
    };
    report(run(), "pointer_pbc_matrix");
}

// bead.c:99 pointer_immobile_pbc_matrix
extern "C" void pointer_immobile_pbc_matrix(double* positions_mobile, double* positions_immobile, int number_of_mobile,
                                            int number_of_immobile, double box_length, double* pointers) {
    pbx_clear_error();
    auto run = [&]() -> int {
        int nm = number_of_mobile, ni = number_of_immobile;
        if (nm <= 0 || ni <= 0) return PBX_OK;
        size_t bytes = sizeof(double) * 3 * (size_t)nm * ni;
        DevBuf dm, di, dP;
        TRY(up(dm, positions_mobile, sizeof(double) * 3 * nm));
        TRY(up(di, positions_immobile, sizeof(double) * 3 * ni));
        TRY(dP.alloc(bytes));
        TRY(pbx_launch_immobile_pointers(dm.as<double>(), di.as<double>(), nm, ni, box_length, dP.as<double>(), 0));
        PBX_HIP(hipMemcpy(pointers, dP.p, bytes, hipMemcpyDeviceToHost));
        return PBX_OK;
    };
    report(run(), "pointer_immobile_pbc_matrix");
}

// bead.c:169 check_overlaps; returns 1 on overlap, 0 otherwise (-1 on error)
extern "C" int check_overlaps(double* positions, double* as, int number_of_beads, double box_length,
                              double overlap_treshold, int* connection_matrix) {
    pbx_clear_error();
    int result = 0;
    auto run = [&]() -> int {
        int N = number_of_beads;
        if (N <= 1) return PBX_OK;
        DevBuf dx, da, dm, df;
        TRY(up(dx, positions, sizeof(double) * 3 * N));
        TRY(up(da, as, sizeof(double) * N));
        if (connection_matrix) TRY(up(dm, connection_matrix, sizeof(int) * (size_t)N * N));
        TRY(df.alloc(sizeof(int)));
        PBX_HIP(hipMemset(df.p, 0, sizeof(int)));
        TRY(pbx_launch_overlaps(dx.as<double>(), da.as<double>(), N, N, box_length, overlap_treshold,
                                connection_matrix ? dm.as<int>() : nullptr, df.as<int>(), 0));
        PBX_HIP(hipMemcpy(&result, df.p, sizeof(int), hipMemcpyDeviceToHost));
        return PBX_OK;
    };
    int rc = run();
    report(rc, "check_overlaps");
    return rc ? -1 : (result ? 1 : 0);
}

// ---------------------------------------------------------------------------
// dense kernels on host arrays (the numpy.linalg call sites of box.py)

// Lower Cholesky of the n x n row-major SPD matrix A (box.py:554).  L gets
// the factor with a zero upper triangle.  Returns 0, k>0 (non-positive pivot
// at column k-1) or a negative error.
static int cholesky_host(const double* A, int n, double* L, int mixed) {
    pbx_clear_error();
    if (n <= 0) return PBX_ERR_ARG;
    int np = (int)pbx_round_up(n, 64), ld = np;
    DevBuf dA, dW, dI;
    TRY(dA.alloc(sizeof(double) * (size_t)np * ld));
    TRY(dW.alloc(sizeof(double) * pbx_dense_work_doubles(np)));
    TRY(dI.alloc(sizeof(int)));
    PBX_HIP(hipMemset(dI.p, 0, sizeof(int)));
    PBX_HIP(hipMemcpy2D(dA.p, sizeof(double) * ld, A, sizeof(double) * n, sizeof(double) * n, n,
                        hipMemcpyHostToDevice));
    TRY(pbx_fill_pad(dA.as<double>(), n, np, ld, 0));
    TRY(pbx_potrf(dA.as<double>(), np, ld, dW.as<double>(), dI.as<int>(), 0, mixed));
    int info = 0;
    PBX_HIP(hipMemcpy(&info, dI.p, sizeof(int), hipMemcpyDeviceToHost));
    PBX_HIP(hipMemcpy2D(L, sizeof(double) * n, dA.p, sizeof(double) * ld, sizeof(double) * n, n,
                        hipMemcpyDeviceToHost));
    for (int r = 0; r < n; ++r)
        for (int c = r + 1; c < n; ++c) L[(size_t)r * n + c] = 0.0;
    if (info == PBX_INFO_DAG_WAIT) {
        pbx_set_error(PBX_ERR_STATE, "a bounded dependency wait ran out (factorisation or solve)");
        return PBX_ERR_STATE;
    }
    if (info) pbx_set_error(info, "Matrix is not positive definite (pivot %d)", info - 1);
    return info;
}

extern "C" int pbx_cholesky(const double* A, int n, double* L) { return cholesky_host(A, n, L, 0); }

// The same factor with the trailing updates on FP32 matrix cores (operands
// rounded to FP32, FP32 accumulation added into the FP64 matrix); diagonal
// tiles factored and inverted in FP64.
extern "C" int pbx_cholesky_mixed(const double* A, int n, double* L) { return cholesky_host(A, n, L, 1); }

static int spd_inverse_host(const double* A, int n, double scale, int precision, double* X) {
    pbx_clear_error();
    if (n <= 0 || precision < 0 || precision > 2) return PBX_ERR_ARG;
    int np = (int)pbx_round_up(n, 64), ld = np;
    DevBuf dA, dX, dW, dI, dC;
    TRY(dA.alloc(sizeof(double) * (size_t)np * ld));
    TRY(dX.alloc(sizeof(double) * (size_t)np * ld));
    if (precision == 2) TRY(dC.alloc(sizeof(double) * (size_t)np * ld));
    TRY(dW.alloc(sizeof(double) * pbx_dense_work_doubles(np)));
    TRY(dI.alloc(sizeof(int)));
    PBX_HIP(hipMemset(dI.p, 0, sizeof(int)));
    PBX_HIP(hipMemcpy2D(dA.p, sizeof(double) * ld, A, sizeof(double) * n, sizeof(double) * n, n,
                        hipMemcpyHostToDevice));
    TRY(pbx_fill_pad(dA.as<double>(), n, np, ld, 0));
    TRY(pbx_spd_inverse(dA.as<double>(), np, ld, dW.as<double>(), dX.as<double>(), dI.as<int>(), 0, scale,
                        precision, precision == 2 ? dC.as<double>() : nullptr));
    int info = 0;
    PBX_HIP(hipMemcpy(&info, dI.p, sizeof(int), hipMemcpyDeviceToHost));
    PBX_HIP(hipMemcpy2D(X, sizeof(double) * n, dA.p, sizeof(double) * ld, sizeof(double) * n, n,
                        hipMemcpyDeviceToHost));
    if (info == PBX_INFO_DAG_WAIT) {
        pbx_set_error(PBX_ERR_STATE, "a bounded dependency wait ran out (factorisation or solve)");
        return PBX_ERR_STATE;
    }
    if (info) pbx_set_error(info, "Matrix is not positive definite (pivot %d)", info - 1);
    return info;
}

// X = scale * inv(A) for SPD A (box.py:514,524,537 np.linalg.inv call sites)
extern "C" int pbx_spd_inverse_host(const double* A, int n, double scale, double* X) {
    return spd_inverse_host(A, n, scale, 0, X);
}

// The same with the factorisation's trailing updates and the inverse's tile
// products on FP32 matrix cores; refine = 1 adds one Newton-Schulz step with
// an FP64 residual.
extern "C" int pbx_spd_inverse_mixed(const double* A, int n, double scale, int refine, double* X) {
    if (refine != 0 && refine != 1) {
        pbx_set_error(PBX_ERR_ARG, "spd_inverse_mixed: refine must be 0 or 1");
        return PBX_ERR_ARG;
    }
    return spd_inverse_host(A, n, scale, 1 + refine, X);
}

// x = inv(A) b for SPD A through its factor A = L L^T and the two chained
// triangular solves the midpoint uses for D_mid w (box.py:303-308): y =
// inv(L) b, then x = inv(L^T) y.  The solves run reps times (fresh epochs on
// the same hand-off words); ms, when non-null, receives their mean device
// time per pair.
extern "C" int pbx_spd_solve(const double* A, int n, const double* b, int reps, double* y, double* x, double* ms) {
    pbx_clear_error();
    if (n <= 0 || reps <= 0 || !A || !b || !y || !x) return PBX_ERR_ARG;
    int np = (int)pbx_round_up(n, 64), ld = np;
    DevBuf dA, dW, dI, dv, dy, dx, dwords, dt;
    TRY(dA.alloc(sizeof(double) * (size_t)np * ld));
    TRY(dW.alloc(sizeof(double) * pbx_dense_work_doubles(np)));
    TRY(dI.alloc(sizeof(int)));
    TRY(dv.alloc(sizeof(double) * np));
    TRY(dy.alloc(sizeof(double) * np));
    TRY(dx.alloc(sizeof(double) * np));
    TRY(dwords.alloc(sizeof(unsigned long long) * 2 * np));
    TRY(dt.alloc(sizeof(int)));
    PBX_HIP(hipMemset(dI.p, 0, sizeof(int)));
    PBX_HIP(hipMemset(dt.p, 0, sizeof(int)));
    PBX_HIP(hipMemset(dwords.p, 0, sizeof(unsigned long long) * 2 * np));
    PBX_HIP(hipMemset(dv.p, 0, sizeof(double) * np));
    PBX_HIP(hipMemcpy(dv.p, b, sizeof(double) * n, hipMemcpyHostToDevice));
    PBX_HIP(hipMemcpy2D(dA.p, sizeof(double) * ld, A, sizeof(double) * n, sizeof(double) * n, n,
                        hipMemcpyHostToDevice));
    TRY(pbx_fill_pad(dA.as<double>(), n, np, ld, 0));
    TRY(pbx_potrf(dA.as<double>(), np, ld, dW.as<double>(), dI.as<int>(), 0, 0));
    hipEvent_t e0, e1;
    PBX_HIP(hipEventCreate(&e0));
    PBX_HIP(hipEventCreate(&e1));
    PBX_HIP(hipEventRecord(e0, 0));
    unsigned ep = 0;
    int rc = PBX_OK;
    for (int r = 0; r < reps && rc == PBX_OK; ++r) {
        rc = pbx_launch_trsv_chain(dA.as<double>(), np, ld, dW.as<double>(), dv.as<double>(), dy.as<double>(),
                                   (unsigned long long*)dwords.p, dt.as<int>(), ++ep, 0, dI.as<int>(), 0, 0, 1.0);
        if (rc == PBX_OK)
            rc = pbx_launch_trsv_chain(dA.as<double>(), np, ld, dW.as<double>(), dy.as<double>(), dx.as<double>(),
                                       (unsigned long long*)dwords.p, dt.as<int>(), ++ep, 1, dI.as<int>(), 0, 0,
                                       1.0);
    }
    (void)hipEventRecord(e1, 0);
    const hipError_t es = hipEventSynchronize(e1);
    float t = 0.f;
    if (es == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    PBX_HIP(es);
    if (ms) *ms = t / reps;
    int info = 0;
    PBX_HIP(hipMemcpy(&info, dI.p, sizeof(int), hipMemcpyDeviceToHost));
    PBX_HIP(hipMemcpy(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    PBX_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (info == PBX_INFO_DAG_WAIT) {
        pbx_set_error(PBX_ERR_STATE, "a bounded dependency wait ran out (factorisation or solve)");
        return PBX_ERR_STATE;
    }
    if (info) pbx_set_error(info, "Matrix is not positive definite (pivot %d)", info - 1);
    return info;
}

// "<name> (<arch>)" of the current device, for records of tuned settings.
extern "C" int pbx_device_name(char* out, int len) {
    pbx_clear_error();
    if (!out || len <= 0) return PBX_ERR_ARG;
    int dev = 0;
    hipDeviceProp_t prop;
    PBX_HIP(hipGetDevice(&dev));
    PBX_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(out, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    return PBX_OK;
}

// Milliseconds per factorisation (pbx_cholesky's device part: the copy of D
// into B and the factorisation) of a synthetic SPD matrix of order n, with
// the launch shapes in force, averaged over `reps` after one untimed run.
extern "C" int pbx_time_potrf(int n, int reps, int mixed, double* ms) {
    pbx_clear_error();
    if (n <= 0 || reps <= 0 || !ms || (mixed != 0 && mixed != 1)) {
        pbx_set_error(PBX_ERR_ARG, "time_potrf: n %d reps %d mixed %d", n, reps, mixed);
        return PBX_ERR_ARG;
    }
    const int np = (int)pbx_round_up(n, 64), ld = np;
    DevBuf dD, dB, dW, dI;
    TRY(dD.alloc(sizeof(double) * (size_t)np * ld));
    TRY(dB.alloc(sizeof(double) * (size_t)np * ld));
    TRY(dW.alloc(sizeof(double) * pbx_dense_work_doubles(np)));
    TRY(dI.alloc(sizeof(int)));
    hipStream_t st = nullptr;
    PBX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PBX_OK, info = 0;
    float t = 0.f;
    auto run = [&]() {
        return pbx_potrf_copy(dD.as<double>(), dB.as<double>(), np, ld, dW.as<double>(), dI.as<int>(), st, mixed);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemsetAsync(dI.p, 0, sizeof(int), st) != hipSuccess) {
        pbx_set_error(PBX_ERR_HIP, "time_potrf: event or memset failed");
        rc = PBX_ERR_HIP;
    }
    if (!rc) rc = pbx_launch_spd_fill(dD.as<double>(), n, np, ld, st);
    if (!rc) rc = run();
    if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = PBX_ERR_HIP;
    for (int r = 0; r < reps && !rc; ++r) rc = run();
    if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess ||
                hipMemcpy(&info, dI.p, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)) {
        pbx_set_error(PBX_ERR_HIP, "time_potrf: timing failed");
        rc = PBX_ERR_HIP;
    }
    (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (rc) return rc;
    if (info) {
        pbx_set_error(PBX_ERR_STATE, "time_potrf: factorisation status %d", info);
        return PBX_ERR_STATE;
    }
    *ms = (double)t / reps;
    return PBX_OK;
}

// C = A B^T, all dimensions multiples of 64 (checks the MFMA tile layout)
extern "C" int pbx_gemm_nt(const double* A, const double* B, double* C, int M, int N, int K) {
    pbx_clear_error();
    DevBuf dA, dB, dC;
    TRY(up(dA, A, sizeof(double) * (size_t)M * K));
    TRY(up(dB, B, sizeof(double) * (size_t)N * K));
    TRY(dC.alloc(sizeof(double) * (size_t)M * N));
    TRY(pbx_gemm_nt_test(dA.as<double>(), dB.as<double>(), dC.as<double>(), M, N, K, 0));
    PBX_HIP(hipMemcpy(C, dC.p, sizeof(double) * (size_t)M * N, hipMemcpyDeviceToHost));
    return PBX_OK;
}

// ---------------------------------------------------------------------------
// 2. device-resident simulation

struct PbxSim {
    pbx_config cfg;
    int N = 0, n = 0, np = 0, ld = 0;
    int Nt = 0;  // N mobile beads, then Nt - N immobile ones (pos and a only)
    double kBT = 0.0;
    int F_nonzero = 0;
    hipStream_t stream = nullptr;
    // device state
    double *pos = nullptr, *a = nullptr, *F = nullptr, *disp = nullptr, *drift = nullptr;
    double *w = nullptr, *xipad = nullptr, *u = nullptr;
    int* label = nullptr;
    long long* flux = nullptr;
    // the other per-step counter buffer: the wrap kernel of a step clears it
    // (flux_alt_zero), and the next step swaps it in instead of clearing
    // `flux` with a fill of its own
    long long* flux_alt = nullptr;
    int flux_alt_zero = 0;
    int *flag = nullptr, *info = nullptr;
    double *D = nullptr, *B = nullptr, *Rff = nullptr, *X = nullptr, *Dmid = nullptr, *RffMid = nullptr;
    double *Wchol = nullptr, *Winv = nullptr;
    // lubrication modes: precision of the two SPD inverses per D (0 FP64,
    // 1 FP32 matrix-core bulk, 2 that plus one Newton-Schulz step, whose
    // copy of the input matrix is Xref)
    int inv_precision = 0;
    double* Xref = nullptr;
    // launch sequences with fixed arguments (the factorisation and the SPD
    // inverses) replayed from HIP graphs when `graphs` is on: captured on
    // first use per (operation, output buffer, Cholesky variant)
    struct GraphSlot {
        int op;
        const void* key;
        int variant;
        hipGraphExec_t exec;
    };
    int graphs = 0;
    int n_graphs = 0;
    GraphSlot graph_slots[8];
    // whole-step graphs of queued runs (graphs on, Ermak without overlap
    // checks, RPY or Smith-Ewald): one per (builds D, factors) pair and time
    // step, replayed with the step's noise slot and flux slot read from the
    // cursor, which the graph's last node advances
    PbxStepCursor* cursor_dev = nullptr;
    PbxStepCursor* cursor_host = nullptr;  // pinned, written once per run
    int capturing = 0;                     // a whole step is being captured
    double step_graph_dt = 0.0;
    // midpoint: the chained triangular solves' hand-off words and row ticket
    // counter (zeroed at create, words written with a per-call epoch; not
    // part of a snapshot)
    unsigned long long* chain_words = nullptr;
    int* chain_ticket = nullptr;
    unsigned chain_epoch = 0;
    // D and B as the lubrication builds leave them (ensure_B / ensure_D form
    // what is read): 0 both formed; 1 B formed, D holds inv(L) of the
    // one-factorisation build; 2 B holds L = chol(J R J) and Wchol the
    // inverses of L's diagonal tiles (B = sqrt(kBT) J L^{-T} J), D unformed
    int mat_state = 0;
    double* bxi = nullptr;  // B xi of the current midpoint attempt (B held as L)
    double *Dd = nullptr, *Bd = nullptr;  // nohi diagonal D and sqrt(D)
    // noise: chunks of xi_cap attempts (3N draws each) in order; the device
    // ring holds two chunks (chunk q in slot q % 2), the pinned host side
    // three (chunk q in buffer q % 3, reusable once xi_ev[q % 3] completes).
    // The uploads run on their own stream `up`, ordered against `stream` by
    // events both ways, so the producer thread's waits on xi_ev never touch
    // `stream` (which may be capturing a graph at the time: waiting on an
    // event of a capturing stream invalidates the capture).
    double* xi_dev = nullptr;
    double* xi_host = nullptr;
    hipEvent_t xi_ev[3] = {nullptr, nullptr, nullptr};
    hipStream_t up = nullptr;
    hipEvent_t rd_ev = nullptr;  // recorded on `stream` before each upload
    long long xi_cap = 0;
    long long xi_used = 0;   // attempts consumed from the current source
    long long xi_taken = 0;  // chunks whose upload has been queued
    pbx_xi_fill_fn xi_fill = nullptr;
    void* xi_user = nullptr;
    long long draw_count = 0;
    // native source: producer thread filling host chunks ahead of use
    int native = 0;
    PbxMT mt;
    PbxMT mt_after[4];        // generator state after chunk q, in slot q % 4
    PbxMT mt_start;           // state at chunk mt_base (seed or restore point)
    long long mt_base = 0;
    std::thread producer;
    std::mutex mu;
    std::condition_variable cv;
    long long xi_produced = 0;
    bool stop = false;
    // batched move attempts (Ermak with overlap checks)
    double *Pbuf = nullptr, *cand = nullptr;
    int* cflag = nullptr;
    int batch_fixed = 0;       // 0: adaptive
    double att_mean = 2.0;     // running mean of attempts per step
    int* flag_host = nullptr;  // pinned: [0] overlap flag, [1] factorisation status
    double* pos_host = nullptr;      // pinned copy of positions (end of step)
    long long* flux_host = nullptr;  // pinned copy of flux counters
    int info_pending = 0;            // a status copy is queued into flag_host[1]
    // device-side aliases of flag_host, pos_host, flux_host and E_host (the
    // end-of-step values are written there by the wrap kernel)
    int* flag_hd = nullptr;
    double* pos_hd = nullptr;
    long long* flux_hd = nullptr;
    double* E_hd = nullptr;
    int last_attempts = 0;
    int have_D = 0, have_B = 0, have_Rff = 0;
    std::vector<void*> owned;
    size_t device_bytes = 0;         // sum of the allocations in `owned`
    // force field (pbx_sim_set_forces): recomputed at the start of every step
    int* ovl_mask = nullptr;         // connection matrix of the overlap test (Nt x Nt), or null
    int forces_set = 0, forces_on = 0;
    PbxForceArgs fa{};
    double* e_part = nullptr;        // per-bead energies
    // divergence term (box.py:560-610): perturbed D / Rff and the accumulator
    double* Ddiv = nullptr;
    double* RffDiv = nullptr;
    double* divv = nullptr;
    double* E_dev = nullptr;
    double* E_host = nullptr;        // pinned: the last step's energy
    // optional stage timing with HIP events on the simulation stream
    int timing = 0;
    // ev[0..3]: stage bounds of the step; ev[4..7]: two pairs around the
    // look-ahead assembly (one pair read while the other is recorded)
    hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int ahead_pending = -1;  // pair recorded at the end of the last step, not read yet
    // Look-ahead (non-lubrication matrix modes): a step that built D ends by
    // assembling D for its final positions into Dspec, after the result
    // copies, while the host reads the results and comes back; the next
    // step, if it builds D and the positions were not replaced since, swaps
    // Dspec in instead of assembling.  The host waits on done_ev (recorded
    // before that assembly), not on the whole stream.
    double* Dspec = nullptr;
    int spec_ok = 0;
    int lookahead = 1;
    int next_builds = -1;  // whether the next step builds D: 1, 0, or -1 unknown (then assumed)
    hipEvent_t done_ev = nullptr;
    double stage_ms[3] = {0, 0, 0};  // assembly (+lubrication), factorisation, propagation
    // queued runs (pbx_sim_run without host decisions): per-step flux slots
    long long* flux_log = nullptr;
    long long flux_log_cap = 0;  // in counters
    long long timed_steps = 0;
};

namespace {

// Graph captures (any simulation's, on its own stream) hold this shared;
// device frees, which the runtime may order against every stream, hold it
// exclusively, so no free lands inside another thread's capture.
std::shared_mutex g_capture_mu;

// Synchronous copies of a simulation go through its own (non-blocking)
// stream, never the legacy one: another thread may be capturing a graph on
// its stream, and the runtime refuses legacy-stream work while any capture is
// in progress.
hipError_t sim_copy(PbxSim* s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s->stream);
}

int sim_alloc(PbxSim* s, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        pbx_set_error(PBX_ERR_OOM, "hipMalloc(%zu) failed", bytes);
        *p = nullptr;
        return PBX_ERR_OOM;
    }
    s->owned.push_back(*p);
    s->device_bytes += bytes;
    return PBX_OK;
}

int sim_matrix(PbxSim* s, double** p) {
    TRY(sim_alloc(s, (void**)p, sizeof(double) * (size_t)s->np * s->ld));
    PBX_HIP(hipMemsetAsync(*p, 0, sizeof(double) * (size_t)s->np * s->ld, s->stream));
    return pbx_fill_pad(*p, s->n, s->np, s->ld, s->stream);
}

bool is_lub(const PbxSim* s) { return s->cfg.hydrodynamics == PBX_RPY_LUB || s->cfg.hydrodynamics == PBX_RPY_SMITH_LUB; }
bool is_smith(const PbxSim* s) {
    return s->cfg.hydrodynamics == PBX_RPY_SMITH || s->cfg.hydrodynamics == PBX_RPY_SMITH_LUB;
}

// Mobility (optionally Ewald) times `scale`, written into M.
int assemble_M(PbxSim* s, double* M, double scale, const int* sel = nullptr, int selK = 0) {
    const pbx_config& c = s->cfg;
    if (is_smith(s))
        return pbx_launch_smith(s->pos, s->a, s->N, c.box_length, c.ewald_alpha, c.ewald_real, c.ewald_imag, scale,
                                M, s->ld, 0, s->stream, sel, selK);
    return pbx_launch_rpy(s->pos, s->a, s->N, c.box_length, scale, M, s->ld, 0, s->stream, sel, selK);
}

// Queue the factorisation status word for the next synchronisation (no
// stall: the step keeps going and the word is examined once the stream has
// been synchronised for another reason).
int post_info(PbxSim* s) {
    PBX_HIP(hipMemcpyAsync(s->flag_host + 1, s->info, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    s->info_pending = 1;
    return PBX_OK;
}

// After a stream synchronisation: report a queued non-zero status.
int take_info(PbxSim* s) {
    if (!s->info_pending) return PBX_OK;
    s->info_pending = 0;
    int info = s->flag_host[1];
    if (!info) return PBX_OK;
    s->flag_host[1] = 0;  // reported once
    PBX_HIP(hipMemsetAsync(s->info, 0, sizeof(int), s->stream));
    if (info == PBX_INFO_DAG_WAIT) {
        pbx_set_error(PBX_ERR_STATE, "a bounded dependency wait ran out (factorisation or solve)");
        return PBX_ERR_STATE;
    }
    pbx_set_error(info, "Matrix is not positive definite (pivot %d)", info - 1);
    return info;
}

int check_info(PbxSim* s) {
    int info = 0;
    PBX_HIP(hipMemcpyAsync(s->flag_host + 1, s->info, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    info = s->flag_host[1];
    if (info) {
        PBX_HIP(hipMemsetAsync(s->info, 0, sizeof(int), s->stream));
        if (info == PBX_INFO_DAG_WAIT) {
            pbx_set_error(PBX_ERR_STATE, "a bounded dependency wait ran out (factorisation or solve)");
            return PBX_ERR_STATE;
        }
        pbx_set_error(info, "Matrix is not positive definite (pivot %d)", info - 1);
    }
    return info;
}

// Run `body` (stream work with fixed arguments) on the simulation's stream:
// eagerly, or with graphs on from a HIP graph captured on first use and
// launched as one unit afterwards.  The sequences captured here hold kernel
// launches only (no memset nodes); the table is small, and a full table
// falls back to eager launches.
template <class F>
int run_graphed(PbxSim* s, int op, const void* key, F&& body) {
    if (!s->graphs || s->capturing) return body();  // (inside a whole-step capture: part of it)
    const int variant = pbx_potrf_config_generation();
    for (int q = 0; q < s->n_graphs; ++q) {
        const PbxSim::GraphSlot& g = s->graph_slots[q];
        if (g.op == op && g.key == key && g.variant == variant) {
            PBX_HIP(hipGraphLaunch(g.exec, s->stream));
            return PBX_OK;
        }
    }
    if (s->n_graphs == (int)(sizeof(s->graph_slots) / sizeof(s->graph_slots[0]))) return body();
    std::shared_lock<std::shared_mutex> capture_lock(g_capture_mu);
    PBX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s->stream, &graph);
    capture_lock.unlock();
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        pbx_set_error(PBX_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        return PBX_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        pbx_set_error(PBX_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ei));
        return PBX_ERR_HIP;
    }
    s->graph_slots[s->n_graphs++] = {op, key, variant, exec};
    PBX_HIP(hipGraphLaunch(exec, s->stream));
    return PBX_OK;
}

void drop_graphs(PbxSim* s) {
    for (int q = 0; q < s->n_graphs; ++q) (void)hipGraphExecDestroy(s->graph_slots[q].exec);
    s->n_graphs = 0;
}

enum {
    GRAPH_FACTOR = 1,
    GRAPH_INV_RFF = 2,
    GRAPH_INV_R = 3,
    GRAPH_INV_FACTOR = 4,
    GRAPH_FACTOR_RMID = 5,
    GRAPH_FACTOR_REV = 6,
    GRAPH_STEP = 8  // + 0..3
};  // GRAPH_STEP + 0..3

// box.py:486-528 _compute_Dff_matrix into (D) or (Rff) depending on mode.
int build_Dff(PbxSim* s, double* Dout, double* RffOut) {
    const pbx_config& c = s->cfg;
    if (!is_lub(s)) {
        // D = M * (kBT 1e19 / eta)
        TRY(assemble_M(s, Dout, s->kBT * 1e19 / c.viscosity));
        return PBX_OK;
    }
    // Mff = M 1e19 / eta ; Rff = inv(Mff)
    TRY(assemble_M(s, RffOut, 1e19 / c.viscosity));
    return run_graphed(s, GRAPH_INV_RFF, RffOut,
                       [&]() {
                           return pbx_spd_inverse(RffOut, s->np, s->ld, s->Winv, s->X, s->info, s->stream, 1.0,
                                                  s->inv_precision, s->Xref);
                       });
}

// A lubrication step that rebuilds D and factors it gets both from one
// factorisation (pbx_inverse_and_factor) when the SPD inverses and the noise
// factor are FP64.
bool fuse_inverse_and_factor(const PbxSim* s) {
    return is_lub(s) && s->inv_precision == 0 && s->cfg.factor_precision == 0 && s->X && s->B;
}

// box.py:533-541: R = dR eta/1e19 + Rff (into Rout)
int build_R(PbxSim* s, const double* RffIn, double* Rout) {
    const pbx_config& c = s->cfg;
    PBX_HIP(hipMemcpyAsync(Rout, RffIn, sizeof(double) * (size_t)s->np * s->ld, hipMemcpyDeviceToDevice, s->stream));
    return pbx_launch_jo_correction(s->pos, nullptr, s->a, s->N, c.box_length, c.lubrication_cutoff,
                                    c.cichocki_correction, c.viscosity, 1e19, Rout, s->ld, s->stream);
}

// B itself, and then D, when a lubrication build left them unformed (the
// same operations as forming them there, so the same values)
int ensure_B(PbxSim* s) {
    if (s->mat_state != 2) return PBX_OK;
    s->mat_state = 1;
    return pbx_factor_reversed_finish(s->B, s->Wchol, s->D, s->np, s->ld, s->kBT, s->stream);
}

int ensure_D(PbxSim* s) {
    TRY(ensure_B(s));
    if (s->mat_state != 1) return PBX_OK;
    s->mat_state = 0;
    return pbx_inverse_and_factor_finish(s->D, s->X, s->n, s->np, s->ld, s->kBT, s->stream);
}

// The lubrication midpoint without forces reads D_mid only in D_mid w
// (box.py:303-308), which is kBT R_mid^{-1} w: a factorisation of R_mid and
// two triangular solves give it without the inverse's triangular inverse
// and X^T X.  Its B0 enters only as B0 xi and B0^{-T} xi, which the factor
// L of J R J gives with one triangular solve and one triangular product.
// FP64 inverses and factor only.
bool lean_midpoint(const PbxSim* s, int bDnf) {
    return is_lub(s) && bDnf && !s->F_nonzero && s->cfg.factor_precision == 0 && s->inv_precision == 0 && s->u;
}

// box.py:533-541 _compute_Dtot_matrix: R = dR eta/1e19 + Rff; D = kBT inv(R);
// with_factor: also box.py:546-556, B = chol(D) (fuse_inverse_and_factor).
// With the factor, D itself is formed only when read (ensure_D): a step
// without forces, divergence term or explicit read propagates with B alone;
// a lean midpoint step (lean_midpoint) keeps B as L too (ensure_B).
int build_Dtot(PbxSim* s, const double* RffIn, double* Dout, bool with_factor = false) {
    if (Dout == s->D) {
        // a new D without a new factor: the last factor stays, formed
        // before D's buffer is reused
        if (!with_factor) TRY(ensure_B(s));
        s->mat_state = 0;
    }
    TRY(build_R(s, RffIn, Dout));
    if (with_factor && Dout == s->D && s->cfg.scheme == PBX_MIDPOINT && lean_midpoint(s, 1)) {
        TRY(run_graphed(s, GRAPH_FACTOR_REV, s->B, [&]() {
            return pbx_factor_reversed(Dout, s->B, s->Wchol, s->n, s->np, s->ld, s->kBT, s->info, s->stream);
        }));
        s->mat_state = 2;
        return PBX_OK;
    }
    if (with_factor) {
        TRY(run_graphed(s, GRAPH_INV_FACTOR, Dout, [&]() {
            return pbx_inverse_and_factor(Dout, s->X, s->B, s->Wchol, s->Winv, s->n, s->np, s->ld, s->kBT, s->info,
                                          s->stream, 0);
        }));
        if (Dout == s->D) s->mat_state = 1;
        else TRY(pbx_inverse_and_factor_finish(Dout, s->X, s->n, s->np, s->ld, s->kBT, s->stream));
        return PBX_OK;
    }
    return run_graphed(s, GRAPH_INV_R, Dout, [&]() {
        TRY(pbx_spd_inverse(Dout, s->np, s->ld, s->Winv, s->X, s->info, s->stream, s->kBT, s->inv_precision,
                            s->Xref));
        return pbx_fill_pad(Dout, s->n, s->np, s->ld, s->stream);
    });
}

// box.py:546-556 _decompose_D_matrix
int factor(PbxSim* s) {
    TRY(ensure_D(s));
    return run_graphed(s, GRAPH_FACTOR, s->B, [&]() {
        return pbx_potrf_copy(s->D, s->B, s->np, s->ld, s->Wchol, s->info, s->stream, s->cfg.factor_precision);
    });
}

// ---- noise ---------------------------------------------------------------

size_t chunk_doubles(const PbxSim* s) { return (size_t)s->xi_cap * s->n; }

void producer_main(PbxSim* s) {
    for (long long q = s->xi_produced;; ++q) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // host buffer q % 3 last held chunk q - 3: wait until its upload is queued
            s->cv.wait(lk, [&] { return s->stop || s->xi_taken >= q - 2; });
            if (s->stop) return;
        }
        if (q >= 3) (void)hipEventSynchronize(s->xi_ev[q % 3]);
        pbx_mt_normals(&s->mt, s->xi_host + (q % 3) * chunk_doubles(s), (long long)chunk_doubles(s));
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->mt_after[q % 4] = s->mt;
            s->xi_produced = q + 1;
        }
        s->cv.notify_all();
    }
}

void noise_stop(PbxSim* s) {
    if (s->producer.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        s->producer.join();
    }
    s->stop = false;
}

// Drop every drawn-but-unused value and start a new source.
int noise_reset(PbxSim* s) {
    noise_stop(s);
    PBX_HIP(hipStreamSynchronize(s->stream));
    if (s->up) PBX_HIP(hipStreamSynchronize(s->up));
    s->xi_used = s->xi_taken = s->xi_produced = 0;
    s->native = 0;
    s->xi_fill = nullptr;
    return PBX_OK;
}

// Queue the upload of chunk q into device slot q % 2.
int noise_take(PbxSim* s, long long q, bool readers_done = false) {
    double* hb = s->xi_host + (q % 3) * chunk_doubles(s);
    if (s->native) {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv.wait(lk, [&] { return s->xi_produced > q; });
    } else {
        if (!s->xi_fill) {
            pbx_set_error(PBX_ERR_STATE, "no noise source registered");
            return PBX_ERR_STATE;
        }
        if (q >= 3) PBX_HIP(hipEventSynchronize(s->xi_ev[q % 3]));
        if (s->xi_fill(s->xi_user, hb, (long long)chunk_doubles(s)) != 0) {
            pbx_set_error(PBX_ERR_STATE, "noise source callback failed");
            return PBX_ERR_STATE;
        }
    }
    // after the stream's readers of the slot's previous chunk (already
    // complete when the caller says so), before its next readers
    if (!readers_done) {
        PBX_HIP(hipEventRecord(s->rd_ev, s->stream));
        PBX_HIP(hipStreamWaitEvent(s->up, s->rd_ev, 0));
    }
    PBX_HIP(hipMemcpyAsync(s->xi_dev + (q % 2) * chunk_doubles(s), hb, sizeof(double) * chunk_doubles(s),
                           hipMemcpyHostToDevice, s->up));
    PBX_HIP(hipEventRecord(s->xi_ev[q % 3], s->up));
    PBX_HIP(hipStreamWaitEvent(s->stream, s->xi_ev[q % 3], 0));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->xi_taken = q + 1;
    }
    s->cv.notify_all();
    return PBX_OK;
}

// Make attempts xi_used .. xi_used + K - 1 resident (K <= xi_cap, so they
// span at most two chunks; uploading chunk q overwrites chunk q - 2, whose
// readers precede it on the stream).
int noise_ensure(PbxSim* s, long long K) {
    const long long q1 = (s->xi_used + K - 1) / s->xi_cap;
    while (s->xi_taken <= q1) TRY(noise_take(s, s->xi_taken));
    return PBX_OK;
}

// Upload the next chunk ahead of need, from the end of a step whose kernels
// have all completed except a look-ahead assembly (which reads no noise):
// when every attempt of the chunk it overwrites is consumed and the producer
// has it ready, its copy overlaps that assembly instead of stalling a later
// attempt batch.
int noise_prefetch(PbxSim* s) {
    if (!s->native) return PBX_OK;
    const long long q = s->xi_taken;
    if (s->xi_used < (q - 1) * s->xi_cap) return PBX_OK;  // slot q % 2 still holds unconsumed attempts
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->xi_produced <= q) return PBX_OK;
    }
    return noise_take(s, q, true);
}

int ring_attempts(const PbxSim* s) { return (int)(2 * s->xi_cap); }
int ring_pos(const PbxSim* s) { return (int)(s->xi_used % (2 * s->xi_cap)); }

void noise_consume(PbxSim* s, long long k) {
    s->xi_used += k;
    s->draw_count += k * s->n;
}

// next 3N draws (one move attempt, box.py:393-399)
int next_xi(PbxSim* s, const double** out) {
    TRY(noise_ensure(s, 1));
    *out = s->xi_dev + (size_t)ring_pos(s) * s->n;
    noise_consume(s, 1);
    return PBX_OK;
}

int overlap_now(PbxSim* s, int* hit) {
    PBX_HIP(hipMemsetAsync(s->flag, 0, sizeof(int), s->stream));
    TRY(pbx_launch_overlaps(s->pos, s->a, s->Nt, s->N, s->cfg.box_length, s->cfg.overlap_treshold, s->ovl_mask,
                            s->flag, s->stream));
    PBX_HIP(hipMemcpyAsync(s->flag_host, s->flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    *hit = s->flag_host[0];
    return take_info(s);
}

struct StepCtx {
    double sq2dt;
    const double* pn;
    const double* pp;
    int flux_on;
    long long* flux;  // per-label counters the step's moves add their crossings to
    int queued;       // steps queued back to back: no per-step status copies
    const PbxStepCursor* cur = nullptr;  // captured step: noise and flux slots from here
    // The step's tail (wrap, read-backs, look-ahead assembly) and the wait
    // for its read-backs.  A batch of move attempts queues the tail behind
    // its selection, the wrap guarded by the batch's outcome, before the host
    // waits for that outcome; tail_done says the accepted batch's tail ran.
    const std::function<int(const int*, int)>* tail = nullptr;
    const std::function<int()>* wait = nullptr;
    int tail_done = 0;
};

int stoch(PbxSim* s, const double* Bm, const double* xi, double mult, StepCtx& c) {
    int diag = s->cfg.hydrodynamics == PBX_NOHI;
    return pbx_launch_stoch(diag ? s->Bd : Bm, s->ld, diag, xi, c.sq2dt, mult, s->pos, s->disp, s->N, c.pn, c.pp,
                            c.flux_on, s->label, c.flux, s->stream, c.cur);
}

int det(PbxSim* s, const double* Dm, double dt, double mult, StepCtx& c) {
    if (!s->F_nonzero) return PBX_OK;  // translating by D.0 leaves positions unchanged
    int diag = s->cfg.hydrodynamics == PBX_NOHI;
    return pbx_launch_det(diag ? s->Dd : Dm, s->ld, diag, s->F, dt / s->kBT, mult, s->pos, s->N, c.pn, c.pp,
                          c.flux_on, s->label, c.flux, s->stream, c.cur);
}

int undo(PbxSim* s, StepCtx& c) {
    return pbx_launch_translate(s->disp, -1.0, s->pos, s->N, c.pn, c.pp, c.flux_on, s->label, c.flux, s->stream);
}

// box.py:560-610 (applied in box.py:211-215): with lubrication, the
// divergence of D by forward differences: each coordinate in turn moved by
// eps = 1e-7, D rebuilt there (the far-field part only when this step builds
// it), its row for that coordinate summed, the coordinate moved back (so it
// rounds as (r + eps) - eps, like the reference), and the beads translated by
// (sum - row sums of D) / eps * dt.  3N rebuilds of D, as in the reference.
int divergence_step(PbxSim* s, double dt, int bDff, int bDnf, StepCtx& c) {
    const double eps = 0.0000001;
    if (!s->divv) {
        TRY(sim_matrix(s, &s->Ddiv));
        TRY(sim_matrix(s, &s->RffDiv));
        TRY(sim_alloc(s, (void**)&s->divv, sizeof(double) * s->np));
    }
    TRY(pbx_launch_div_init(s->D, s->n, s->ld, s->divv, s->stream));
    for (int k = 0; k < s->n; ++k) {
        TRY(pbx_launch_div_shift(s->pos, k, eps, s->stream));
        if (bDff) TRY(build_Dff(s, nullptr, s->RffDiv));
        if (bDnf) TRY(build_Dtot(s, bDff ? s->RffDiv : s->Rff, s->Ddiv));
        TRY(pbx_launch_div_add_row(bDnf ? s->Ddiv : s->D, k, s->n, s->ld, s->divv, s->stream));
        TRY(pbx_launch_div_shift(s->pos, k, -eps, s->stream));
    }
    TRY(pbx_launch_div_finish(s->divv, s->n, eps, dt, s->drift, s->stream));
    return pbx_launch_translate(s->drift, 1.0, s->pos, s->N, c.pn, c.pp, c.flux_on, s->label, c.flux, s->stream);
}

// box.py:202-250.  With overlap checks the attempts run in batches of K:
// one pass over B yields K candidate moves, one kernel flags those that
// overlap, and the first clean one is applied; the host reads one index per
// batch instead of one flag per attempt.  Results do not depend on K.
int ermak(PbxSim* s, double dt, int bDff, int bDnf, StepCtx& c) {
    TRY(ensure_B(s));
    if (s->F_nonzero || (s->cfg.divergence_term && is_lub(s))) TRY(ensure_D(s));
    TRY(det(s, s->D, dt, 1.0, c));
    if (s->cfg.divergence_term && is_lub(s)) {
        TRY(divergence_step(s, dt, bDff, bDnf, c));
        // the 3N rebuilds factorised after the step's status copy was queued
        if (!c.queued) TRY(post_info(s));
    }
    if (!s->cfg.check_overlaps) {
        const double* xi = s->xi_dev;  // a captured step reads its slot from the cursor; the
        if (!c.cur) TRY(next_xi(s, &xi));  // host makes it resident and consumes it around the replay
        TRY(stoch(s, s->B, xi, 1.0, c));
        s->last_attempts = 1;
        return PBX_OK;
    }
    const int nohi = s->cfg.hydrodynamics == PBX_NOHI;
    const int kmax = pbx_attempt_kmax();
    int want = s->batch_fixed;
    if (want <= 0) {
        want = 2;
        while (want < 2.0 * s->att_mean && want < kmax) want *= 2;
    }
    long long attempts = 0;
    while (true) {
        long long K = want;
        if (K > kmax) K = kmax;
        if (K > s->xi_cap) K = s->xi_cap;
        if (K > s->cfg.max_move_attempts - attempts) K = s->cfg.max_move_attempts - attempts;
        if (K < 1) K = 1;
        TRY(noise_ensure(s, K));
        TRY(pbx_launch_attempt_batch(nohi ? nullptr : s->B, nohi ? s->Bd : nullptr, s->ld, s->np, s->N, s->Nt, s->pos,
                                     s->xi_dev, ring_pos(s), ring_attempts(s), (int)K, c.sq2dt, s->Pbuf, s->cand,
                                     s->cflag, s->a, s->cfg.box_length, s->cfg.overlap_treshold, s->ovl_mask,
                                     s->stream));
        TRY(pbx_launch_attempt_select(s->cand, s->N, (int)K, s->cflag, s->pos, c.pn, c.pp, c.flux_on, s->label,
                                      c.flux, s->flag, s->stream));
        if (c.tail) {  // the tail's wrap kernel publishes the outcome
            TRY((*c.tail)(s->flag, (int)K));
            TRY((*c.wait)());
        } else {
            PBX_HIP(hipMemcpyAsync(s->flag_host, s->flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
            PBX_HIP(hipStreamSynchronize(s->stream));
        }
        TRY(take_info(s));
        const int a = s->flag_host[0];
        if (a < K) {
            attempts += a + 1;
            noise_consume(s, a + 1);
            c.tail_done = c.tail != nullptr;
            break;
        }
        attempts += K;
        noise_consume(s, K);
        if (attempts >= s->cfg.max_move_attempts) {
            s->last_attempts = (int)attempts;
            pbx_set_error(PBX_ERR_MOVES, "maximal number of move attepts exceeded");
            return PBX_ERR_MOVES;
        }
    }
    s->last_attempts = (int)attempts;
    s->att_mean = 0.75 * s->att_mean + 0.25 * (double)attempts;
    return PBX_OK;
}

int chain_solve(PbxSim* s, const double* Lm, const double* W, const double* x, double* y, int trans, int rev = 0,
                double yscale = 1.0) {
    if (++s->chain_epoch == 0) ++s->chain_epoch;
    return pbx_launch_trsv_chain(Lm, s->np, s->ld, W, x, y, s->chain_words, s->chain_ticket, s->chain_epoch, trans,
                                 s->info, s->stream, rev, yscale);
}

// box.py:254-349
int midpoint(PbxSim* s, double dt, int bDff, int bDnf, StepCtx& c) {
    const pbx_config& cf = s->cfg;
    if (cf.hydrodynamics == PBX_NOHI) {
        pbx_set_error(PBX_ERR_ARG, "midpoint scheme needs a matrix diffusion tensor (hydrodynamics != nohi)");
        return PBX_ERR_ARG;
    }
    const double m = cf.m_midpoint;
    const bool lean = lean_midpoint(s, bDnf);
    if (!lean) TRY(ensure_D(s));
    // B0 held as L = chol(J R0 J): B0 xi = sqrt(kBT) J L^{-T} J xi (one
    // solve per attempt) and B0^{-T} xi = J L J xi / sqrt(kBT)
    const bool implicit = lean && s->mat_state == 2;
    const double cB = sqrt(s->kBT);
    const double* xi = nullptr;
    auto stoch0 = [&](double mult) {
        if (implicit)
            return pbx_launch_stoch_vec(s->bxi, c.sq2dt, mult, s->pos, s->disp, s->N, c.pn, c.pp, c.flux_on, s->label,
                                        c.flux, s->stream);
        return stoch(s, s->B, xi, mult, c);
    };
    TRY(det(s, s->D, dt, 1.0 / m, c));
    int attempts = 0;
    while (true) {
        ++attempts;
        TRY(next_xi(s, &xi));
        PBX_HIP(hipMemcpyAsync(s->xipad, xi, sizeof(double) * s->n, hipMemcpyDeviceToDevice, s->stream));
        if (implicit) TRY(chain_solve(s, s->B, s->Wchol, s->xipad, s->bxi, 1, 1, cB));
        TRY(stoch0(1.0 / m));
        if (cf.check_overlaps) {
            int hit = 0;
            TRY(overlap_now(s, &hit));
            if (hit) {
                TRY(undo(s, c));
                continue;
            }
        }
        // D at the midpoint (the reference's midpoint Cholesky is discarded by
        // the B0 restore, box.py:303-316, and is not computed here)
        const double* Dm = s->D;
        if (bDff || (is_lub(s) && bDnf)) {
            if (bDff) TRY(build_Dff(s, s->Dmid, s->RffMid));
            if (lean) {
                // Dmid := L_mid with R_mid = L_mid L_mid^T (tile inverses in Winv)
                TRY(build_R(s, bDff ? s->RffMid : s->Rff, s->Dmid));
                TRY(run_graphed(s, GRAPH_FACTOR_RMID, s->Dmid, [&]() {
                    return pbx_potrf(s->Dmid, s->np, s->ld, s->Winv, s->info, s->stream, 0);
                }));
            } else if (is_lub(s) && bDnf) {
                TRY(build_Dtot(s, bDff ? s->RffMid : s->Rff, s->Dmid));
            }
            if (is_lub(s) || bDff) Dm = s->Dmid;
            if (is_lub(s) && bDff && !bDnf) Dm = s->D;  // self.D untouched by Dff in lub modes
        }
        // stochastic drift: m/2 sqrt(2dt) (D_mid B0^{-T} - B0) xi, w = B0^{-T} xi
        if (implicit)
            TRY(pbx_launch_rev_lower_gemv(s->B, s->np, s->ld, s->xipad, 1.0 / cB, s->w, s->stream));
        else
            TRY(chain_solve(s, s->B, s->Wchol, s->xipad, s->w, 1));
        if (lean) {
            // D_mid w = kBT L_mid^{-T} L_mid^{-1} w
            TRY(chain_solve(s, s->Dmid, s->Winv, s->w, s->u, 0));
            TRY(chain_solve(s, s->Dmid, s->Winv, s->u, s->w, 1));
        }
        // status of the midpoint inverses and of the solves' bounded waits,
        // read at this attempt's overlap check or at the step's end
        if (!c.queued) TRY(post_info(s));
        if (implicit)
            TRY(pbx_launch_drift_vec(s->w, s->bxi, s->n, s->kBT, m / 2 * sqrt(2 * dt), s->drift, s->stream));
        else if (lean)
            TRY(pbx_launch_drift_solved(s->B, s->n, s->ld, s->w, s->xipad, s->kBT, m / 2 * sqrt(2 * dt), s->drift,
                                        s->stream));
        else if (cf.factor_precision)
            // FP32-bulk factor: the same drift as (D_mid - D0) B0^{-T} xi
            TRY(pbx_launch_drift_diff(Dm, s->D, s->n, s->ld, s->w, m / 2 * sqrt(2 * dt), s->drift, s->stream));
        else
            TRY(pbx_launch_drift(Dm, s->B, s->n, s->ld, s->w, s->xipad, m / 2 * sqrt(2 * dt), s->drift, s->stream));
        if (s->F_nonzero)
            TRY(pbx_launch_det_drift(Dm, s->D, s->n, s->ld, s->F, m * dt / (2 * s->kBT), s->drift, s->stream));
        TRY(det(s, s->D, dt, 1.0 - 1.0 / m, c));
        TRY(stoch0(1.0 - 1.0 / m));
        TRY(pbx_launch_translate(s->drift, 1.0, s->pos, s->N, c.pn, c.pp, c.flux_on, s->label, c.flux, s->stream));
        if (!cf.check_overlaps) break;
        int hit = 0;
        TRY(overlap_now(s, &hit));
        if (!hit) break;
        TRY(det(s, s->D, dt, -(1.0 - 1.0 / m), c));
        TRY(stoch0(-1.0));
        TRY(pbx_launch_translate(s->drift, -1.0, s->pos, s->N, c.pn, c.pp, c.flux_on, s->label, c.flux, s->stream));
        if (attempts >= cf.max_move_attempts) {
            s->last_attempts = attempts;
            pbx_set_error(PBX_ERR_MOVES, "maximal number of move attepts exceeded");
            return PBX_ERR_MOVES;
        }
    }
    s->last_attempts = attempts;
    return PBX_OK;
}

}  // namespace

extern "C" int pbx_sim_create(const pbx_config* cfg, int N, const double* radii, const double* xyz,
                              const int* label_idx, pbx_sim** out) {
    pbx_clear_error();
    *out = nullptr;
    if (!cfg || N <= 0 || !radii || !xyz) {
        pbx_set_error(PBX_ERR_ARG, "pbx_sim_create: bad arguments");
        return PBX_ERR_ARG;
    }
    if (cfg->hydrodynamics < PBX_NOHI || cfg->hydrodynamics > PBX_RPY_SMITH_LUB || cfg->scheme < 0 ||
        cfg->scheme > 1) {
        pbx_set_error(PBX_ERR_ARG, "pbx_sim_create: unknown hydrodynamics/scheme");
        return PBX_ERR_ARG;
    }
    if (cfg->factor_precision != 0 && cfg->factor_precision != 1) {
        pbx_set_error(PBX_ERR_ARG, "pbx_sim_create: factor_precision must be 0 (FP64) or 1 (mixed)");
        return PBX_ERR_ARG;
    }
    if (cfg->n_immobile < 0 || (cfg->n_immobile > 0 && (!cfg->immobile_xyz || !cfg->immobile_radii)) ||
        (long long)N + cfg->n_immobile > (1 << 30)) {
        pbx_set_error(PBX_ERR_ARG, "pbx_sim_create: bad immobile beads");
        return PBX_ERR_ARG;
    }
    PbxSim* s = new PbxSim();
    s->cfg = *cfg;
    s->N = N;
    s->Nt = N + cfg->n_immobile;
    s->n = 3 * N;
    s->np = (int)pbx_round_up(s->n, 64);
    s->ld = s->np;
    s->kBT = 1.380649e-23 * cfg->T;  // scipy.constants.Boltzmann * T (box.py:636)
    s->xi_cap = cfg->xi_pool_attempts > 0 ? cfg->xi_pool_attempts : 64;
    int rc = PBX_OK;
    auto fail = [&](int code) {
        pbx_sim_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        pbx_set_error(PBX_ERR_HIP, "hipStreamCreate failed");
        delete s;
        return PBX_ERR_HIP;
    }
    const size_t n3 = s->n;
    const int Ni = cfg->n_immobile;
    if ((rc = sim_alloc(s, (void**)&s->pos, sizeof(double) * 3 * (size_t)s->Nt))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->a, sizeof(double) * s->Nt))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->F, sizeof(double) * s->np))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->disp, sizeof(double) * s->np))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->drift, sizeof(double) * s->np))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->w, sizeof(double) * s->np))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->xipad, sizeof(double) * s->np))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->label, sizeof(int) * N))) return fail(rc);
    int nl = cfg->n_labels > 0 ? cfg->n_labels : 1;
    if ((rc = sim_alloc(s, (void**)&s->flux, sizeof(long long) * nl))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->flux_alt, sizeof(long long) * nl))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->flag, sizeof(int) * 2))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->info, sizeof(int)))) return fail(rc);
    if ((rc = sim_alloc(s, (void**)&s->xi_dev, sizeof(double) * 2 * s->xi_cap * n3))) return fail(rc);
    if (cfg->scheme == PBX_ERMAK && cfg->check_overlaps) {
        const int kmax = pbx_attempt_kmax();
        // K + 1 candidate rows, then N displacement bounds
        if ((rc = sim_alloc(s, (void**)&s->cand, sizeof(double) * ((kmax + 1) * n3 + N)))) return fail(rc);
        if ((rc = sim_alloc(s, (void**)&s->cflag, sizeof(int) * kmax))) return fail(rc);
        if (cfg->hydrodynamics != PBX_NOHI &&
            (rc = sim_alloc(s, (void**)&s->Pbuf, sizeof(double) * pbx_attempt_partial_doubles(s->np))))
            return fail(rc);
    }
    for (int k = 0; k < 3; ++k)
        if (hipEventCreateWithFlags(&s->xi_ev[k], hipEventDisableTiming) != hipSuccess) {
            pbx_set_error(PBX_ERR_HIP, "hipEventCreate failed");
            return fail(PBX_ERR_HIP);
        }
    if (hipEventCreateWithFlags(&s->rd_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->done_ev, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking) != hipSuccess) {
        pbx_set_error(PBX_ERR_HIP, "upload stream or event creation failed");
        return fail(PBX_ERR_HIP);
    }
    if (hipHostMalloc((void**)&s->xi_host, sizeof(double) * 3 * s->xi_cap * n3, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s->flag_host, sizeof(int) * 4, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s->pos_host, sizeof(double) * n3, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s->flux_host, sizeof(long long) * nl, hipHostMallocDefault) != hipSuccess) {
        pbx_set_error(PBX_ERR_OOM, "hipHostMalloc failed");
        return fail(PBX_ERR_OOM);
    }
    if (hipHostGetDevicePointer((void**)&s->flag_hd, s->flag_host, 0) != hipSuccess ||
        hipHostGetDevicePointer((void**)&s->pos_hd, s->pos_host, 0) != hipSuccess ||
        hipHostGetDevicePointer((void**)&s->flux_hd, s->flux_host, 0) != hipSuccess) {
        pbx_set_error(PBX_ERR_HIP, "hipHostGetDevicePointer failed");
        return fail(PBX_ERR_HIP);
    }
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e != hipSuccess) {
            pbx_set_error(PBX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
            return false;
        }
        return true;
    };
    // immobile beads (box.py:649-668) follow the mobile ones and never move
    if (!hip_ok(hipMemcpyAsync(s->pos, xyz, sizeof(double) * n3, hipMemcpyHostToDevice, s->stream), "pos") ||
        !hip_ok(hipMemcpyAsync(s->a, radii, sizeof(double) * N, hipMemcpyHostToDevice, s->stream), "radii") ||
        (Ni > 0 && !hip_ok(hipMemcpyAsync(s->pos + n3, cfg->immobile_xyz, sizeof(double) * 3 * (size_t)Ni,
                                          hipMemcpyHostToDevice, s->stream),
                           "immobile xyz")) ||
        (Ni > 0 && !hip_ok(hipMemcpyAsync(s->a + N, cfg->immobile_radii, sizeof(double) * Ni, hipMemcpyHostToDevice,
                                          s->stream),
                           "immobile radii")) ||
        !hip_ok(hipMemsetAsync(s->xipad, 0, sizeof(double) * s->np, s->stream), "xipad") ||
        !hip_ok(hipMemsetAsync(s->info, 0, sizeof(int), s->stream), "info") ||
        !hip_ok(hipMemsetAsync(s->flux, 0, sizeof(long long) * nl, s->stream), "flux"))
        return fail(PBX_ERR_HIP);
    std::vector<int> lab(N, 0);
    if (label_idx)
        for (int i = 0; i < N; ++i) lab[i] = label_idx[i];
    std::vector<double> F(s->np, 0.0);
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) {
            F[3 * i + c] = cfg->external_force[c];
            if (cfg->external_force[c] != 0.0) s->F_nonzero = 1;
        }
    if (!hip_ok(sim_copy(s, s->label, lab.data(), sizeof(int) * N, hipMemcpyHostToDevice), "labels") ||
        !hip_ok(sim_copy(s, s->F, F.data(), sizeof(double) * s->np, hipMemcpyHostToDevice), "F"))
        return fail(PBX_ERR_HIP);

    if (cfg->hydrodynamics == PBX_NOHI) {
        // box.py:488-492 and :550-552: D = kBT 1e19/6/pi/a/eta, B = sqrt(D)
        std::vector<double> Dd(n3), Bd(n3);
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) {
                double d = s->kBT * 1e19 / 6 / 3.14159265358979323846 / radii[i] / cfg->viscosity;
                Dd[3 * i + c] = d;
                Bd[3 * i + c] = sqrt(d);
            }
        if ((rc = sim_alloc(s, (void**)&s->Dd, sizeof(double) * n3))) return fail(rc);
        if ((rc = sim_alloc(s, (void**)&s->Bd, sizeof(double) * n3))) return fail(rc);
        if (!hip_ok(sim_copy(s, s->Dd, Dd.data(), sizeof(double) * n3, hipMemcpyHostToDevice), "Dd") ||
            !hip_ok(sim_copy(s, s->Bd, Bd.data(), sizeof(double) * n3, hipMemcpyHostToDevice), "Bd"))
            return fail(PBX_ERR_HIP);
        s->have_D = s->have_B = 1;
    } else {
        if ((rc = sim_matrix(s, &s->D))) return fail(rc);
        if ((rc = sim_matrix(s, &s->B))) return fail(rc);
        if (!is_lub(s) && (rc = sim_matrix(s, &s->Dspec))) return fail(rc);
        if ((rc = sim_alloc(s, (void**)&s->Wchol, sizeof(double) * pbx_dense_work_doubles(s->np)))) return fail(rc);
        if (cfg->scheme == PBX_MIDPOINT) {
            if ((rc = sim_matrix(s, &s->Dmid))) return fail(rc);
            if ((rc = sim_alloc(s, (void**)&s->chain_words, sizeof(unsigned long long) * 2 * s->np)))
                return fail(rc);
            if ((rc = sim_alloc(s, (void**)&s->chain_ticket, sizeof(int)))) return fail(rc);
            if ((rc = sim_alloc(s, (void**)&s->u, sizeof(double) * s->np))) return fail(rc);
            if ((rc = sim_alloc(s, (void**)&s->bxi, sizeof(double) * s->np))) return fail(rc);
            if (!hip_ok(hipMemsetAsync(s->chain_words, 0, sizeof(unsigned long long) * 2 * s->np, s->stream),
                        "solve words") ||
                !hip_ok(hipMemsetAsync(s->chain_ticket, 0, sizeof(int), s->stream), "solve ticket") ||
                !hip_ok(hipMemsetAsync(s->u, 0, sizeof(double) * s->np, s->stream), "solve vector"))
                return fail(PBX_ERR_HIP);
        }
        if (is_lub(s)) {
            if ((rc = sim_matrix(s, &s->Rff))) return fail(rc);
            if ((rc = sim_matrix(s, &s->X))) return fail(rc);
            if ((rc = sim_alloc(s, (void**)&s->Winv, sizeof(double) * pbx_dense_work_doubles(s->np))))
                return fail(rc);
            if (cfg->scheme == PBX_MIDPOINT)
                if ((rc = sim_matrix(s, &s->RffMid))) return fail(rc);
        }
    }
    if (cfg->n_overlap_exempt > 0) {
        // bead.py:401-413 build_connection_matrix: pairs the overlap test
        // skips (pairs of two immobile beads are skipped without it)
        const int Nt = s->Nt;
        if (!cfg->overlap_exempt) {
            pbx_set_error(PBX_ERR_ARG, "pbx_sim_create: overlap_exempt is null");
            return fail(PBX_ERR_ARG);
        }
        std::vector<int> m((size_t)Nt * Nt, 0);
        for (int q = 0; q < cfg->n_overlap_exempt; ++q) {
            const int i = cfg->overlap_exempt[2 * q], j = cfg->overlap_exempt[2 * q + 1];
            if (i < 0 || i >= Nt || j < 0 || j >= Nt) {
                pbx_set_error(PBX_ERR_ARG, "pbx_sim_create: overlap exemption index out of range");
                return fail(PBX_ERR_ARG);
            }
            m[(size_t)i * Nt + j] = m[(size_t)j * Nt + i] = 1;
        }
        if ((rc = sim_alloc(s, (void**)&s->ovl_mask, sizeof(int) * m.size()))) return fail(rc);
        if (!hip_ok(sim_copy(s, s->ovl_mask, m.data(), sizeof(int) * m.size(), hipMemcpyHostToDevice), "mask"))
            return fail(PBX_ERR_HIP);
    }
    if (!hip_ok(hipStreamSynchronize(s->stream), "create sync")) return fail(PBX_ERR_HIP);
    if (cfg->check_overlaps) {
        int hit = 0;
        if ((rc = overlap_now(s, &hit))) return fail(rc);
        if (hit) {
            pbx_set_error(PBX_ERR_OVERLAP, "starting configuration contains overlaps");
            return fail(PBX_ERR_OVERLAP);
        }
    }
    *out = s;
    return PBX_OK;
}

extern "C" int pbx_sim_destroy(pbx_sim* s) {
    if (!s) return PBX_OK;
    noise_stop(s);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->up) (void)hipStreamSynchronize(s->up);
    for (int k = 0; k < 8; ++k)
        if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
    if (s->done_ev) (void)hipEventDestroy(s->done_ev);
    for (int k = 0; k < 3; ++k)
        if (s->xi_ev[k]) (void)hipEventDestroy(s->xi_ev[k]);
    drop_graphs(s);
    {
        std::unique_lock<std::shared_mutex> free_lock(g_capture_mu);
        for (void* p : s->owned) (void)hipFree(p);
    }
    if (s->xi_host) (void)hipHostFree(s->xi_host);
    if (s->flag_host) (void)hipHostFree(s->flag_host);
    if (s->pos_host) (void)hipHostFree(s->pos_host);
    if (s->flux_host) (void)hipHostFree(s->flux_host);
    if (s->E_host) (void)hipHostFree(s->E_host);
    if (s->cursor_host) (void)hipHostFree(s->cursor_host);
    if (s->rd_ev) (void)hipEventDestroy(s->rd_ev);
    if (s->up) (void)hipStreamDestroy(s->up);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return PBX_OK;
}

extern "C" int pbx_sim_set_noise_source(pbx_sim* s, pbx_xi_fill_fn fn, void* user) {
    pbx_clear_error();
    TRY(noise_reset(s));
    s->xi_fill = fn;
    s->xi_user = user;
    return PBX_OK;
}

extern "C" int pbx_sim_set_seed(pbx_sim* s, unsigned int seed) {
    pbx_clear_error();
    TRY(noise_reset(s));
    pbx_mt_seed(&s->mt, seed);
    s->mt_start = s->mt;
    s->mt_base = 0;
    s->native = 1;
    try {
        s->producer = std::thread(producer_main, s);
    } catch (...) {
        s->native = 0;
        pbx_set_error(PBX_ERR_STATE, "could not start the noise producer thread");
        return PBX_ERR_STATE;
    }
    return PBX_OK;
}

// ---- snapshot / restore ------------------------------------------------------
// Layout: header | positions 3N | flux n_labels | unconsumed draws |
// [generator state] | D, B [, Rff] (np x ld each) | diagonal inverses.

namespace {
struct SnapHeader {
    char magic[8];
    int N, np, hydro, scheme, n_labels, native;
    int have_D, have_B, have_Rff, last_attempts;
    long long draw_count, xi_used, xi_taken, xi_cap;
    long long pending, nmat, wdoubles;
    double att_mean;
};
const char SNAP_MAGIC[8] = {'P', 'B', 'X', 'S', 'N', 'A', 'P', '1'};

int snap_nmat(const PbxSim* s) {
    if (s->cfg.hydrodynamics == PBX_NOHI) return 0;
    return is_lub(s) ? 3 : 2;
}
size_t snap_bytes(const PbxSim* s, long long pending) {
    const int nl = s->cfg.n_labels > 0 ? s->cfg.n_labels : 1;
    const size_t mat = sizeof(double) * (size_t)s->np * s->ld;
    const size_t w = s->Wchol ? sizeof(double) * pbx_dense_work_doubles(s->np) : 0;
    return sizeof(SnapHeader) + sizeof(double) * s->n + sizeof(long long) * nl + sizeof(double) * pending * s->n +
           (s->native ? sizeof(PbxMT) : 0) + mat * snap_nmat(s) + w;
}
long long pending_attempts(const PbxSim* s) { return s->xi_taken * s->xi_cap - s->xi_used; }

// copy attempts [a0, a0 + k) of the device ring to/from host memory
int ring_copy(PbxSim* s, long long a0, long long k, double* host, bool to_host) {
    for (long long a = a0; a < a0 + k; ++a) {
        double* dev = s->xi_dev + (size_t)(a % (2 * s->xi_cap)) * s->n;
        double* h = host + (size_t)(a - a0) * s->n;
        if (to_host)
            PBX_HIP(sim_copy(s, h, dev, sizeof(double) * s->n, hipMemcpyDeviceToHost));
        else
            PBX_HIP(sim_copy(s, dev, h, sizeof(double) * s->n, hipMemcpyHostToDevice));
    }
    return PBX_OK;
}
}  // namespace

extern "C" long long pbx_sim_snapshot_size(pbx_sim* s) {
    std::lock_guard<std::mutex> lk(s->mu);
    return (long long)snap_bytes(s, pending_attempts(s));
}

extern "C" int pbx_sim_snapshot(pbx_sim* s, void* buf, long long bytes) {
    pbx_clear_error();
    TRY(ensure_D(s));
    PBX_HIP(hipStreamSynchronize(s->stream));
    SnapHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, SNAP_MAGIC, 8);
    PbxMT mt_now;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        h.xi_taken = s->xi_taken;
        if (s->native) mt_now = (s->xi_taken == s->mt_base) ? s->mt_start : s->mt_after[(s->xi_taken - 1) % 4];
    }
    h.N = s->N;
    h.np = s->np;
    h.hydro = s->cfg.hydrodynamics;
    h.scheme = s->cfg.scheme;
    h.n_labels = s->cfg.n_labels > 0 ? s->cfg.n_labels : 1;
    h.native = s->native;
    h.have_D = s->have_D;
    h.have_B = s->have_B;
    h.have_Rff = s->have_Rff;
    h.last_attempts = s->last_attempts;
    h.draw_count = s->draw_count;
    h.xi_used = s->xi_used;
    h.xi_cap = s->xi_cap;
    h.pending = pending_attempts(s);
    h.nmat = snap_nmat(s);
    h.wdoubles = s->Wchol ? (long long)pbx_dense_work_doubles(s->np) : 0;
    h.att_mean = s->att_mean;
    if (!buf || bytes < (long long)snap_bytes(s, h.pending)) {
        pbx_set_error(PBX_ERR_ARG, "snapshot buffer too small (%lld < %zu bytes)", bytes, snap_bytes(s, h.pending));
        return PBX_ERR_ARG;
    }
    char* p = (char*)buf;
    memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    PBX_HIP(sim_copy(s, p, s->pos, sizeof(double) * s->n, hipMemcpyDeviceToHost));
    p += sizeof(double) * s->n;
    PBX_HIP(sim_copy(s, p, s->flux, sizeof(long long) * h.n_labels, hipMemcpyDeviceToHost));
    p += sizeof(long long) * h.n_labels;
    TRY(ring_copy(s, s->xi_used, h.pending, (double*)p, true));
    p += sizeof(double) * h.pending * s->n;
    if (s->native) {
        memcpy(p, &mt_now, sizeof(PbxMT));
        p += sizeof(PbxMT);
    }
    double* mats[3] = {s->D, s->B, s->Rff};
    const size_t mat = sizeof(double) * (size_t)s->np * s->ld;
    for (int m = 0; m < h.nmat; ++m) {
        PBX_HIP(sim_copy(s, p, mats[m], mat, hipMemcpyDeviceToHost));
        p += mat;
    }
    if (h.wdoubles) PBX_HIP(sim_copy(s, p, s->Wchol, sizeof(double) * h.wdoubles, hipMemcpyDeviceToHost));
    return PBX_OK;
}

extern "C" int pbx_sim_restore(pbx_sim* s, const void* buf, long long bytes) {
    pbx_clear_error();
    s->spec_ok = 0;
    SnapHeader h;
    if (!buf || bytes < (long long)sizeof(h)) {
        pbx_set_error(PBX_ERR_ARG, "snapshot too short");
        return PBX_ERR_ARG;
    }
    memcpy(&h, buf, sizeof(h));
    const int nl = s->cfg.n_labels > 0 ? s->cfg.n_labels : 1;
    if (memcmp(h.magic, SNAP_MAGIC, 8) || h.N != s->N || h.np != s->np || h.hydro != s->cfg.hydrodynamics ||
        h.scheme != s->cfg.scheme || h.n_labels != nl || h.xi_cap != s->xi_cap || h.nmat != snap_nmat(s) ||
        h.pending < 0 || h.pending > 2 * h.xi_cap) {
        pbx_set_error(PBX_ERR_ARG, "snapshot does not match this simulation (size, mode or noise chunking)");
        return PBX_ERR_ARG;
    }
    if (!h.native && !s->xi_fill) {
        pbx_set_error(PBX_ERR_STATE, "snapshot uses a noise callback: register one before restoring");
        return PBX_ERR_STATE;
    }
    {
        const int was = s->native;
        s->native = h.native;  // the expected size depends on it
        const bool ok = bytes >= (long long)snap_bytes(s, h.pending);
        s->native = was;
        if (!ok) {
            pbx_set_error(PBX_ERR_ARG, "snapshot too short");
            return PBX_ERR_ARG;
        }
    }
    // stop the producer, keep a registered callback
    pbx_xi_fill_fn fill = s->xi_fill;
    void* user = s->xi_user;
    TRY(noise_reset(s));
    s->xi_fill = fill;
    s->xi_user = user;
    s->native = h.native;
    const char* p = (const char*)buf + sizeof(h);
    PBX_HIP(sim_copy(s, s->pos, p, sizeof(double) * s->n, hipMemcpyHostToDevice));
    p += sizeof(double) * s->n;
    PBX_HIP(sim_copy(s, s->flux, p, sizeof(long long) * nl, hipMemcpyHostToDevice));
    p += sizeof(long long) * nl;
    s->xi_used = h.xi_used;
    s->xi_taken = h.xi_taken;
    s->xi_produced = h.xi_taken;
    TRY(ring_copy(s, h.xi_used, h.pending, (double*)p, false));
    p += sizeof(double) * h.pending * s->n;
    if (h.native) {
        memcpy(&s->mt, p, sizeof(PbxMT));
        p += sizeof(PbxMT);
        s->mt_start = s->mt;
        s->mt_base = h.xi_taken;
    }
    double* mats[3] = {s->D, s->B, s->Rff};
    const size_t mat = sizeof(double) * (size_t)s->np * s->ld;
    for (int m = 0; m < h.nmat; ++m) {
        PBX_HIP(sim_copy(s, mats[m], p, mat, hipMemcpyHostToDevice));
        p += mat;
    }
    if (h.wdoubles && s->Wchol)
        PBX_HIP(sim_copy(s, s->Wchol, p, sizeof(double) * std::min<long long>(h.wdoubles, pbx_dense_work_doubles(s->np)),
                          hipMemcpyHostToDevice));
    s->have_D = h.have_D;
    s->mat_state = 0;
    s->have_B = h.have_B;
    s->have_Rff = h.have_Rff;
    s->last_attempts = h.last_attempts;
    s->draw_count = h.draw_count;
    s->att_mean = h.att_mean;
    if (h.native) {
        try {
            s->producer = std::thread(producer_main, s);
        } catch (...) {
            pbx_set_error(PBX_ERR_STATE, "could not start the noise producer thread");
            return PBX_ERR_STATE;
        }
    }
    return PBX_OK;
}

extern "C" int pbx_sim_set_attempt_batch(pbx_sim* s, int K) {
    if (K < 0 || K > pbx_attempt_kmax()) {
        pbx_set_error(PBX_ERR_ARG, "attempt batch must lie in [0, %d] (0 = adaptive)", pbx_attempt_kmax());
        return PBX_ERR_ARG;
    }
    s->batch_fixed = K;
    return PBX_OK;
}

extern "C" int pbx_sim_set_graphs(pbx_sim* s, int on) {
    pbx_clear_error();
    if (!s || (on != 0 && on != 1)) {
        pbx_set_error(PBX_ERR_ARG, "set_graphs: on must be 0 or 1");
        return PBX_ERR_ARG;
    }
    s->spec_ok = 0;  // graphs hold the matrix pointers the look-ahead swaps
    if (!on) {
        PBX_HIP(hipStreamSynchronize(s->stream));
        drop_graphs(s);
    }
    s->graphs = on;
    return PBX_OK;
}

extern "C" int pbx_sim_set_inverse_precision(pbx_sim* s, int precision) {
    pbx_clear_error();
    if (!s || precision < 0 || precision > 2) {
        pbx_set_error(PBX_ERR_ARG, "set_inverse_precision: precision must be 0, 1 or 2");
        return PBX_ERR_ARG;
    }
    if (precision == 2 && is_lub(s) && !s->Xref) TRY(sim_alloc(s, (void**)&s->Xref, sizeof(double) * (size_t)s->np * s->ld));
    if (precision != s->inv_precision) {
        // captured inverse sequences hold the previous precision's kernels
        PBX_HIP(hipStreamSynchronize(s->stream));
        drop_graphs(s);
        s->spec_ok = 0;
    }
    s->inv_precision = precision;
    return PBX_OK;
}

// box.py:150-198 propagate
extern "C" int pbx_sim_propagate(pbx_sim* s, double dt, int build_Dff_flag, int build_Dnf_flag, int cholesky_flag) {
    pbx_clear_error();
    const pbx_config& cf = s->cfg;
    StepCtx c;
    c.sq2dt = sqrt(2 * dt);
    c.pn = cf.flux_normal;
    c.pp = cf.flux_plane_point;
    c.flux_on = cf.measure_flux;
    c.queued = 0;
    if (cf.measure_flux) {
        if (s->flux_alt_zero) {
            std::swap(s->flux, s->flux_alt);  // cleared by the last step's wrap kernel
        } else {
            int nl = cf.n_labels > 0 ? cf.n_labels : 1;
            PBX_HIP(hipMemsetAsync(s->flux, 0, sizeof(long long) * nl, s->stream));
        }
    }
    s->flux_alt_zero = 0;
    c.flux = s->flux;
    if (s->timing == 1) PBX_HIP(hipEventRecord(s->ev[0], s->stream));
    const int spec = s->spec_ok;
    s->spec_ok = 0;
    if (cf.hydrodynamics != PBX_NOHI) {
        if (build_Dff_flag) {
            if (spec)
                std::swap(s->D, s->Dspec);  // assembled for these positions at the end of the last step
            else
                TRY(build_Dff(s, s->D, s->Rff));
            if (is_lub(s)) s->have_Rff = 1; else s->have_D = 1;
        }
        const bool fused = build_Dnf_flag && cholesky_flag && fuse_inverse_and_factor(s);
        if (is_lub(s) && build_Dnf_flag) {
            if (!s->have_Rff) {
                pbx_set_error(PBX_ERR_STATE, "lubrication step before the far-field matrix was built");
                return PBX_ERR_STATE;
            }
            TRY(build_Dtot(s, s->Rff, s->D, fused));
            s->have_D = 1;
            if (fused) s->have_B = 1;
        }
        if (s->timing) PBX_HIP(hipEventRecord(s->ev[1], s->stream));
        if (cholesky_flag && !fused) {
            if (!s->have_D) {
                pbx_set_error(PBX_ERR_STATE, "Cholesky requested before D was built");
                return PBX_ERR_STATE;
            }
            TRY(factor(s));
            s->have_B = 1;
        }
        if (s->timing) PBX_HIP(hipEventRecord(s->ev[2], s->stream));
        if (!s->have_B) {
            pbx_set_error(PBX_ERR_STATE, "propagation without a noise factor");
            return PBX_ERR_STATE;
        }
        if (cf.scheme == PBX_ERMAK)
            s->info_pending = 1;  // the step's wrap kernel publishes the status word
        else
            TRY(post_info(s));
    }
    // box.py:190: forces from the positions at the start of the step
    if (s->forces_on) TRY(pbx_launch_forces(s->fa, s->pos, s->F, s->e_part, s->E_dev, s->stream));
    const bool timed = s->timing && cf.hydrodynamics != PBX_NOHI;
    const bool staged = timed && s->timing == 1;  // all stage bounds, not just the factorisation's
    const bool ahead = s->Dspec && s->lookahead && !s->graphs && build_Dff_flag && s->next_builds != 0 &&
                       cf.hydrodynamics != PBX_NOHI;
    const int prev = s->ahead_pending, q = prev == 0 ? 1 : 0;
    // wrap (bead.py:170-183), the positions, energy and flux read-backs, and
    // the next step's far-field assembly from the wrapped positions
    PbxWrapOut wo{};
    wo.pos = s->pos_hd;
    wo.status = s->flag_hd;
    wo.info = s->info;
    if (cf.measure_flux) {
        wo.flux = s->flux_hd;
        wo.flux_dev = s->flux;
        wo.flux_clear = s->flux_alt;
        wo.nl = cf.n_labels > 0 ? cf.n_labels : 1;
    }
    if (s->forces_on && s->E_hd) {
        wo.E = s->E_hd;
        wo.E_dev = s->E_dev;
    }
    const std::function<int(const int*, int)> tail = [&](const int* sel, int K) -> int {
        TRY(pbx_launch_wrap_publish(s->pos, s->N, cf.box_length, sel, K, wo, s->stream));
        if (staged) PBX_HIP(hipEventRecord(s->ev[3], s->stream));
        if (ahead) {
            PBX_HIP(hipEventRecord(s->done_ev, s->stream));
            if (s->timing == 1) PBX_HIP(hipEventRecord(s->ev[4 + 2 * q], s->stream));
            TRY(assemble_M(s, s->Dspec, s->kBT * 1e19 / cf.viscosity, sel, K));  // as build_Dff
            if (s->timing == 1) PBX_HIP(hipEventRecord(s->ev[5 + 2 * q], s->stream));
        }
        return PBX_OK;
    };
    // the read-backs are complete; a look-ahead assembly may still run
    const std::function<int()> wait = [&]() -> int {
        if (ahead)
            PBX_HIP(hipEventSynchronize(s->done_ev));
        else
            PBX_HIP(hipStreamSynchronize(s->stream));
        return PBX_OK;
    };
    if (cf.scheme == PBX_ERMAK) {
        c.tail = &tail;
        c.wait = &wait;
        TRY(ermak(s, dt, build_Dff_flag, build_Dnf_flag, c));
    } else {
        TRY(midpoint(s, dt, build_Dff_flag, build_Dnf_flag, c));
    }
    if (!c.tail_done) {
        TRY(tail(nullptr, 0));
        TRY(wait());
    }
    if (timed) {  // the events are complete: read them after the one synchronisation
        float t01 = 0, t12 = 0, t23 = 0;
        PBX_HIP(hipEventElapsedTime(&t12, s->ev[1], s->ev[2]));
        if (staged) {
            PBX_HIP(hipEventElapsedTime(&t01, s->ev[0], s->ev[1]));
            PBX_HIP(hipEventElapsedTime(&t23, s->ev[2], s->ev[3]));
        }
        s->stage_ms[0] += t01;
        s->stage_ms[1] += t12;
        s->stage_ms[2] += t23;
        s->timed_steps += 1;
    }
    if (s->timing == 1 && prev >= 0) {  // the last step's look-ahead assembly preceded this step's kernels
        float ta = 0;
        PBX_HIP(hipEventElapsedTime(&ta, s->ev[4 + 2 * prev], s->ev[5 + 2 * prev]));
        s->stage_ms[0] += ta;
    }
    s->ahead_pending = (ahead && s->timing == 1) ? q : -1;
    const int rc = take_info(s);
    if (rc == PBX_OK && ahead) s->spec_ok = 1;
    if (rc == PBX_OK && cf.measure_flux) s->flux_alt_zero = 1;
    // the wait above covered every reader of the ring (the move's kernels)
    if (rc == PBX_OK && cf.scheme == PBX_ERMAK) TRY(noise_prefetch(s));
    return rc;
}

namespace {

// One Box.propagate queued on the stream with no host synchronisation: the
// matrix builds and factorisation its flags ask for, then the Ermak or
// midpoint move with no overlap test, then the wrap.  Crossings add into
// flux_slot (zero beforehand).  The same kernels as pbx_sim_propagate with
// check_overlaps off, so the same bits.
int queued_step(PbxSim* s, double dt, int bDff, int bDnf, int chol, long long* flux_slot,
                const PbxStepCursor* cur = nullptr) {
    const pbx_config& cf = s->cfg;
    StepCtx c;
    c.sq2dt = sqrt(2 * dt);
    c.pn = cf.flux_normal;
    c.pp = cf.flux_plane_point;
    c.flux_on = cf.measure_flux;
    c.flux = flux_slot;
    c.queued = 1;
    c.cur = cur;
    if (cf.hydrodynamics != PBX_NOHI) {
        if (bDff) {
            TRY(build_Dff(s, s->D, s->Rff));
            if (is_lub(s)) s->have_Rff = 1; else s->have_D = 1;
        }
        const bool fused = bDnf && chol && fuse_inverse_and_factor(s);
        if (is_lub(s) && bDnf) {
            if (!s->have_Rff) {
                pbx_set_error(PBX_ERR_STATE, "lubrication step before the far-field matrix was built");
                return PBX_ERR_STATE;
            }
            TRY(build_Dtot(s, s->Rff, s->D, fused));
            s->have_D = 1;
            if (fused) s->have_B = 1;
        }
        if (chol && !fused) {
            if (!s->have_D) {
                pbx_set_error(PBX_ERR_STATE, "Cholesky requested before D was built");
                return PBX_ERR_STATE;
            }
            TRY(factor(s));
            s->have_B = 1;
        }
        if (!s->have_B) {
            pbx_set_error(PBX_ERR_STATE, "propagation without a noise factor");
            return PBX_ERR_STATE;
        }
    }
    if (s->forces_on) TRY(pbx_launch_forces(s->fa, s->pos, s->F, s->e_part, s->E_dev, s->stream));
    if (cf.scheme == PBX_ERMAK)
        TRY(ermak(s, dt, bDff, bDnf, c));
    else
        TRY(midpoint(s, dt, bDff, bDnf, c));
    return pbx_launch_wrap(s->pos, s->N, cf.box_length, s->stream);
}

// Whole steps replay from graphs when graphs are on and the step is a fixed
// launch sequence: Ermak without overlap checks (one attempt per step, no
// host decision) in the RPY or Smith-Ewald modes without the divergence
// term.
bool step_graphs_apply(const PbxSim* s) {
    const pbx_config& cf = s->cfg;
    return s->graphs && cf.scheme == PBX_ERMAK && !cf.check_overlaps && !cf.divergence_term &&
           (cf.hydrodynamics == PBX_RPY || cf.hydrodynamics == PBX_RPY_SMITH);
}

// Replay (capturing on first use) the step with flags (bDff, chol) at time
// step dt; the noise slot the cursor points at is resident.
int step_graph_launch(PbxSim* s, double dt, int bDff, int chol) {
    if (s->step_graph_dt != dt) {  // dt is baked into the captured launches
        PBX_HIP(hipStreamSynchronize(s->stream));
        int k = 0;
        for (int q = 0; q < s->n_graphs; ++q) {
            if (s->graph_slots[q].op >= GRAPH_STEP)
                (void)hipGraphExecDestroy(s->graph_slots[q].exec);
            else
                s->graph_slots[k++] = s->graph_slots[q];
        }
        s->n_graphs = k;
        s->step_graph_dt = dt;
    }
    const int op = GRAPH_STEP + 2 * (bDff != 0) + (chol != 0);
    const int gen = pbx_potrf_config_generation();
    for (int q = 0; q < s->n_graphs; ++q) {
        const PbxSim::GraphSlot& g = s->graph_slots[q];
        if (g.op == op && g.key == (const void*)s->cursor_dev && g.variant == gen) {
            PBX_HIP(hipGraphLaunch(g.exec, s->stream));
            return PBX_OK;
        }
    }
    auto body = [&]() {
        TRY(queued_step(s, dt, bDff, bDff, chol, nullptr, s->cursor_dev));
        return pbx_launch_cursor_advance(s->cursor_dev, s->stream);
    };
    if (s->n_graphs == (int)(sizeof(s->graph_slots) / sizeof(s->graph_slots[0]))) return body();
    std::shared_lock<std::shared_mutex> capture_lock(g_capture_mu);
    PBX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    s->capturing = 1;
    const int rc = body();
    s->capturing = 0;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s->stream, &graph);
    capture_lock.unlock();
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        pbx_set_error(PBX_ERR_HIP, "step graph capture: %s", hipGetErrorString(e));
        return PBX_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        pbx_set_error(PBX_ERR_HIP, "step graph instantiate: %s", hipGetErrorString(ei));
        return PBX_ERR_HIP;
    }
    s->graph_slots[s->n_graphs++] = {op, (const void*)s->cursor_dev, gen, exec};
    PBX_HIP(hipGraphLaunch(exec, s->stream));
    return PBX_OK;
}

// Steps j0 .. j0+steps-1 queued back to back (pbx_sim_run when no step
// needs a host decision): one flux slot per step, zeroed once; the status
// word is copied every 32 steps and read one copy later, so the host stays up
// to 64 steps ahead of the device and a failure is reported within 64 steps;
// positions, flux and energy come back once at the end.
int run_queued(PbxSim* s, double dt, long long j0, long long steps, int diff_freq, int lub_freq, int chol_freq,
               long long* flux_out) {
    const int nl = s->cfg.n_labels > 0 ? s->cfg.n_labels : 1;
    const long long need = steps * nl;
    if (need > s->flux_log_cap) {
        if (s->flux_log) {
            PBX_HIP(hipStreamSynchronize(s->stream));
            {
                std::unique_lock<std::shared_mutex> free_lock(g_capture_mu);
                PBX_HIP(hipFree(s->flux_log));
            }
            s->owned.erase(std::find(s->owned.begin(), s->owned.end(), (void*)s->flux_log));
            s->device_bytes -= sizeof(long long) * (size_t)s->flux_log_cap;
            s->flux_log = nullptr;
        }
        TRY(sim_alloc(s, (void**)&s->flux_log, sizeof(long long) * (size_t)need));
        s->flux_log_cap = need;
    }
    PBX_HIP(hipMemsetAsync(s->flux_log, 0, sizeof(long long) * (size_t)need, s->stream));
    s->spec_ok = 0;  // D is assembled at the start of each step here
    const bool graphed = step_graphs_apply(s);
    if (graphed) {
        if (!s->cursor_dev) {
            TRY(sim_alloc(s, (void**)&s->cursor_dev, sizeof(PbxStepCursor)));
            if (hipHostMalloc((void**)&s->cursor_host, sizeof(PbxStepCursor), hipHostMallocDefault) != hipSuccess) {
                s->cursor_host = nullptr;
                pbx_set_error(PBX_ERR_OOM, "hipHostMalloc failed");
                return PBX_ERR_OOM;
            }
        }
        // the previous run ended in a stream synchronisation: the pinned
        // struct is free to rewrite
        *s->cursor_host = {0, (long long)ring_pos(s), (long long)ring_attempts(s), (long long)s->n, (long long)nl,
                           s->xi_dev, s->flux_log};
        PBX_HIP(hipMemcpyAsync(s->cursor_dev, s->cursor_host, sizeof(PbxStepCursor), hipMemcpyHostToDevice,
                               s->stream));
    }
    PBX_HIP(hipEventRecord(s->done_ev, s->stream));
    bool posted = false;
    for (long long t = 0; t < steps; ++t) {
        const long long j = j0 + t;
        if (graphed) {
            // the step's draws resident (uploads are queued on the stream
            // ahead of the replay), the replay, then the draws consumed
            const int bD = j % diff_freq == 0, ch = j % chol_freq == 0;
            if (bD) s->have_D = 1;
            if (ch) {
                if (!s->have_D) {
                    pbx_set_error(PBX_ERR_STATE, "Cholesky requested before D was built");
                    return PBX_ERR_STATE;
                }
                s->have_B = 1;
            }
            if (!s->have_B) {
                pbx_set_error(PBX_ERR_STATE, "propagation without a noise factor");
                return PBX_ERR_STATE;
            }
            TRY(noise_ensure(s, 1));
            TRY(step_graph_launch(s, dt, bD, ch));
            noise_consume(s, 1);
            s->last_attempts = 1;
        } else {
            TRY(queued_step(s, dt, j % diff_freq == 0, j % lub_freq == 0, j % chol_freq == 0, s->flux_log + t * nl));
        }
        if ((t + 1) % 32 == 0) {
            if (posted) {  // the copy queued 32 steps ago: read it
                PBX_HIP(hipEventSynchronize(s->done_ev));
                TRY(take_info(s));
            }
            TRY(post_info(s));
            PBX_HIP(hipEventRecord(s->done_ev, s->stream));
            posted = true;
        }
    }
    TRY(post_info(s));
    PBX_HIP(hipMemcpyAsync(s->pos_host, s->pos, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->stream));
    if (s->forces_on) PBX_HIP(hipMemcpyAsync(s->E_host, s->E_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    if (steps > 0) {  // the last step's counters, as pbx_sim_get_flux and snapshots read them
        PBX_HIP(hipMemcpyAsync(s->flux, s->flux_log + (steps - 1) * nl, sizeof(long long) * nl,
                               hipMemcpyDeviceToDevice, s->stream));
        PBX_HIP(hipMemcpyAsync(s->flux_host, s->flux_log + (steps - 1) * nl, sizeof(long long) * nl,
                               hipMemcpyDeviceToHost, s->stream));
    }
    if (flux_out && s->cfg.measure_flux)
        PBX_HIP(hipMemcpyAsync(flux_out, s->flux_log, sizeof(long long) * (size_t)need, hipMemcpyDeviceToHost,
                               s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    if (flux_out && !s->cfg.measure_flux) memset(flux_out, 0, sizeof(long long) * (size_t)need);
    s->last_attempts = 1;
    return take_info(s);
}

}  // namespace

// Box.propagate plus the read-back BD.py needs after every step, in the
// step's final synchronisation: positions (N x 3) and, when flux is
// measured, the per-label counters (either pointer may be NULL).
extern "C" int pbx_sim_propagate_out(pbx_sim* s, double dt, int build_Dff, int build_Dnf, int cholesky,
                                     double* xyz_out, long long* flux_out) {
    int rc = pbx_sim_propagate(s, dt, build_Dff, build_Dnf, cholesky);
    if (rc) return rc;
    if (xyz_out) memcpy(xyz_out, s->pos_host, sizeof(double) * s->n);
    if (flux_out && s->cfg.measure_flux) {
        int nl = s->cfg.n_labels > 0 ? s->cfg.n_labels : 1;
        memcpy(flux_out, s->flux_host, sizeof(long long) * nl);
    }
    return PBX_OK;
}

// The driver loop of BD.py:164-185 without file output, inside the library:
// steps j = j0 .. j0+steps-1, each Box.propagate with the cadence flags of
// BD.py:177.  flux_out[t * n_labels + l] = label l's flux of step j0+t (the
// value BD.py writes on .flx line j0+t+1); xyz_out receives the positions
// before every step j with j % xyz_every == 0 (BD.py:170-171), one N x 3
// frame each, in order.
extern "C" int pbx_sim_run(pbx_sim* s, double dt, long long j0, long long steps, int diff_freq, int lub_freq,
                           int chol_freq, long long* flux_out, double* xyz_out, int xyz_every) {
    pbx_clear_error();
    if (!s || steps < 0 || j0 < 0 || diff_freq < 1 || lub_freq < 1 || chol_freq < 1 || (xyz_out && xyz_every < 1)) {
        pbx_set_error(PBX_ERR_ARG, "run: bad arguments");
        return PBX_ERR_ARG;
    }
    const int nl = s->cfg.n_labels > 0 ? s->cfg.n_labels : 1;
    // no step needs a host decision (no overlap test, no divergence term, no
    // frames to hand back): queue them all
    if (!xyz_out && !s->cfg.check_overlaps && !(s->cfg.divergence_term && is_lub(s)))
        return run_queued(s, dt, j0, steps, diff_freq, lub_freq, chol_freq, flux_out);
    double* frame = xyz_out;
    for (long long t = 0; t < steps; ++t) {
        const long long j = j0 + t;
        if (xyz_out && j % xyz_every == 0) {
            if (t == 0)
                TRY(pbx_sim_get_positions(s, frame));
            else
                memcpy(frame, s->pos_host, sizeof(double) * s->n);  // positions after step j-1
            frame += s->n;
        }
        s->next_builds = (t + 1 < steps && (j + 1) % diff_freq == 0) ? 1 : 0;
        const int rc = pbx_sim_propagate(s, dt, j % diff_freq == 0, j % lub_freq == 0, j % chol_freq == 0);
        s->next_builds = -1;
        if (rc) return rc;
        if (flux_out) {
            if (s->cfg.measure_flux)
                memcpy(flux_out + t * nl, s->flux_host, sizeof(long long) * nl);
            else
                memset(flux_out + t * nl, 0, sizeof(long long) * nl);
        }
    }
    return PBX_OK;
}

// box.py:404-461, 685-723 and interactions.py:483-589: the force field the
// step recomputes from the positions at the start of every propagate.
extern "C" int pbx_sim_set_forces(pbx_sim* s, const pbx_forces* f) {
    pbx_clear_error();
    if (!s || !f || !(f->unit > 0.0) || f->n_bonds < 0 || f->n_angles < 0 ||
        (f->n_bonds && (!f->bond_beads || !f->bond_params)) || (f->n_angles && (!f->angle_beads || !f->angle_params))) {
        pbx_set_error(PBX_ERR_ARG, "set_forces: bad arguments");
        return PBX_ERR_ARG;
    }
    if (s->forces_set) {
        pbx_set_error(PBX_ERR_STATE, "set_forces: the force field is set once per simulation");
        return PBX_ERR_STATE;
    }
    const int N = s->N, Nt = s->Nt;
    for (int q = 0; q < 2 * f->n_bonds; ++q)
        if (f->bond_beads[q] < 0 || f->bond_beads[q] >= Nt) {
            pbx_set_error(PBX_ERR_ARG, "set_forces: bond bead index out of range");
            return PBX_ERR_ARG;
        }
    for (int q = 0; q < 3 * f->n_angles; ++q)
        if (f->angle_beads[q] < 0 || f->angle_beads[q] >= Nt) {
            pbx_set_error(PBX_ERR_ARG, "set_forces: angle bead index out of range");
            return PBX_ERR_ARG;
        }
    PbxForceArgs& a = s->fa;
    a.N = N;
    a.Nt = Nt;
    a.L = s->cfg.box_length;
    a.unit = f->unit;
    a.lj = (f->lennard_jones_6 && f->lennard_jones_12) ? 3 : f->lennard_jones_6 ? 1 : f->lennard_jones_12 ? 2 : 0;
    a.alpha = f->lennard_jones_alpha;
    int region = 0;
    for (int c = 0; c < 3; ++c) {
        a.fex[c] = s->cfg.external_force[c];
        a.reg_on[c] = f->region_on[c] ? 1 : 0;
        a.reg_lo[c] = f->region[2 * c];
        a.reg_hi[c] = f->region[2 * c + 1];
        region |= a.reg_on[c];
    }
    s->forces_set = 1;
    // captured steps were recorded without the force launches
    PBX_HIP(hipStreamSynchronize(s->stream));
    drop_graphs(s);
    s->forces_on = a.lj || f->n_bonds || f->n_angles || region;
    if (!s->forces_on) {
        // constant F = unit * Fex on every bead (box.py:459-460, 414-428)
        std::vector<double> F(s->np, 0.0);
        s->F_nonzero = 0;
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) {
                F[3 * i + c] = f->unit == 1.0 ? a.fex[c] : f->unit * a.fex[c];
                if (F[3 * i + c] != 0.0) s->F_nonzero = 1;
            }
        PBX_HIP(sim_copy(s, s->F, F.data(), sizeof(double) * s->np, hipMemcpyHostToDevice));
        return PBX_OK;
    }
    s->F_nonzero = 1;
    auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
        TRY(sim_alloc(s, dst, bytes ? bytes : 8));
        if (bytes) PBX_HIP(sim_copy(s, *dst, src, bytes, hipMemcpyHostToDevice));
        return PBX_OK;
    };
    if (a.lj) {
        std::vector<double> hc(Nt), ep(Nt);
        for (int i = 0; i < Nt; ++i) {
            hc[i] = f->hard_core_radius ? f->hard_core_radius[i] : 0.0;
            ep[i] = f->epsilon_LJ ? f->epsilon_LJ[i] : 0.0;
        }
        TRY(upload((void**)&a.hc, hc.data(), sizeof(double) * Nt));
        TRY(upload((void**)&a.eps, ep.data(), sizeof(double) * Nt));
    }
    // per-bead incidence lists of the mobile members (counting sort by bead,
    // entries in list order); immobile beads get no force
    std::vector<int> bptr(N + 1, 0), aptr(N + 1, 0);
    for (int q = 0; q < f->n_bonds; ++q)
        for (int r = 0; r < 2; ++r)
            if (f->bond_beads[2 * q + r] < N) ++bptr[f->bond_beads[2 * q + r] + 1];
    for (int q = 0; q < f->n_angles; ++q)
        for (int r = 0; r < 3; ++r)
            if (f->angle_beads[3 * q + r] < N) ++aptr[f->angle_beads[3 * q + r] + 1];
    for (int i = 0; i < N; ++i) {
        bptr[i + 1] += bptr[i];
        aptr[i + 1] += aptr[i];
    }
    std::vector<int> bent(bptr[N]), aent(aptr[N]), bfill(bptr.begin(), bptr.end() - 1),
        afill(aptr.begin(), aptr.end() - 1);
    for (int q = 0; q < f->n_bonds; ++q)
        for (int r = 0; r < 2; ++r)
            if (f->bond_beads[2 * q + r] < N) bent[bfill[f->bond_beads[2 * q + r]]++] = 2 * q + r;
    for (int q = 0; q < f->n_angles; ++q)
        for (int r = 0; r < 3; ++r)
            if (f->angle_beads[3 * q + r] < N) aent[afill[f->angle_beads[3 * q + r]]++] = 4 * q + r;
    TRY(upload((void**)&a.bond_ptr, bptr.data(), sizeof(int) * (N + 1)));
    TRY(upload((void**)&a.bond_ent, bent.data(), sizeof(int) * bent.size()));
    TRY(upload((void**)&a.bond_ij, f->bond_beads, sizeof(int) * 2 * f->n_bonds));
    TRY(upload((void**)&a.bond_par, f->bond_params, sizeof(double) * 2 * f->n_bonds));
    TRY(upload((void**)&a.ang_ptr, aptr.data(), sizeof(int) * (N + 1)));
    TRY(upload((void**)&a.ang_ent, aent.data(), sizeof(int) * aent.size()));
    TRY(upload((void**)&a.ang_ijk, f->angle_beads, sizeof(int) * 3 * f->n_angles));
    TRY(upload((void**)&a.ang_par, f->angle_params, sizeof(double) * 2 * f->n_angles));
    TRY(sim_alloc(s, (void**)&s->e_part, sizeof(double) * N));
    TRY(sim_alloc(s, (void**)&s->E_dev, sizeof(double)));
    if (!s->E_host) {
        PBX_HIP(hipHostMalloc((void**)&s->E_host, sizeof(double)));
        *s->E_host = 0.0;
        PBX_HIP(hipHostGetDevicePointer((void**)&s->E_hd, s->E_host, 0));
    }
    return PBX_OK;
}

// The force vector of the last step (3N, box.py:404-428).
extern "C" int pbx_sim_get_force(pbx_sim* s, double* F) {
    if (!s || !F) return PBX_ERR_ARG;
    PBX_HIP(hipMemcpyAsync(F, s->F, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    return PBX_OK;
}

// The energy of the last step's force field (box.py:406-412; 0 without
// interactions), in the input's energy unit.
extern "C" int pbx_sim_get_energy(pbx_sim* s, double* E) {
    if (!s || !E) return PBX_ERR_ARG;
    *E = (s->forces_on && s->E_host) ? *s->E_host : 0.0;
    return PBX_OK;
}

// ||D - B B^T||_F / ||D||_F of the current diffusion matrix and noise factor.
extern "C" int pbx_sim_factor_residual(pbx_sim* s, double* rel) {
    pbx_clear_error();
    if (!s || !rel || s->cfg.hydrodynamics == PBX_NOHI || !s->have_D || !s->have_B) {
        pbx_set_error(PBX_ERR_STATE, "factor_residual: needs a dense D and its factor");
        return PBX_ERR_STATE;
    }
    TRY(ensure_D(s));
    DevBuf acc;
    TRY(acc.alloc(2 * sizeof(double)));
    return pbx_factor_residual(s->D, s->B, s->n, s->np, s->ld, acc.as<double>(), rel, s->stream);
}

extern "C" long long pbx_sim_device_bytes(pbx_sim* s) { return s ? (long long)s->device_bytes : 0; }

// Stage timing with HIP events recorded on the simulation's stream:
// [0] mobility assembly (+ lubrication and inverses), [1] Cholesky,
// [2] propagation (noise products, overlap checks, flux, wrap).
extern "C" int pbx_sim_set_timing(pbx_sim* s, int on) {
    // timing only: no system-scope fence (no L2 writeback and invalidation
    // on the stream at each mark); the host never waits on these events
    if (on && !s->ev[0])
        for (int k = 0; k < 8; ++k) PBX_HIP(hipEventCreateWithFlags(&s->ev[k], hipEventDisableSystemFence));
    if (on < 0 || on > 2) return PBX_ERR_ARG;
    s->timing = on;
    s->ahead_pending = -1;
    s->stage_ms[0] = s->stage_ms[1] = s->stage_ms[2] = 0.0;
    s->timed_steps = 0;
    return PBX_OK;
}

extern "C" long long pbx_sim_get_timing(pbx_sim* s, double* stage_ms) {
    for (int k = 0; k < 3; ++k) stage_ms[k] = s->stage_ms[k];
    return s->timed_steps;
}

extern "C" int pbx_sim_get_positions(pbx_sim* s, double* xyz) {
    PBX_HIP(hipMemcpyAsync(xyz, s->pos, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    return PBX_OK;
}

extern "C" int pbx_sim_set_positions(pbx_sim* s, const double* xyz) {
    s->spec_ok = 0;
    PBX_HIP(hipMemcpyAsync(s->pos, xyz, sizeof(double) * s->n, hipMemcpyHostToDevice, s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    return PBX_OK;
}

extern "C" int pbx_sim_get_flux(pbx_sim* s, long long* out) {
    int nl = s->cfg.n_labels > 0 ? s->cfg.n_labels : 1;
    PBX_HIP(hipMemcpyAsync(out, s->flux, sizeof(long long) * nl, hipMemcpyDeviceToHost, s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    return PBX_OK;
}

// which: 0 = D, 1 = B (upper triangle zeroed), 2 = Rff
extern "C" int pbx_sim_get_matrix(pbx_sim* s, int which, double* out) {
    const double* src = nullptr;
    if (s->cfg.hydrodynamics == PBX_NOHI) {
        pbx_set_error(PBX_ERR_ARG, "nohi keeps D and B as vectors (pbx_sim_get_diag)");
        return PBX_ERR_ARG;
    }
    if (which == 0) {
        TRY(ensure_D(s));
        src = s->D;
    }
    if (which == 1) {
        TRY(ensure_B(s));
        src = s->B;
    }
    if (which == 2) src = s->Rff;
    if (!src) {
        pbx_set_error(PBX_ERR_ARG, "matrix %d not held by this simulation", which);
        return PBX_ERR_ARG;
    }
    PBX_HIP(hipMemcpy2DAsync(out, sizeof(double) * s->n, src, sizeof(double) * s->ld, sizeof(double) * s->n, s->n,
                             hipMemcpyDeviceToHost, s->stream));
    PBX_HIP(hipStreamSynchronize(s->stream));
    if (which == 1)
        for (int r = 0; r < s->n; ++r)
            for (int c = r + 1; c < s->n; ++c) out[(size_t)r * s->n + c] = 0.0;
    return PBX_OK;
}

extern "C" int pbx_sim_get_diag(pbx_sim* s, int which, double* out) {
    const double* src = which == 0 ? s->Dd : s->Bd;
    if (!src) {
        pbx_set_error(PBX_ERR_ARG, "diagonal D/B exist only for hydrodynamics nohi");
        return PBX_ERR_ARG;
    }
    PBX_HIP(sim_copy(s, out, src, sizeof(double) * s->n, hipMemcpyDeviceToHost));
    return PBX_OK;
}

extern "C" long long pbx_sim_draw_count(pbx_sim* s) { return s->draw_count; }
extern "C" int pbx_sim_last_attempts(pbx_sim* s) { return s->last_attempts; }

extern "C" int pbx_sim_sync(pbx_sim* s) {
    PBX_HIP(hipStreamSynchronize(s->stream));
    return PBX_OK;
}

// Look-ahead assembly of D after each step that built it (on by default;
// non-lubrication matrix modes, HIP graphs off).
extern "C" int pbx_sim_set_lookahead(pbx_sim* s, int on) {
    s->lookahead = on != 0;
    s->spec_ok = 0;
    return PBX_OK;
}

// Lower Cholesky of a device-resident matrix with no allocation: the entry
// point for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, hosts that forbid it).  It runs the same
// two-column factorisation as pbx_cholesky, whose work area (diagonal-tile
// inverses, look-ahead words, one scratch tile) and status word live here in
// a file-scope device array instead of an allocation.
//
// Each GPU holds its own copy of `scratch` (6.3 MB) from the moment the
// library's code object is loaded on it until the process ends.  Host code
// never passes the array's name to a kernel (in host code it is a host-side
// stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device, on the first call with that device
// current, and is kept per device.  Every call on a GPU uses that GPU's one
// array, so calls that overlap on one GPU race; the words that must start at
// zero are cleared by the first kernel of every call.
#include <mutex>

#include "pbx_common.h"

namespace {
constexpr int kMaxOrder = 12288;  // largest n (and padded order) accepted
constexpr size_t kWorkDoubles = (size_t)kMaxOrder * 64 + PBX_LA_INTS / 2 + 64 * 64;
constexpr size_t kStatusDoubles = 32;  // status word, on a 256-byte block of its own
}  // namespace

static __device__ __align__(256) double scratch[kWorkDoubles + kStatusDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// The current device's address of `scratch` (looked up once per device).
int scratch_ptr(double** out) {
    int dev = 0;
    PBX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        pbx_set_error(PBX_ERR_ARG, "cholesky_noalloc: device %d out of range", dev);
        return PBX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            pbx_set_error(PBX_ERR_HIP, "cholesky_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return PBX_ERR_HIP;
        }
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return PBX_OK;
}

// B (order np, leading dimension ldB) := the lower tiles of D (n x n,
// leading dimension ldD) with the identity in rows and columns n..np-1;
// diagonal tiles whole.  Block (0, 0) also clears the look-ahead words and
// the status word.  Grid (T, T), tiles above the diagonal leave at once.
__global__ __launch_bounds__(256) void noalloc_stage_kernel(const double* __restrict__ D, int n, int ldD, double* B,
                                                            int np, int ldB, int* la, int* status) {
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj > ti) return;
    if (ti == 0 && tj == 0) {
        if (threadIdx.x < PBX_LA_INTS) la[threadIdx.x] = 0;
        if (threadIdx.x == 0) *status = 0;
    }
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = ti * 64 + (e >> 6), c = tj * 64 + (e & 63);
        const double v = (r < n && c < n) ? D[(size_t)r * ldD + c] : (r == c ? 1.0 : 0.0);
        B[(size_t)r * ldB + c] = v;
    }
}

// Zero the strict upper triangle of B (order np), as numpy.linalg.cholesky
// returns it.  Grid (T, T), tiles below the diagonal leave at once.
__global__ __launch_bounds__(256) void noalloc_zero_upper_kernel(double* B, int ldB) {
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj < ti) return;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = ti * 64 + (e >> 6), c = tj * 64 + (e & 63);
        if (c > r) B[(size_t)r * ldB + c] = 0.0;
    }
}

}  // namespace

// See include/pybrown_amd.h.
extern "C" int pbx_cholesky_noalloc(const double* D_dev, int n, int ld_D, double* B_dev, int ld_B, void* stream) {
    pbx_clear_error();
    if (n <= 0 || n > kMaxOrder) {
        pbx_set_error(PBX_ERR_ARG, "cholesky_noalloc: order %d outside 1..%d (the size of its static work area)", n,
                      kMaxOrder);
        return PBX_ERR_ARG;
    }
    const int np = (int)pbx_round_up(n, 64), T = np / 64;
    if (!D_dev || !B_dev || ld_D < n || ld_B < np || (ld_B & 1) || ((uintptr_t)B_dev & 15)) {
        pbx_set_error(PBX_ERR_ARG,
                      "cholesky_noalloc: need D (n x n, ld_D >= n) and B (%d rows, even ld_B >= %d, 16-byte aligned)",
                      np, np);
        return PBX_ERR_ARG;
    }
    if (pbx_dense_work_doubles(np) > kWorkDoubles) {
        pbx_set_error(PBX_ERR_STATE, "cholesky_noalloc: work area of order %d exceeds the static array", np);
        return PBX_ERR_STATE;
    }
    double* work = nullptr;
    int rc = scratch_ptr(&work);
    if (rc) return rc;
    int* la = (int*)(work + (size_t)np * 64);
    int* status = (int*)(work + kWorkDoubles);
    hipStream_t s = (hipStream_t)stream;
    PBX_LAUNCH(noalloc_stage_kernel, dim3(T, T), dim3(256), 0, s, D_dev, n, ld_D, B_dev, np, ld_B, la, status);
    PBX_HIP(hipGetLastError());
    rc = pbx_potrf_two_col_ready(B_dev, np, ld_B, work, status, s);
    if (rc) return rc;
    PBX_LAUNCH(noalloc_zero_upper_kernel, dim3(T, T), dim3(256), 0, s, B_dev, ld_B);
    PBX_HIP(hipGetLastError());
    int info = 0;
    PBX_HIP(hipMemcpyAsync(&info, status, sizeof(int), hipMemcpyDeviceToHost, s));
    PBX_HIP(hipStreamSynchronize(s));
    if (info == PBX_INFO_DAG_WAIT) {
        pbx_set_error(PBX_ERR_STATE, "a bounded dependency wait ran out (factorisation or solve)");
        return PBX_ERR_STATE;
    }
    if (info) pbx_set_error(info, "Matrix is not positive definite (pivot %d)", info - 1);
    return info;
}

// Pair kernels on gfx950: minimum-image pointers, overlap scan, free-space RPY
// mobility and Smith-Ewald periodic RPY mobility.
//
// Work decomposition: the lower triangle of bead pairs is cut into 16 x 16
// bead tiles (256 pairs = one 256-thread workgroup, one pair per lane).  Each
// lane evaluates one 3x3 block; the 48 x 48 tile is staged in LDS and leaves
// as coalesced 48-double row segments, once as computed and once transposed,
// so the full symmetric matrix is written without strided scatters.
#include "pbx_common.h"

#include <math.h>

namespace {

constexpr int TB = 16;          // beads per tile side
constexpr int TE = 3 * TB;      // matrix entries per tile side
constexpr int TLD = TE + 1;     // LDS row stride (odd, avoids bank aliasing)

__device__ __forceinline__ void tile_decode(int k, int& bi, int& bj) {
    // k enumerates lower-triangle tiles row by row: (0,0),(1,0),(1,1),(2,0)...
    int r = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= k) ++r;
    while (r * (r + 1) / 2 > k) --r;
    bi = r;
    bj = k - r * (r + 1) / 2;
}

// ---------------------------------------------------------------------------
// free-space RPY block, diff_tensor.c:490-549.  FP contraction is disabled so
// the arithmetic rounds exactly as the reference's C does.
__device__ void rpy_pair(double ai, double aj, double rx, double ry, double rz, double m[6]) {
#pragma clang fp contract(off)
    const double PI = 3.14159265358979323846;
    double big = ai > aj ? ai : aj;
    double small = ai > aj ? aj : ai;
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double s2 = ai * ai + aj * aj;
    if (r > (ai + aj)) {
        double c1 = 1.0 / (8 * PI * r);
        double c2 = 1.0 + s2 / (3 * r2);
        double c3 = (1.0 - s2 / r2) / r2;
        m[0] = c1 * (c2 + c3 * rx * rx);
        m[1] = c1 * (c2 + c3 * ry * ry);
        m[2] = c1 * (c2 + c3 * rz * rz);
        m[3] = c1 * c3 * rx * ry;
        m[4] = c1 * c3 * rx * rz;
        m[5] = c1 * c3 * ry * rz;
    } else if (r <= (big - small)) {
        double t = 1.0 / (6 * PI * big);
        m[0] = m[1] = m[2] = t;
        m[3] = m[4] = m[5] = 0.0;
    } else {
        double r3 = r2 * r;
        double c1 = 1.0 / (6 * PI * ai * aj);
        double c2 = 16 * r3 * (ai + aj);
        double c3 = (ai - aj) * (ai - aj) + 3 * r2;
        c3 *= c3;
        double c4 = (c2 - c3) / 32 / r3;
        double c5 = ((ai - aj) * (ai - aj) - r2);
        c5 *= 3 * c5;
        double c6 = c5 / (32 * r3) / r2;
        m[0] = c1 * (c4 + c6 * rx * rx);
        m[1] = c1 * (c4 + c6 * ry * ry);
        m[2] = c1 * (c4 + c6 * rz * rz);
        m[3] = c1 * c6 * rx * ry;
        m[4] = c1 * c6 * rx * rz;
        m[5] = c1 * c6 * ry * rz;
    }
}

// Free-space far-field form (the r > ai+aj branch without the branch test),
// used to remove the Ewald sum's point-particle value for overlapping pairs
// (diff_tensor.c:626-641).
__device__ __forceinline__ void rpy_far(double ai, double aj, double rx, double ry, double rz, double m[6]) {
    const double PI = 3.14159265358979323846;
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double s2 = ai * ai + aj * aj;
    double c1 = 1.0 / (8 * PI * r);
    double c2 = 1.0 + s2 / (3 * r2);
    double c3 = (1.0 - s2 / r2) / r2;
    m[0] = c1 * (c2 + c3 * rx * rx);
    m[1] = c1 * (c2 + c3 * ry * ry);
    m[2] = c1 * (c2 + c3 * rz * rz);
    m[3] = c1 * c3 * rx * ry;
    m[4] = c1 * c3 * rx * rz;
    m[5] = c1 * c3 * ry * rz;
}

// ---------------------------------------------------------------------------
// Smith-Ewald sums (Appendix of the reference: diff_tensor.c:553-1087),
// reorganised for the GPU: one pass over the real-space images evaluates
// sqrt/erfc/exp once per image for both the O and Q tensors, the reciprocal
// sum runs over half of the (n, -n) pairs with doubled weight (cos is even),
// and the |n|^2-dependent reciprocal coefficients come from a per-workgroup
// LDS table.  Accumulators: a scalar on the identity and a symmetric tensor
// on y y^T / n n^T, weighted by c1*c2 (O) and c1*c3 (Q).

constexpr int RECIP_TAB = 129;  // supports n_max <= 8 (|n|^2 <= 64) -> 3 doubles each

struct RecipCoef {
    double A;   // (2/pi)/n^2 exp(-pi^2 n^2/alpha^2)
    double AB;  // A * (1 + pi^2 n^2/alpha^2)/n^2
    double C;   // 4 pi exp(-pi^2 n^2/alpha^2)/n^2
};

__device__ void recip_table_fill(RecipCoef* tab, int n_max, double alpha) {
    const double PI = 3.14159265358979323846;
    int cnt = n_max * n_max + 1;
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        RecipCoef c = {0.0, 0.0, 0.0};
        if (q > 0) {
            double n2 = (double)q;
            double kq = PI * PI / (alpha * alpha) * n2;
            double e = exp(-kq);
            c.A = 2.0 / PI / n2 * e;
            c.AB = c.A * (1.0 + kq) / n2;
            c.C = 4.0 * PI * e / n2;
        }
        tab[q] = c;
    }
}

// Accumulate the Ewald lattice sums for separation sigma (box units).
// self: sigma = 0 and the m = 0 image is excluded.  MM / NM >= 0 fix the
// truncation at compile time (loops unroll; the reciprocal cosines come from
// three sincospi calls and angle-addition recurrences); -1 reads m_max /
// n_max at run time and evaluates each cosine directly.
template <int MM, int NM>
__device__ void smith_sums(double sx, double sy, double sz, bool self, double alpha, int m_max, int n_max,
                           double wO, double wQ, const RecipCoef* tab, double& sI, double T[6]) {
    const double SQRT_PI = 1.7724538509055160273;
    const double k0 = 2.0 * alpha / SQRT_PI;
    const double k1 = 2.0 / SQRT_PI;
    const double k2 = 4.0 * alpha * alpha * alpha / SQRT_PI;
    const double a2 = alpha * alpha;
    const int mm = MM >= 0 ? MM : m_max;
    const int nm = NM >= 0 ? NM : n_max;
#pragma unroll
    for (int mx = -mm; mx <= mm; ++mx) {
        const int rx = mm - (mx < 0 ? -mx : mx);
#pragma unroll
        for (int my = -rx; my <= rx; ++my) {
            const int ry = rx - (my < 0 ? -my : my);
#pragma unroll
            for (int mz = -ry; mz <= ry; ++mz) {
                if (self && mx == 0 && my == 0 && mz == 0) continue;
                double x = mx + sx, y = my + sy, z = mz + sz;
                double r2 = x * x + y * y + z * z;
                double r = sqrt(r2);
                double ir = 1.0 / r;
                double ir2 = ir * ir;
                double ir3 = ir2 * ir;
                double E = erfc(alpha * r);
                double G = exp(-a2 * r2);
                double H = E + k1 * alpha * r * G;
                double cI = wO * E * ir + wQ * H * ir3;
                double cY = wO * (E * ir3 + k0 * G * ir2) - wQ * (3.0 * H * ir3 * ir2 + k2 * G * ir2);
                sI += cI;
                T[0] += cY * x * x;
                T[1] += cY * y * y;
                T[2] += cY * z * z;
                T[3] += cY * x * y;
                T[4] += cY * x * z;
                T[5] += cY * y * z;
            }
        }
    }
    // reciprocal sum over half of the (n, -n) pairs, doubled (cos is even)
    if constexpr (NM >= 0) {
        double cx[NM + 1], sxn[NM + 1], cy[NM + 1], syn[NM + 1], cz[NM + 1], szn[NM + 1];
        cx[0] = cy[0] = cz[0] = 1.0;
        sxn[0] = syn[0] = szn[0] = 0.0;
        if (NM >= 1) {
            if (self) {
                cx[1] = cy[1] = cz[1] = 1.0;
                sxn[1] = syn[1] = szn[1] = 0.0;
            } else {
                sincospi(2.0 * sx, &sxn[1], &cx[1]);
                sincospi(2.0 * sy, &syn[1], &cy[1]);
                sincospi(2.0 * sz, &szn[1], &cz[1]);
            }
        }
#pragma unroll
        for (int k = 2; k <= NM; ++k) {
            cx[k] = cx[k - 1] * cx[1] - sxn[k - 1] * sxn[1];
            sxn[k] = sxn[k - 1] * cx[1] + cx[k - 1] * sxn[1];
            cy[k] = cy[k - 1] * cy[1] - syn[k - 1] * syn[1];
            syn[k] = syn[k - 1] * cy[1] + cy[k - 1] * syn[1];
            cz[k] = cz[k - 1] * cz[1] - szn[k - 1] * szn[1];
            szn[k] = szn[k - 1] * cz[1] + cz[k - 1] * szn[1];
        }
#pragma unroll
        for (int nx = 0; nx <= NM; ++nx) {
            const int rx = NM - nx;
#pragma unroll
            for (int ny = (nx == 0 ? 0 : -rx); ny <= rx; ++ny) {
                const int ay = ny < 0 ? -ny : ny;
                const int ry = rx - ay;
                // e^{i(nx tx + ny ty)}
                const double ysn = ny < 0 ? -syn[ay] : syn[ay];
                const double pr = cx[nx] * cy[ay] - sxn[nx] * ysn;
                const double pi = sxn[nx] * cy[ay] + cx[nx] * ysn;
#pragma unroll
                for (int nz = ((nx == 0 && ny == 0) ? 1 : -ry); nz <= ry; ++nz) {
                    const int az = nz < 0 ? -nz : nz;
                    const double zsn = nz < 0 ? -szn[az] : szn[az];
                    const double cs = pr * cz[az] - pi * zsn;
                    const RecipCoef c = tab[nx * nx + ny * ny + nz * nz];
                    sI += 2.0 * wO * c.A * cs;
                    const double cT = 2.0 * cs * (wQ * c.C - wO * c.AB);
                    T[0] += cT * (double)(nx * nx);
                    T[1] += cT * (double)(ny * ny);
                    T[2] += cT * (double)(nz * nz);
                    T[3] += cT * (double)(nx * ny);
                    T[4] += cT * (double)(nx * nz);
                    T[5] += cT * (double)(ny * nz);
                }
            }
        }
    } else {
        for (int nx = 0; nx <= nm; ++nx) {
            int rx = nm - nx;
            for (int ny = (nx == 0 ? 0 : -rx); ny <= rx; ++ny) {
                int ry = rx - abs(ny);
                int z0 = (nx == 0 && ny == 0) ? 1 : -ry;
                for (int nz = z0; nz <= ry; ++nz) {
                    int q = nx * nx + ny * ny + nz * nz;
                    RecipCoef c = tab[q];
                    double cs = self ? 1.0 : cospi(2.0 * (nx * sx + ny * sy + nz * sz));
                    sI += 2.0 * wO * c.A * cs;
                    double cT = 2.0 * cs * (wQ * c.C - wO * c.AB);
                    T[0] += cT * (double)(nx * nx);
                    T[1] += cT * (double)(ny * ny);
                    T[2] += cT * (double)(nz * nz);
                    T[3] += cT * (double)(nx * ny);
                    T[4] += cT * (double)(nx * nz);
                    T[5] += cT * (double)(ny * nz);
                }
            }
        }
    }
}

// diff_tensor.c:553-575 with the diagonal-only copy of :194-199.
template <int MM, int NM>
__device__ void smith_self_block(double a, double L, double alpha, int m_max, int n_max, const RecipCoef* tab,
                                 double m[6]) {
    const double PI = 3.14159265358979323846;
    const double SQRT_PI = 1.7724538509055160273;
    double c1 = 1.0 / (6 * PI * a);
    double c2 = 3 * a / (4 * L);
    double c3 = a * a * a / (2 * L * L * L);
    double sI = 0.0, T[6] = {0, 0, 0, 0, 0, 0};
    smith_sums<MM, NM>(0.0, 0.0, 0.0, true, alpha, m_max, n_max, c1 * c2, c1 * c3, tab, sI, T);
    double corrO = 1.5 * alpha * a / (SQRT_PI * L);
    double corrQ = alpha * alpha * alpha * a * a * a / (L * L * L * 3.0 * SQRT_PI);
    double diag = c1 * (1.0 - c2 * corrO - c3 * corrQ) + sI;
    m[0] = diag + T[0];
    m[1] = diag + T[1];
    m[2] = diag + T[2];
    m[3] = m[4] = m[5] = 0.0;
}

// diff_tensor.c:579-648 RPY_Smith_Mij_block
template <int MM, int NM>
__device__ void smith_pair_block(double ai, double aj, double rx, double ry, double rz, double L, double alpha,
                                 int m_max, int n_max, const RecipCoef* tab, double m[6]) {
    const double PI = 3.14159265358979323846;
    double c1 = 1.0 / (6 * PI * ai);
    double c2 = 3 * ai / (4 * L);
    double c3 = (ai == aj) ? ai * ai * ai / (2 * L * L * L) : ai * (ai * ai + aj * aj) / (4 * L * L * L);
    double sI = 0.0, T[6] = {0, 0, 0, 0, 0, 0};
    smith_sums<MM, NM>(rx / L, ry / L, rz / L, false, alpha, m_max, n_max, c1 * c2, c1 * c3, tab, sI, T);
    m[0] = sI + T[0];
    m[1] = sI + T[1];
    m[2] = sI + T[2];
    m[3] = T[3];
    m[4] = T[4];
    m[5] = T[5];
    double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 < (ai + aj) * (ai + aj)) {
        double nearb[6], farb[6];
        rpy_pair(ai, aj, rx, ry, rz, nearb);
        rpy_far(ai, aj, rx, ry, rz, farb);
        for (int k = 0; k < 6; ++k) m[k] += nearb[k] - farb[k];
    }
}

// ---------------------------------------------------------------------------
// geometry sources: either positions + box (min image) or the reference's
// packed pair list.

struct GeomXYZ {
    const double* xyz;
    double L;
    int N;
    // optional guard: the whole launch is skipped when *sel >= selK (a batch
    // of move attempts that accepted no move leaves the positions unchanged)
    const int* sel = nullptr;
    int selK = 0;
    __device__ bool skipped() const { return sel && *sel >= selK; }
    // separation vector the reference hands to its pair routines for i > j:
    // -wrap(r_i - r_j) (bead.c:42-89 then diffusion.py:40)
    __device__ void sep(int i, int j, double& x, double& y, double& z) const {
        x = -pbx_min_image(xyz[3 * i + 0] - xyz[3 * j + 0], L);
        y = -pbx_min_image(xyz[3 * i + 1] - xyz[3 * j + 1], L);
        z = -pbx_min_image(xyz[3 * i + 2] - xyz[3 * j + 2], L);
    }
};

struct GeomPacked {
    const double* P;
    int N;
    __device__ bool skipped() const { return false; }
    __device__ void sep(int i, int j, double& x, double& y, double& z) const {
        // pair (j < i) sits at results_position(i-1, j, N-1) (diff_tensor.c:107)
        size_t s = (size_t)j * (size_t)(N - 1) - (size_t)j * (size_t)(j + 1) / 2 + (size_t)(i - 1);
        x = P[3 * s + 0];
        y = P[3 * s + 1];
        z = P[3 * s + 2];
    }
};

enum Kind { K_RPY = 0, K_SMITH = 1 };

template <int KIND, class Geom, int MM = -1, int NM = -1>
__global__ __launch_bounds__(256) void assemble_tile_kernel(Geom g, const double* __restrict__ a, double L,
                                                            double alpha, int m_max, int n_max, double scale,
                                                            double* __restrict__ M, int ld) {
    __shared__ double tile[TE * TLD];
    __shared__ RecipCoef tab[RECIP_TAB];
    const int N = g.N;
    if (g.skipped()) return;
    int bi, bj;
    tile_decode(blockIdx.x, bi, bj);
    if (KIND == K_SMITH) recip_table_fill(tab, n_max, alpha);
    __syncthreads();

    const int ti = threadIdx.x / TB, tj = threadIdx.x % TB;
    const int i = bi * TB + ti, j = bj * TB + tj;
    double b[6] = {0, 0, 0, 0, 0, 0};
    bool live = (i < N) && (j < N) && (bi > bj || i >= j);
    if (live) {
        if (i == j) {
            if (KIND == K_RPY) {
                double d = 1.0 / (6 * 3.14159265358979323846 * a[i]);
                b[0] = b[1] = b[2] = d;
            } else {
                smith_self_block<MM, NM>(a[i], L, alpha, m_max, n_max, tab, b);
            }
        } else {
            double x, y, z;
            g.sep(i, j, x, y, z);
            if (KIND == K_RPY)
                rpy_pair(a[i], a[j], x, y, z, b);
            else
                smith_pair_block<MM, NM>(a[i], a[j], x, y, z, L, alpha, m_max, n_max, tab, b);
        }
        for (int k = 0; k < 6; ++k) b[k] = b[k] * scale;
    }
    // stage the 3x3 block; on diagonal tiles mirror it into the upper part
    const double full[9] = {b[0], b[3], b[4], b[3], b[1], b[5], b[4], b[5], b[2]};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            if (bi != bj || ti >= tj) tile[(3 * ti + r) * TLD + 3 * tj + c] = full[3 * r + c];
            if (bi == bj && ti > tj) tile[(3 * tj + c) * TLD + 3 * ti + r] = full[3 * r + c];
        }
    __syncthreads();

    const int n = 3 * N;
    const int R0 = bi * TE, C0 = bj * TE;
    for (int e = threadIdx.x; e < TE * TE; e += blockDim.x) {
        int r = e / TE, c = e % TE;
        if (R0 + r < n && C0 + c < n) M[(size_t)(R0 + r) * ld + C0 + c] = tile[r * TLD + c];
    }
    if (bi != bj) {
        for (int e = threadIdx.x; e < TE * TE; e += blockDim.x) {
            int r = e / TE, c = e % TE;  // r: row of the transposed tile
            if (C0 + r < n && R0 + c < n) M[(size_t)(C0 + r) * ld + R0 + c] = tile[c * TLD + r];
        }
    }
}

// ---------------------------------------------------------------------------
// clib/bead.c:27-95
__global__ void pointers_kernel(const double* __restrict__ xyz, int N, double L, double* __restrict__ P3) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)N * N;
    if (t >= total) return;
    int j = (int)(t / N), i = (int)(t % N);
    if (i <= j) return;
    for (int c = 0; c < 3; ++c) {
        double d = pbx_min_image(xyz[3 * i + c] - xyz[3 * j + c], L);
        P3[3 * ((size_t)j * N + i) + c] = d;
        P3[3 * ((size_t)i * N + j) + c] = -d;
    }
}

// clib/bead.c:99-158: P3[i][j] = wrap(r_immobile_j - r_mobile_i), layout
// (n_mobile, n_immobile, 3); thread per (i, j), j fastest (coalesced writes).
__global__ void immobile_pointers_kernel(const double* __restrict__ xm, const double* __restrict__ xi, int nm, int ni,
                                         double L, double* __restrict__ P3) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nm * ni) return;
    const int i = (int)(t / ni), j = (int)(t % ni);
    for (int c = 0; c < 3; ++c) P3[3 * t + c] = pbx_min_image(xi[3 * j + c] - xm[3 * i + c], L);
}

// clib/bead.c:169-260: overlap iff no raw axis difference lies in the open
// skip band (S, L-S) / (-(L-S), -S) and the minimum-image distance is at most
// S + threshold (S = a_i + a_j).  Beads Nm.. are immobile: their mutual pairs
// are not tested (the connection matrix of bead.py:401-413 exempts them).
__global__ __launch_bounds__(256) void overlaps_kernel(const double* __restrict__ xyz, const double* __restrict__ a,
                                                       int N, int Nm, double L, double thr,
                                                       const int* __restrict__ mask, int* flag) {
    int bi, bj;
    pbx_pair_tile(blockIdx.x, (Nm + TB - 1) / TB, bi, bj);
    const int ti = threadIdx.x / TB, tj = threadIdx.x % TB;
    const int i = bi * TB + ti, j = bj * TB + tj;
    bool hit = false;
    if (i < N && j < i && j < Nm && !(mask && mask[(size_t)j + (size_t)N * i]))
        hit = pbx_pair_overlaps(xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], xyz[3 * i + 2] - xyz[3 * j + 2],
                                a[i], a[j], L, thr);
    if (__any(hit) && (threadIdx.x % 64) == 0) atomicOr(flag, 1);
}

// Launch the Smith kernel with compile-time truncation for m_max, n_max <= 3.
template <class Geom>
void launch_smith(Geom g, const double* a, double L, double alpha, int m_max, int n_max, double scale, double* M,
                  int ld, int tiles, hipStream_t s) {
#define PBX_SMITH_CASE(MMV, NMV)                                                                              \
    if (m_max == MMV && n_max == NMV) {                                                                     \
        PBX_LAUNCH((assemble_tile_kernel<K_SMITH, Geom, MMV, NMV>), dim3(tiles), dim3(256), 0, s, g, a, \
                           L, alpha, m_max, n_max, scale, M, ld);                                           \
        return;                                                                                              \
    }
#define PBX_SMITH_ROW(MMV) PBX_SMITH_CASE(MMV, 0) PBX_SMITH_CASE(MMV, 1) PBX_SMITH_CASE(MMV, 2) PBX_SMITH_CASE(MMV, 3)
    PBX_SMITH_ROW(0)
    PBX_SMITH_ROW(1)
    PBX_SMITH_ROW(2)
    PBX_SMITH_ROW(3)
#undef PBX_SMITH_ROW
#undef PBX_SMITH_CASE
    PBX_LAUNCH((assemble_tile_kernel<K_SMITH, Geom, -1, -1>), dim3(tiles), dim3(256), 0, s, g, a, L, alpha,
                       m_max, n_max, scale, M, ld);
}

inline int tiles_for(int N) {
    int T = (N + TB - 1) / TB;
    return T * (T + 1) / 2;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers

int pbx_launch_rpy(const double* xyz, const double* a, int N, double L, double scale, double* M, int ld, int,
                   hipStream_t s, const int* sel, int selK) {
    if (N <= 0) return PBX_OK;
    PBX_CHECK_BOX(L);
    GeomXYZ g{xyz, L, N, sel, selK};
    PBX_LAUNCH((assemble_tile_kernel<K_RPY, GeomXYZ>), dim3(tiles_for(N)), dim3(256), 0, s, g, a, L, 0.0, 0,
                       0, scale, M, ld);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_smith(const double* xyz, const double* a, int N, double L, double alpha, int m_max, int n_max,
                     double scale, double* M, int ld, int, hipStream_t s, const int* sel, int selK) {
    if (N <= 0) return PBX_OK;
    PBX_CHECK_BOX(L);
    if (n_max < 0 || n_max > 8 || m_max < 0 || m_max > 32) {
        pbx_set_error(PBX_ERR_ARG, "Ewald truncation out of range (m_max=%d <= 32, n_max=%d <= 8)", m_max, n_max);
        return PBX_ERR_ARG;
    }
    GeomXYZ g{xyz, L, N, sel, selK};
    launch_smith(g, a, L, alpha, m_max, n_max, scale, M, ld, tiles_for(N), s);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_rpy_packed(const double* a, const double* P, int N, double* M, int ld, hipStream_t s) {
    if (N <= 0) return PBX_OK;
    GeomPacked g{P, N};
    PBX_LAUNCH((assemble_tile_kernel<K_RPY, GeomPacked>), dim3(tiles_for(N)), dim3(256), 0, s, g, a, 0.0,
                       0.0, 0, 0, 1.0, M, ld);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_smith_packed(const double* a, const double* P, int N, double L, double alpha, int m_max, int n_max,
                            double* M, int ld, hipStream_t s) {
    if (N <= 0) return PBX_OK;
    PBX_CHECK_BOX(L);
    if (n_max < 0 || n_max > 8 || m_max < 0 || m_max > 32) {
        pbx_set_error(PBX_ERR_ARG, "Ewald truncation out of range (m_max=%d <= 32, n_max=%d <= 8)", m_max, n_max);
        return PBX_ERR_ARG;
    }
    GeomPacked g{P, N};
    launch_smith(g, a, L, alpha, m_max, n_max, 1.0, M, ld, tiles_for(N), s);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_pointers(const double* xyz, int N, double L, double* P3, hipStream_t s) {
    if (N <= 1) return PBX_OK;
    PBX_CHECK_BOX(L);
    size_t total = (size_t)N * N;
    PBX_LAUNCH(pointers_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, xyz, N, L, P3);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_immobile_pointers(const double* xm, const double* xi, int nm, int ni, double L, double* P3,
                                 hipStream_t s) {
    if (nm <= 0 || ni <= 0) return PBX_OK;
    PBX_CHECK_BOX(L);
    size_t total = (size_t)nm * ni;
    PBX_LAUNCH(immobile_pointers_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, xm, xi, nm, ni,
                       L, P3);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_overlaps(const double* xyz, const double* a, int Nt, int Nm, double L, double thr, const int* mask,
                        int* flag, hipStream_t s) {
    if (Nt <= 1 || Nm <= 0) return PBX_OK;
    if (Nm > Nt) {
        pbx_set_error(PBX_ERR_ARG, "overlaps: %d mobile beads of %d", Nm, Nt);
        return PBX_ERR_ARG;
    }
    PBX_CHECK_BOX(L);
    PBX_LAUNCH(overlaps_kernel, dim3(pbx_pair_tiles(Nm, Nt)), dim3(256), 0, s, xyz, a, Nt, Nm, L, thr, mask,
                       flag);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// Jeffrey-Onishi lubrication correction on gfx950 (reference:
// src/bdsim/clib/diff_tensor.c:270-479, :1091-1625).
//
// For every pair with (r - a_i - a_j)/(a_i + a_j) <= cutoff the two-sphere
// resistance (JO series) minus the inverse of the two-sphere free-space RPY
// mobility, optionally Cichocki-projected, is added to a 3N x 3N matrix.
//
// All three 3x3 blocks of a pair are combinations of P = r r^T/r^2 and
// Q = I - P, so a pair is carried as six scalars (P and Q weights of blocks
// 00, 11 and 10).  The 6x6 RPY inverse (LAPACK LU in the reference) is the
// exact closed form in that basis: each of the parallel and perpendicular
// subspaces is a 2x2 system [[m_i, b],[b, m_j]].
//
// Decomposition: one wave per bead k scans all partners p; lanes take
// partners p = lane, lane+64, ...  The coupling block (k,p) is written by
// bead k's wave alone; the diagonal block (k,k) is the wave's fixed-order
// reduction of its lanes' sums.  No atomics, so the result is deterministic.
#include "pbx_common.h"

namespace {

// Jeffrey & Onishi (1984) f_m(lambda) coefficients (lambda^k at [m][k]),
// the table of diff_tensor.c:1385-1573.
__constant__ double c_XF[12][11] = {
    {1},
    {0, 3},
    {0, 9},
    {0, -4, 27, -4},
    {0, -24, 81, 36},
    {0, 0, 72, 243, 72},
    {0, 16, 108, 281, 648, 144},
    {0, 0, 288, 1620, 1515, 1620, 288},
    {0, 0, 576, 4848, 5409, 4524, 3888, 576},
    {0, 0, 1152, 9072, 14752, 26163, 14752, 9072, 1152},
    {0, 0, 2304, 20736, 42804, 115849, 76176, 39264, 20736, 2304},
    {0, 0, 4608, 46656, 108912, 269100, 319899, 269100, 108912, 46656, 4608},
};
__constant__ double c_YF[12][11] = {
    {1},
    {0, 1.5},
    {0, 2.25},
    {0, 2, 3.375, 2},
    {0, 6, 5.0625, 18},
    {0, 0, 31.5, 7.59375, 31.5},
    {0, 4, 54, 19.390625, 81, 72},
    {0, 0, 144, 131.625, 149.0859375, 131.625, 144},
    {0, 0, 279, 532.625, 493.62890625, -14.625, 648, 288},
    {0, 0, 576, 1134, 1888.84375, 1496.443359375, 1888.84375, 1134, 576},
    {0, 0, 1152, 1964.25, 6155.4375, 10301.1650390625, 8452.125, -175.5, 3888, 1152},
    {0, 0, 2304, 7128, 11035.5, 21441.4453125, 46486.24755859375, 21441.4453125, 11035.5, 7128, 2304},
};

__device__ __forceinline__ double fpoly(const double (*tab)[11], int m, double l) {
    double v = 0.0;
    for (int k = 10; k >= 0; --k) v = v * l + tab[m][k];
    return v;
}

struct GPoly {
    double xg1, xg2, xg3, yg2, yg3;
};

// diff_tensor.c:1577-1625
__device__ GPoly gpoly(double l) {
    double l2 = l * l, l3 = l2 * l, l4 = l2 * l2;
    double p3 = (1 + l) * (1 + l) * (1 + l);
    GPoly g;
    g.xg1 = 2 * l2 / p3;
    g.xg2 = 0.2 * l * (1 + 7 * l + l2) / p3;
    g.xg3 = (1.0 / 42) * (1 + 18 * l - 29 * l2 + 18 * l3 + l4) / p3;
    g.yg2 = (4.0 / 15) * l * (2 + l + 2 * l2) / p3;
    g.yg3 = (2.0 / 375) * (16 - 45 * l + 58 * l2 - 45 * l3 + 16 * l4) / p3;
    return g;
}

// X^A_11 and Y^A_11 (diff_tensor.c:1190-1285): even m = 2..10.
__device__ void jo_11(double s, double l, double& X, double& Y) {
    GPoly g = gpoly(l);
    double s2 = 1.0 - 4.0 / (s * s);
    double ls = log(s2);
    double x = g.xg1 / s2 - g.xg2 * ls - g.xg3 * s2 * ls + (1.0 - g.xg1);
    double y = -g.yg2 * ls - g.yg3 * s2 * ls + 1.0;
    double t = 2.0 / s, t2 = t * t, u = 1.0 / (1.0 + l), u2 = u * u;
    double tm = 1.0, um = 1.0;
    for (int m = 2; m < 12; m += 2) {
        tm *= t2;
        um *= u2;
        double m1 = (m == 2) ? -2.0 : (double)(m - 2);
        double hm = ldexp(um, -m);  // 2^-m (1+l)^-m
        x += tm * (hm * fpoly(c_XF, m, l) - g.xg1) + tm * (4.0 / (m * m1) * g.xg3 - 2.0 / m * g.xg2);
        y += tm * (hm * fpoly(c_YF, m, l) - 2.0 / m * g.yg2) + tm * 4.0 / (m * m1) * g.yg3;
    }
    X = x;
    Y = y;
}

// X^A_12 and Y^A_12 (diff_tensor.c:1289-1381): odd m = 1..11, sign minus.
__device__ void jo_12(double s, double l, double& X, double& Y) {
    GPoly g = gpoly(l);
    double s2 = 1.0 - 4.0 / (s * s);
    double ls = log((s + 2) / (s - 2));
    double x = 2.0 / s * g.xg1 / s2 + g.xg2 * ls + g.xg3 * s2 * ls + 4 * g.xg3 / s;
    double y = g.yg2 * ls + g.yg3 * s2 * ls + 4 * g.yg3 / s;
    double t = 2.0 / s, t2 = t * t, u = 1.0 / (1.0 + l), u2 = u * u;
    double tm = t, um = u;
    for (int m = 1; m < 12; m += 2) {
        if (m > 1) {
            tm *= t2;
            um *= u2;
        }
        double m1 = (double)(m - 2);
        double hm = ldexp(um, -m);
        x += tm * (hm * fpoly(c_XF, m, l) - g.xg1) + tm * (4.0 / (m * m1) * g.xg3 - 2.0 / m * g.xg2);
        y += tm * (hm * fpoly(c_YF, m, l) - 2.0 / m * g.yg2) + tm * (4.0 / (m * m1) * g.yg3);
    }
    X = -x;
    Y = -y;
}

// Pair result as (P weight, Q weight) for blocks 00, 11, 10.
struct PairPQ {
    double p[3], q[3];
};

// Two-sphere JO resistance for radii (ai, aj) (diff_tensor.c:270-342).
__device__ PairPQ jo_pair(double ai, double aj, double r) {
    const double PI = 3.14159265358979323846;
    double s = 2 * r / (ai + aj);
    double l = aj / ai;
    double x11, y11, x22, y22, x12, y12;
    jo_11(s, l, x11, y11);
    jo_11(s, 1 / l, x22, y22);
    x22 *= l;
    y22 *= l;
    jo_12(s, l, x12, y12);
    double sc = 6 * PI * ai;
    PairPQ o;
    o.p[0] = x11 * sc; o.q[0] = y11 * sc;
    o.p[1] = x22 * sc; o.q[1] = y22 * sc;
    o.p[2] = x12 * sc; o.q[2] = y12 * sc;
    return o;
}

// Inverse of the 6x6 two-sphere free-space RPY mobility (diff_tensor.c:1091-
// 1162) in closed form.
__device__ PairPQ rpy2b_inverse(double ai, double aj, double r) {
    const double PI = 3.14159265358979323846;
    double mi = 1.0 / (6 * PI * ai), mj = 1.0 / (6 * PI * aj);
    double big = ai > aj ? ai : aj, small = ai > aj ? aj : ai;
    double r2 = r * r;
    double s2 = ai * ai + aj * aj;
    double bpar, bperp;
    if (r > ai + aj) {
        double c1 = 1.0 / (8 * PI * r);
        double c2 = 1.0 + s2 / (3 * r2);
        double c3 = (1.0 - s2 / r2) / r2;
        bperp = c1 * c2;
        bpar = c1 * (c2 + c3 * r2);
    } else if (r <= big - small) {
        bpar = bperp = 1.0 / (6 * PI * big);
    } else {
        double r3 = r2 * r;
        double c1 = 1.0 / (6 * PI * ai * aj);
        double c4 = (16 * r3 * (ai + aj) - ((ai - aj) * (ai - aj) + 3 * r2) * ((ai - aj) * (ai - aj) + 3 * r2)) /
                    (32 * r3);
        double c5 = (ai - aj) * (ai - aj) - r2;
        double c6 = 3 * c5 * c5 / (32 * r3) / r2;
        bperp = c1 * c4;
        bpar = c1 * (c4 + c6 * r2);
    }
    double dpar = 1.0 / (mi * mj - bpar * bpar);
    double dperp = 1.0 / (mi * mj - bperp * bperp);
    PairPQ o;
    o.p[0] = mj * dpar;   o.q[0] = mj * dperp;
    o.p[1] = mi * dpar;   o.q[1] = mi * dperp;
    o.p[2] = -bpar * dpar; o.q[2] = -bperp * dperp;
    return o;
}

// diff_tensor.c:1166-1186 in P/Q weights
__device__ __forceinline__ PairPQ cichocki(PairPQ t) {
    PairPQ o;
    double dp = (t.p[0] + t.p[1] - 2 * t.p[2]) / 4;
    double dq = (t.q[0] + t.q[1] - 2 * t.q[2]) / 4;
    o.p[0] = o.p[1] = dp;
    o.q[0] = o.q[1] = dq;
    o.p[2] = -dp;
    o.q[2] = -dq;
    return o;
}

// ΔR for radii (a_lo, a_hi) and separation r (lo < hi in bead order).
__device__ PairPQ lub_pair(double alo, double ahi, double r, int cich) {
    PairPQ nf = jo_pair(alo, ahi, r);
    PairPQ ff = rpy2b_inverse(alo, ahi, r);
    if (cich) {
        nf = cichocki(nf);
        ff = cichocki(ff);
    }
    PairPQ d;
    for (int b = 0; b < 3; ++b) {
        d.p[b] = nf.p[b] - ff.p[b];
        d.q[b] = nf.q[b] - ff.q[b];
    }
    return d;
}

// Expand P/Q weights into (xx,yy,zz,xy,xz,yz).
__device__ __forceinline__ void expand(double p, double q, double rx, double ry, double rz, double r2, double o[6]) {
    double ex = rx * rx / r2, ey = ry * ry / r2, ez = rz * rz / r2;
    double pq = p - q;
    o[0] = p * ex + q * (1 - ex);
    o[1] = p * ey + q * (1 - ey);
    o[2] = p * ez + q * (1 - ez);
    o[3] = pq * rx * ry / r2;
    o[4] = pq * rx * rz / r2;
    o[5] = pq * ry * rz / r2;
}

struct GeomXYZ {
    const double* xyz;
    double L;
    __device__ void sep(int hi, int lo, double& x, double& y, double& z) const {
        x = -pbx_min_image(xyz[3 * hi + 0] - xyz[3 * lo + 0], L);
        y = -pbx_min_image(xyz[3 * hi + 1] - xyz[3 * lo + 1], L);
        z = -pbx_min_image(xyz[3 * hi + 2] - xyz[3 * lo + 2], L);
    }
};

struct GeomPacked {
    const double* P;
    int N;
    __device__ void sep(int hi, int lo, double& x, double& y, double& z) const {
        size_t s = (size_t)lo * (size_t)(N - 1) - (size_t)lo * (size_t)(lo + 1) / 2 + (size_t)(hi - 1);
        x = P[3 * s + 0];
        y = P[3 * s + 1];
        z = P[3 * s + 2];
    }
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// C[block(k, .)] = ((ΔR * mul) / div) + C for bead k = one wave.
template <class Geom>
__global__ __launch_bounds__(256) void lub_rows_kernel(Geom g, const double* __restrict__ a, int N, double cutoff,
                                                       int cich, double mul, double div, double* C, int ld) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (k >= N) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double ak = a[k];
    for (int p = lane; p < N; p += 64) {
        if (p == k) continue;
        int lo = p < k ? p : k, hi = p < k ? k : p;
        double rx, ry, rz;
        g.sep(hi, lo, rx, ry, rz);
        double r2 = rx * rx + ry * ry + rz * rz;
        double r = sqrt(r2);
        double alo = a[lo], ahi = a[hi];
        if (!((r - alo - ahi) / (alo + ahi) <= cutoff)) continue;
        PairPQ d = lub_pair(alo, ahi, r, cich);
        double blk[6];
        int own = (k == lo) ? 0 : 1;
        expand(d.p[own], d.q[own], rx, ry, rz, r2, blk);
        for (int e = 0; e < 6; ++e) acc[e] += blk[e];
        expand(d.p[2], d.q[2], rx, ry, rz, r2, blk);
        const double full[9] = {blk[0], blk[3], blk[4], blk[3], blk[1], blk[5], blk[4], blk[5], blk[2]};
        for (int rr = 0; rr < 3; ++rr)
            for (int cc = 0; cc < 3; ++cc) {
                double* dst = C + (size_t)(3 * k + rr) * ld + 3 * p + cc;
                *dst = ((full[3 * rr + cc] * mul) / div) + *dst;
            }
    }
    (void)ak;
    for (int e = 0; e < 6; ++e) acc[e] = wave_sum(acc[e]);
    if (lane == 0) {
        const double full[9] = {acc[0], acc[3], acc[4], acc[3], acc[1], acc[5], acc[4], acc[5], acc[2]};
        for (int rr = 0; rr < 3; ++rr)
            for (int cc = 0; cc < 3; ++cc) {
                double* dst = C + (size_t)(3 * k + rr) * ld + 3 * k + cc;
                *dst = ((full[3 * rr + cc] * mul) / div) + *dst;
            }
    }
}

// JO_2B_R_matrix for a single pair, packed 18 values (diff_tensor.c:270-342)
__global__ void jo_2b_kernel(double ai, double aj, double rx, double ry, double rz, double* out) {
    if (threadIdx.x != 0) return;
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    PairPQ d = jo_pair(ai, aj, r);
    for (int b = 0; b < 3; ++b) expand(d.p[b], d.q[b], rx, ry, rz, r2, out + 6 * b);
}

}  // namespace

int pbx_launch_jo_correction(const double* xyz, const double* P, const double* a, int N, double L, double cutoff,
                             int cich, double mul, double div, double* C, int ld, hipStream_t s) {
    if (N <= 1) return PBX_OK;
    if (xyz) PBX_CHECK_BOX(L);
    dim3 grid((N + 3) / 4), block(256);
    if (xyz) {
        GeomXYZ g{xyz, L};
        PBX_LAUNCH(lub_rows_kernel<GeomXYZ>, grid, block, 0, s, g, a, N, cutoff, cich, mul, div, C, ld);
    } else {
        GeomPacked g{P, N};
        PBX_LAUNCH(lub_rows_kernel<GeomPacked>, grid, block, 0, s, g, a, N, cutoff, cich, mul, div, C, ld);
    }
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

int pbx_launch_jo_2b(double ai, double aj, double rx, double ry, double rz, double* out18, hipStream_t s) {
    PBX_LAUNCH(jo_2b_kernel, dim3(1), dim3(64), 0, s, ai, aj, rx, ry, rz, out18);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

/*
 * oracle/orc.c -- CPU restatement of the reference's hydrodynamic kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the engine under pybrown_amd/ never calls into it.
 *
 * Every function restates one reference function.  Citations are relative to
 * the reference's src/bdsim/ directory.  Floating-point expressions keep the
 * reference's evaluation order (left-to-right products, the same temporaries)
 * so that, built with the same compiler flags and without FMA contraction, the
 * results agree with the reference bit for bit where libm is shared.
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   orc_pointer_pbc_matrix      bit-exact vs the compiled reference (oracle/_ref)
 *   orc_pointer_immobile_...    bit-exact vs the compiled reference (oracle/_ref)
 *   orc_check_overlaps          exact vs oracle/_ref on random configurations
 *   orc_rpy_m_matrix            bit-exact vs oracle/_ref; pecnut golden values
 *   orc_rpy_smith_m_matrix      bit-exact vs oracle/_ref; BD_BOX golden ratios
 *   orc_jo_2b_r_matrix          bit-exact vs oracle/_ref; HYDROLIB golden values
 *   orc_rpy_2b_r_matrix         6x6 inverse vs numpy.linalg.inv (the reference
 *                               uses LAPACK dgetrf/dgetri, absent from this
 *                               image, so its own build cannot run this one)
 *   orc_jo_r_lubrication_...    composition of the two above; HYDROLIB golden
 *                               values, Cichocki null-space and permutation
 *                               invariants
 *
 * Matrix layout: dense 3N x 3N, row-major (the reference reshapes its output
 * buffer with numpy in C order, pyBrown/diffusion.py:54).  The mobility is
 * symmetric, so the reference's column-major writes give the same matrix.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* index helpers                                                             */

/* Offset of pair (p<q) in the packed upper-triangle pointer list
 * (clib/diff_tensor.c:1645-1650 results_position(q-1, p, N-1)). */
static inline size_t pair_slot(int p, int q, int N)
{
	return (size_t)p * (size_t)(N - 1) - (size_t)p * (size_t)(p + 1) / 2 + (size_t)(q - 1);
}

/* Minimum-image reduction with the reference's loop semantics
 * (clib/bead.c:44-50): result in (-L/2, L/2]. */
static inline double min_image(double d, double L)
{
	while (d >= L / 2.0) d -= L;
	while (d <= -L / 2.0) d += L;
	return d;
}

/* Scatter a symmetric 3x3 block (xx,yy,zz,xy,xz,yz) at block (bi,bj) and its
 * transpose.  n = 3N is the matrix order. */
static void put_block(double *M, size_t n, int bi, int bj, const double b[6])
{
	size_t I = 3 * (size_t)bi, J = 3 * (size_t)bj;
	double full[3][3] = {{b[0], b[3], b[4]}, {b[3], b[1], b[5]}, {b[4], b[5], b[2]}};
	for (int r = 0; r < 3; ++r)
		for (int c = 0; c < 3; ++c) {
			M[(I + r) * n + J + c] = full[r][c];
			M[(J + c) * n + I + r] = full[r][c];
		}
}

static void add_block(double *M, size_t n, int bi, int bj, const double b[6])
{
	size_t I = 3 * (size_t)bi, J = 3 * (size_t)bj;
	double full[3][3] = {{b[0], b[3], b[4]}, {b[3], b[1], b[5]}, {b[4], b[5], b[2]}};
	if (bi == bj) {
		for (int r = 0; r < 3; ++r)
			for (int c = 0; c < 3; ++c) M[(I + r) * n + I + c] += full[r][c];
		return;
	}
	for (int r = 0; r < 3; ++r)
		for (int c = 0; c < 3; ++c) {
			M[(I + r) * n + J + c] += full[r][c];
			M[(J + c) * n + I + r] += full[r][c];
		}
}

/* ------------------------------------------------------------------------ */
/* geometry: clib/bead.c                                                     */

/* clib/bead.c:27-95 pointer_pbc_matrix: P3[a][b] = wrap(r_b - r_a) for b>a and
 * its negation for a>b.  The diagonal is left as the caller set it. */
void orc_pointer_pbc_matrix(const double *xyz, int N, double L, double *P3)
{
	for (int j = 0; j < N; ++j)
		for (int i = j + 1; i < N; ++i) {
			double d[3];
			for (int c = 0; c < 3; ++c) d[c] = min_image(xyz[3 * i + c] - xyz[3 * j + c], L);
			double *up = P3 + 3 * ((size_t)j * N + i);
			double *lo = P3 + 3 * ((size_t)i * N + j);
			for (int c = 0; c < 3; ++c) {
				up[c] = d[c];
				lo[c] = -d[c];
			}
		}
}

/* clib/bead.c:99-158 pointer_immobile_pbc_matrix: P3[i][j] = wrap(r_immobile_j -
 * r_mobile_i), layout (n_mobile, n_immobile, 3), same wrap loops. */
void orc_pointer_immobile_pbc_matrix(const double *xm, const double *xi, int nm, int ni, double L, double *P3)
{
	for (int j = 0; j < ni; ++j)
		for (int i = 0; i < nm; ++i)
			for (int c = 0; c < 3; ++c)
				P3[3 * j + 3 * (size_t)i * ni + c] = min_image(xi[3 * j + c] - xm[3 * i + c], L);
}

/* pyBrown/diffusion.py:40 -- the packed pair list handed to every mobility
 * routine: -rij[p][q] for p<q in row-major upper-triangle order. */
void orc_packed_pointers(const double *xyz, int N, double L, double *P)
{
	for (int p = 0; p < N; ++p)
		for (int q = p + 1; q < N; ++q) {
			double *out = P + 3 * pair_slot(p, q, N);
			for (int c = 0; c < 3; ++c) out[c] = -min_image(xyz[3 * q + c] - xyz[3 * p + c], L);
		}
}

/* clib/bead.c:169-218 check_overlaps with overlap_pbc/distance_pbc
 * (:222-260).  A pair is skipped when any raw axis difference lies strictly
 * inside (S, L-S) or (-(L-S), -S), S = a_i + a_j without the threshold; only
 * then is the minimum-image distance compared with S + threshold. */
int orc_check_overlaps(const double *xyz, const double *a, int N, double L, double thr, const int *mask)
{
	for (int j = 0; j < N; ++j)
		for (int i = j + 1; i < N; ++i) {
			if (mask && mask[(size_t)j + (size_t)N * i]) continue;
			double d[3];
			for (int c = 0; c < 3; ++c) d[c] = xyz[3 * i + c] - xyz[3 * j + c];
			double S = a[i] + a[j];
			double Sp = L - S;
			int skip = 0;
			for (int c = 0; c < 3 && !skip; ++c)
				if ((d[c] > S && d[c] < Sp) || (d[c] < -S && d[c] > -Sp)) skip = 1;
			if (skip) continue;
			for (int c = 0; c < 3; ++c) d[c] = min_image(d[c], L);
			double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
			if (dist <= a[i] + a[j] + thr) return 1;
		}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* free-space RPY: clib/diff_tensor.c:483-549                                */

/* Off-diagonal RPY block for radii (ai, aj) and separation r, written as
 * (xx,yy,zz,xy,xz,yz).  Three branches: disjoint, one sphere inside the
 * other, partial overlap (diff_tensor.c:490-549).  The contained branch
 * leaves the off-diagonal components zero. */
void orc_rpy_pair(double ai, double aj, double rx, double ry, double rz, double m[6])
{
	double big = ai > aj ? ai : aj;
	double small = ai > aj ? aj : ai;
	double r2 = rx * rx + ry * ry + rz * rz;
	double r = sqrt(r2);
	double s2 = ai * ai + aj * aj;

	if (r > (ai + aj)) {
		double c1 = 1.0 / (8 * M_PI * r);
		double c2 = 1.0 + s2 / (3 * r2);
		double c3 = (1.0 - s2 / r2) / r2;
		m[0] = c1 * (c2 + c3 * rx * rx);
		m[1] = c1 * (c2 + c3 * ry * ry);
		m[2] = c1 * (c2 + c3 * rz * rz);
		m[3] = c1 * c3 * rx * ry;
		m[4] = c1 * c3 * rx * rz;
		m[5] = c1 * c3 * ry * rz;
	} else if (r <= (big - small)) {
		double t = 1.0 / (6 * M_PI * big);
		m[0] = m[1] = m[2] = t;
		m[3] = m[4] = m[5] = 0.0;
	} else {
		double r3 = r2 * r;
		double c1 = 1.0 / (6 * M_PI * ai * aj);
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

/* clib/diff_tensor.c:73-166 RPY_M_matrix.  Diagonal blocks carry only their
 * diagonal 1/(6 pi a) (:129-135); the buffer must be zeroed by the caller. */
void orc_rpy_m_matrix(const double *a, const double *P, int N, double *M)
{
	size_t n = 3 * (size_t)N;
	for (int j = 0; j < N; ++j) {
		double d = 1.0 / (6 * M_PI * a[j]);
		for (int c = 0; c < 3; ++c) M[(3 * (size_t)j + c) * n + 3 * (size_t)j + c] = d;
		for (int i = j + 1; i < N; ++i) {
			const double *r = P + 3 * pair_slot(j, i, N);
			double b[6];
			orc_rpy_pair(a[i], a[j], r[0], r[1], r[2], b);
			put_block(M, n, i, j, b);
		}
	}
}

/* ------------------------------------------------------------------------ */
/* Smith-Ewald periodic RPY: clib/diff_tensor.c:553-1087                     */

/* O(x) = (I + x x^T/|x|^2)/|x| (diff_tensor.c:1043-1062) */
static void o_tensor(double x, double y, double z, double o[6])
{
	double r2 = x * x + y * y + z * z;
	double r = sqrt(r2);
	double r3 = r2 * r;
	o[0] = (1 + x * x / r2) / r;
	o[1] = (1 + y * y / r2) / r;
	o[2] = (1 + z * z / r2) / r;
	o[3] = x * y / r3;
	o[4] = x * z / r3;
	o[5] = y * z / r3;
}

/* Q(x) = (I - 3 x x^T/|x|^2)/|x|^3 (diff_tensor.c:1066-1087) */
static void q_tensor(double x, double y, double z, double q[6])
{
	double r2 = x * x + y * y + z * z;
	double r = sqrt(r2);
	double r3 = r2 * r;
	double r5 = r3 * r2;
	q[0] = (1 - 3 * x * x / r2) / r3;
	q[1] = (1 - 3 * y * y / r2) / r3;
	q[2] = (1 - 3 * z * z / r2) / r3;
	q[3] = -3 * x * y / r5;
	q[4] = -3 * x * z / r5;
	q[5] = -3 * y * z / r5;
}

static inline void acc6(double t[6], double s, const double v[6])
{
	for (int k = 0; k < 6; ++k) t[k] += s * v[k];
}

/* Outer product scaled: t += s * (x x^T) with the reference's (s*x)*y order. */
static inline void acc_outer(double t[6], double s, double x, double y, double z)
{
	t[0] += s * x * x;
	t[1] += s * y * y;
	t[2] += s * z * z;
	t[3] += s * x * y;
	t[4] += s * x * z;
	t[5] += s * y * z;
}

static inline void sub_outer(double t[6], double s, double x, double y, double z)
{
	t[0] -= s * x * x;
	t[1] -= s * y * y;
	t[2] -= s * z * z;
	t[3] -= s * x * y;
	t[4] -= s * x * z;
	t[5] -= s * y * z;
}

/* The truncation sets are L1 balls: |mx|+|my|+|mz| <= m_max
 * (diff_tensor.c:675-681). */
#define L1_BALL(mx, my, mz, lim)                                   \
	for (int mx = -(lim); mx <= (lim); ++mx)                       \
		for (int my = -((lim) - abs(mx)); my <= (lim) - abs(mx); ++my) \
			for (int mz = -((lim) - abs(mx) - abs(my)); mz <= (lim) - abs(mx) - abs(my); ++mz)

/* diff_tensor.c:652-744 RPY_Smith_Oii_block */
static void smith_o_self(double a, double L, double alpha, int mmax, int nmax, double o[6])
{
	double sqrt_pi = sqrt(M_PI);
	double alpha2 = alpha * alpha;
	double k0 = 2 * alpha / sqrt_pi;
	double kexp = -M_PI * M_PI / alpha2;
	double v[6];
	memset(o, 0, 6 * sizeof(double));
	L1_BALL(mx, my, mz, mmax)
	{
		if (mx == 0 && my == 0 && mz == 0) continue;
		double len2 = mx * mx + my * my + mz * mz;
		double len = sqrt(len2);
		o_tensor(mx, my, mz, v);
		acc6(o, erfc(alpha * len), v);
		double s = k0 * exp(-alpha2 * len2) / len2;
		acc_outer(o, s, mx, my, mz);
	}
	double k1 = 2.0 / M_PI;
	L1_BALL(nx, ny, nz, nmax)
	{
		if (nx == 0 && ny == 0 && nz == 0) continue;
		double len2 = nx * nx + ny * ny + nz * nz;
		double s = k1 / len2 * exp(kexp * len2);
		double t = 1.0 / len2 - kexp;
		o[0] += s * (1.0 - t * nx * nx);
		o[1] += s * (1.0 - t * ny * ny);
		o[2] += s * (1.0 - t * nz * nz);
		o[3] += s * (-t * nx * ny);
		o[4] += s * (-t * nx * nz);
		o[5] += s * (-t * ny * nz);
	}
	double corr = 1.5 * alpha * a / (sqrt_pi * L);
	o[0] -= corr;
	o[1] -= corr;
	o[2] -= corr;
}

/* diff_tensor.c:748-849 RPY_Smith_Qii_block */
static void smith_q_self(double a, double L, double alpha, int mmax, int nmax, double q[6])
{
	double sqrt_pi = sqrt(M_PI);
	double alpha2 = alpha * alpha;
	double alpha3 = alpha * alpha2;
	double a3 = a * a * a;
	double L3 = L * L * L;
	double k1 = 2 / sqrt_pi;
	double k2 = 2 * alpha3;
	double v[6];
	memset(q, 0, 6 * sizeof(double));
	L1_BALL(mx, my, mz, mmax)
	{
		if (mx == 0 && my == 0 && mz == 0) continue;
		double len2 = mx * mx + my * my + mz * mz;
		double al = alpha * sqrt(len2);
		double al2 = al * al;
		double e = exp(-al2);
		q_tensor(mx, my, mz, v);
		acc6(q, erfc(al) + k1 * al * e, v);
		double s = k1 * k2 * e / len2;
		sub_outer(q, s, mx, my, mz);
	}
	double kexp = -M_PI * M_PI / alpha2;
	double k4 = 4 * M_PI;
	L1_BALL(nx, ny, nz, nmax)
	{
		if (nx == 0 && ny == 0 && nz == 0) continue;
		double len2 = nx * nx + ny * ny + nz * nz;
		double s = k4 * exp(kexp * len2) / len2;
		acc_outer(q, s, nx, ny, nz);
	}
	double corr = alpha3 * a3 / (L3 * 3.0 * sqrt_pi);
	q[0] -= corr;
	q[1] -= corr;
	q[2] -= corr;
}

/* diff_tensor.c:853-942 RPY_Smith_Oij_block at sigma = r/L; the real-space
 * sum includes the m = 0 image. */
static void smith_o_pair(double sx, double sy, double sz, double alpha, int mmax, int nmax, double o[6])
{
	double alpha2 = alpha * alpha;
	double sqrt_pi = sqrt(M_PI);
	double k0 = 2 * alpha / sqrt_pi;
	double v[6];
	memset(o, 0, 6 * sizeof(double));
	L1_BALL(mx, my, mz, mmax)
	{
		double x = mx + sx, y = my + sy, z = mz + sz;
		double len2 = x * x + y * y + z * z;
		double len = sqrt(len2);
		double s = erfc(alpha * len);
		o_tensor(x, y, z, v);
		acc6(o, s, v);
		s = k0 * exp(-alpha2 * len2) / len2;
		acc_outer(o, s, x, y, z);
	}
	double kq = M_PI * M_PI / alpha2;
	double k1 = 2.0 / M_PI;
	L1_BALL(nx, ny, nz, nmax)
	{
		if (nx == 0 && ny == 0 && nz == 0) continue;
		double len2 = nx * nx + ny * ny + nz * nz;
		double dot = nx * sx + ny * sy + nz * sz;
		double ke = kq * len2;
		double s = k1 / len2 * exp(-ke) * cos(2 * M_PI * dot);
		double t = (1.0 + ke) / len2;
		o[0] += s * (1.0 - t * nx * nx);
		o[1] += s * (1.0 - t * ny * ny);
		o[2] += s * (1.0 - t * nz * nz);
		o[3] += s * (-t * nx * ny);
		o[4] += s * (-t * nx * nz);
		o[5] += s * (-t * ny * nz);
	}
}

/* diff_tensor.c:946-1039 RPY_Smith_Qij_block */
static void smith_q_pair(double sx, double sy, double sz, double alpha, int mmax, int nmax, double q[6])
{
	double alpha2 = alpha * alpha;
	double alpha3 = alpha2 * alpha;
	double sqrt_pi = sqrt(M_PI);
	double k1 = 2.0 / sqrt_pi;
	double k2 = 4 * alpha3 / sqrt_pi;
	double v[6];
	memset(q, 0, 6 * sizeof(double));
	L1_BALL(mx, my, mz, mmax)
	{
		double x = mx + sx, y = my + sy, z = mz + sz;
		double len2 = x * x + y * y + z * z;
		double al = alpha * sqrt(len2);
		double al2 = al * al;
		double e = exp(-al2);
		double s = erfc(al) + k1 * al * e;
		q_tensor(x, y, z, v);
		acc6(q, s, v);
		s = k2 * e / len2;
		sub_outer(q, s, x, y, z);
	}
	double kexp = -M_PI * M_PI / alpha2;
	double k4 = 4 * M_PI;
	L1_BALL(nx, ny, nz, nmax)
	{
		if (nx == 0 && ny == 0 && nz == 0) continue;
		double len2 = nx * nx + ny * ny + nz * nz;
		double dot = nx * sx + ny * sy + nz * sz;
		double s = k4 * exp(kexp * len2) * cos(2 * M_PI * dot) / len2;
		acc_outer(q, s, nx, ny, nz);
	}
}

/* diff_tensor.c:553-575 RPY_Smith_Mii_block */
void orc_smith_self(double a, double L, double alpha, int mmax, int nmax, double m[6])
{
	double c1 = 1.0 / (6 * M_PI * a);
	double c2 = 3 * a / (4 * L);
	double c3 = a * a * a / (2 * L * L * L);
	double o[6], q[6];
	smith_o_self(a, L, alpha, mmax, nmax, o);
	smith_q_self(a, L, alpha, mmax, nmax, q);
	for (int k = 0; k < 3; ++k) m[k] = c1 * (1.0 + c2 * o[k] + c3 * q[k]);
	for (int k = 3; k < 6; ++k) m[k] = c1 * (c2 * o[k] + c3 * q[k]);
}

/* diff_tensor.c:579-648 RPY_Smith_Mij_block; overlapping pairs get the
 * free-space overlap/contained correction minus the free-space far form. */
void orc_smith_pair(double ai, double aj, double rx, double ry, double rz, double L, double alpha, int mmax,
		    int nmax, double m[6])
{
	double r2 = rx * rx + ry * ry + rz * rz;
	double c1 = 1.0 / (6 * M_PI * ai);
	double c2 = 3 * ai / (4 * L);
	double c3 = (ai == aj) ? ai * ai * ai / (2 * L * L * L) : ai * (ai * ai + aj * aj) / (4 * L * L * L);
	double o[6], q[6];
	smith_o_pair(rx / L, ry / L, rz / L, alpha, mmax, nmax, o);
	smith_q_pair(rx / L, ry / L, rz / L, alpha, mmax, nmax, q);
	for (int k = 0; k < 6; ++k) m[k] = c1 * (c2 * o[k] + c3 * q[k]);

	if (r2 < (ai + aj) * (ai + aj)) {
		double r = sqrt(r2);
		double s2 = ai * ai + aj * aj;
		double near[6];
		orc_rpy_pair(ai, aj, rx, ry, rz, near);
		for (int k = 0; k < 6; ++k) m[k] += near[k];
		double f1 = 1.0 / (8 * M_PI * r);
		double f2 = 1.0 + s2 / (3 * r2);
		double f3 = (1.0 - s2 / r2) / r2;
		m[0] -= f1 * (f2 + f3 * rx * rx);
		m[1] -= f1 * (f2 + f3 * ry * ry);
		m[2] -= f1 * (f2 + f3 * rz * rz);
		m[3] -= f1 * f3 * rx * ry;
		m[4] -= f1 * f3 * rx * rz;
		m[5] -= f1 * f3 * ry * rz;
	}
}

/* diff_tensor.c:170-266 RPY_Smith_M_matrix.  Only the three diagonal entries
 * of each self block are copied into the packed result (:194-199), so the
 * self-block off-diagonals written at :227-235 are exactly zero (their exact
 * value by cubic symmetry; the lattice sums leave ~1e-18 rounding residue). */
void orc_rpy_smith_m_matrix(const double *a, const double *P, double L, double alpha, int mmax, int nmax, int N,
			    double *M)
{
	size_t n = 3 * (size_t)N;
	for (int j = 0; j < N; ++j) {
		double b[6];
		orc_smith_self(a[j], L, alpha, mmax, nmax, b);
		b[3] = b[4] = b[5] = 0.0;
		put_block(M, n, j, j, b);
		for (int i = j + 1; i < N; ++i) {
			const double *r = P + 3 * pair_slot(j, i, N);
			orc_smith_pair(a[i], a[j], r[0], r[1], r[2], L, alpha, mmax, nmax, b);
			put_block(M, n, i, j, b);
		}
	}
}

/* ------------------------------------------------------------------------ */
/* Jeffrey-Onishi two-sphere resistance: clib/diff_tensor.c:1190-1625        */

/* Power table with the reference's products (l2=l*l, l3=l2*l, l4=l2*l2,
 * l5=l3*l2, l6=l3*l3, l7=l4*l3, l8=l4*l4, l9=l5*l4, l10=l5*l5). */
static void powers(double l, double p[11])
{
	p[0] = 1.0;
	p[1] = l;
	p[2] = l * l;
	p[3] = p[2] * l;
	p[4] = p[2] * p[2];
	p[5] = p[3] * p[2];
	p[6] = p[3] * p[3];
	p[7] = p[4] * p[3];
	p[8] = p[4] * p[4];
	p[9] = p[5] * p[4];
	p[10] = p[5] * p[5];
}

/* Jeffrey & Onishi (1984) f_m(lambda) polynomial coefficients for X^A and
 * Y^A, degree m = 0..11, coefficient of lambda^k at [m][k]
 * (diff_tensor.c:1385-1573).  Sums run in ascending k. */
static const double XF[12][11] = {
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
static const double YF[12][11] = {
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

static double fpoly(const double tab[12][11], int m, const double p[11])
{
	if (m == 0) return 1.0;
	int k = 1;
	while (tab[m][k] == 0.0) ++k;
	double v = tab[m][k] * p[k];
	for (++k; k < 11; ++k)
		if (tab[m][k] != 0.0) v = v + tab[m][k] * p[k];
	return v;
}

/* g-polynomials (diff_tensor.c:1577-1625) */
static double xg(int deg, double l)
{
	double l2 = l * l, plus = 1 + l, plus3 = plus * plus * plus;
	if (deg == 1) return 2 * l2 / plus3;
	if (deg == 2) return 0.2 * l * (1 + 7 * l + l2) / plus3;
	double l3 = l2 * l, l4 = l2 * l2;
	return 1.0 / 42 * (1 + 18 * l - 29 * l2 + 18 * l3 + l4) / plus3;
}

static double yg(int deg, double l)
{
	double l2 = l * l, plus = 1 + l, plus3 = plus * plus * plus;
	if (deg == 2) return 4.0 / 15 * l * (2 + l + 2 * l2) / plus3;
	double l3 = l2 * l, l4 = l2 * l2;
	return 2.0 / 375 * (16 - 45 * l + 58 * l2 - 45 * l3 + 16 * l4) / plus3;
}

/* diff_tensor.c:1190-1238 */
double orc_jo_xa11(double s, double l)
{
	double p[11];
	powers(l, p);
	double g1 = xg(1, l), g2 = xg(2, l), g3 = xg(3, l);
	double s2 = 1.0 - 4.0 / (s * s);
	double ls = log(s2);
	double v = 0.0;
	v += g1 / s2;
	v -= g2 * ls;
	v -= g3 * s2 * ls;
	v += 1.0 - g1;
	for (int m = 2; m < 12; m += 2) {
		int m1 = (m == 2) ? -2 : m - 2;
		double w = pow(2 / s, m);
		v += w * (pow(2, -m) * pow(1 + l, -m) * fpoly(XF, m, p) - g1);
		v += w * (4.0 / (m * m1) * g3 - 2.0 / m * g2);
	}
	return v;
}

/* diff_tensor.c:1242-1285 */
double orc_jo_ya11(double s, double l)
{
	double p[11];
	powers(l, p);
	double g2 = yg(2, l), g3 = yg(3, l);
	double s2 = 1.0 - 4.0 / (s * s);
	double ls = log(s2);
	double v = 0.0;
	v -= g2 * ls;
	v -= g3 * s2 * ls;
	v += 1.0;
	for (int m = 2; m < 12; m += 2) {
		int m1 = (m == 2) ? -2 : m - 2;
		double w = pow(2 / s, m);
		v += w * (pow(2, -m) * pow(1 + l, -m) * fpoly(YF, m, p) - 2.0 / m * g2);
		v += w * 4.0 / (m * m1) * g3;
	}
	return v;
}

/* diff_tensor.c:1289-1334 (odd m; overall sign minus) */
double orc_jo_xa12(double s, double l)
{
	double p[11];
	powers(l, p);
	double g1 = xg(1, l), g2 = xg(2, l), g3 = xg(3, l);
	double s2 = 1.0 - 4.0 / (s * s);
	double ls = log((s + 2) / (s - 2));
	double v = 0.0;
	v += 2.0 / s * g1 / s2;
	v += g2 * ls;
	v += g3 * s2 * ls + 4 * g3 / s;
	for (int m = 1; m < 12; m += 2) {
		int m1 = m - 2;
		double w = pow(2 / s, m);
		v += w * (pow(2, -m) * pow(1 + l, -m) * fpoly(XF, m, p) - g1);
		v += w * (4.0 / (m * m1) * g3 - 2.0 / m * g2);
	}
	return -v;
}

/* diff_tensor.c:1338-1381 */
double orc_jo_ya12(double s, double l)
{
	double p[11];
	powers(l, p);
	double g2 = yg(2, l), g3 = yg(3, l);
	double s2 = 1.0 - 4.0 / (s * s);
	double ls = log((s + 2) / (s - 2));
	double v = 0.0;
	v += g2 * ls;
	v += g3 * s2 * ls;
	v += 4 * g3 / s;
	for (int m = 1; m < 12; m += 2) {
		int m1 = m - 2;
		double w = pow(2 / s, m);
		v += w * (pow(2, -m) * pow(1 + l, -m) * fpoly(YF, m, p) - 2.0 / m * g2);
		v += w * (4.0 / (m * m1) * g3);
	}
	return -v;
}

/* Fill one symmetric 3x3 block X rr^T/r^2 + Y (I - rr^T/r^2) into out[6]
 * with the reference's expression order (diff_tensor.c:302-336). */
static void xy_block(double X, double Y, double rx, double ry, double rz, double r2, double out[6])
{
	out[0] = X * rx * rx / r2 + Y * (1 - rx * rx / r2);
	out[1] = X * ry * ry / r2 + Y * (1 - ry * ry / r2);
	out[2] = X * rz * rz / r2 + Y * (1 - rz * rz / r2);
	out[3] = (X - Y) * rx * ry / r2;
	out[4] = (X - Y) * rx * rz / r2;
	out[5] = (X - Y) * ry * rz / r2;
}

/* diff_tensor.c:270-342 JO_2B_R_matrix: 18 values = block 00 (sphere i),
 * block 11 (sphere j), block 10 (coupling), each (xx,yy,zz,xy,xz,yz), times
 * 6 pi a_i. */
void orc_jo_2b_r_matrix(double ai, double aj, double rx, double ry, double rz, double R[18])
{
	double r2 = rx * rx + ry * ry + rz * rz;
	double r = sqrt(r2);
	double s = 2 * r / (ai + aj);
	double l = aj / ai;
	double x11 = orc_jo_xa11(s, l);
	double y11 = orc_jo_ya11(s, l);
	double x22 = l * orc_jo_xa11(s, 1 / l);
	double y22 = l * orc_jo_ya11(s, 1 / l);
	double x12 = orc_jo_xa12(s, l);
	double y12 = orc_jo_ya12(s, l);
	double scale = 6 * M_PI * ai;
	xy_block(x11, y11, rx, ry, rz, r2, R);
	xy_block(x22, y22, rx, ry, rz, r2, R + 6);
	xy_block(x12, y12, rx, ry, rz, r2, R + 12);
	for (int k = 0; k < 18; ++k) R[k] *= scale;
}

/* Gauss-Jordan inverse with partial pivoting of a row-major n x n matrix
 * (n <= 8).  Returns 0 on success, 1 if a zero pivot appears. */
int orc_invert_small(double *A, int n)
{
	double W[8][16];
	for (int r = 0; r < n; ++r)
		for (int c = 0; c < 2 * n; ++c) W[r][c] = c < n ? A[r * n + c] : (c - n == r ? 1.0 : 0.0);
	for (int k = 0; k < n; ++k) {
		int piv = k;
		for (int r = k + 1; r < n; ++r)
			if (fabs(W[r][k]) > fabs(W[piv][k])) piv = r;
		if (W[piv][k] == 0.0) return 1;
		if (piv != k)
			for (int c = 0; c < 2 * n; ++c) {
				double t = W[k][c];
				W[k][c] = W[piv][c];
				W[piv][c] = t;
			}
		double inv = 1.0 / W[k][k];
		for (int c = 0; c < 2 * n; ++c) W[k][c] *= inv;
		for (int r = 0; r < n; ++r) {
			if (r == k) continue;
			double f = W[r][k];
			if (f == 0.0) continue;
			for (int c = 0; c < 2 * n; ++c) W[r][c] -= f * W[k][c];
		}
	}
	for (int r = 0; r < n; ++r)
		for (int c = 0; c < n; ++c) A[r * n + c] = W[r][n + c];
	return 0;
}

/* diff_tensor.c:1091-1162 RPY_2B_R_matrix: inverse of the 6x6 two-sphere
 * free-space RPY mobility, packed like orc_jo_2b_r_matrix.  The reference
 * inverts with LAPACK LU (dgetrf/dgetri); this uses Gauss-Jordan. */
void orc_rpy_2b_r_matrix(double ai, double aj, double rx, double ry, double rz, double R[18])
{
	double mi = 1.0 / (6 * M_PI * ai);
	double mj = 1.0 / (6 * M_PI * aj);
	double b[6];
	orc_rpy_pair(ai, aj, rx, ry, rz, b);
	double full[3][3] = {{b[0], b[3], b[4]}, {b[3], b[1], b[5]}, {b[4], b[5], b[2]}};
	double X[36];
	memset(X, 0, sizeof(X));
	for (int k = 0; k < 3; ++k) {
		X[k * 6 + k] = mi;
		X[(k + 3) * 6 + k + 3] = mj;
	}
	for (int r = 0; r < 3; ++r)
		for (int c = 0; c < 3; ++c) {
			X[r * 6 + 3 + c] = full[r][c];
			X[(3 + r) * 6 + c] = full[r][c];
		}
	orc_invert_small(X, 6);
	/* element picks follow diff_tensor.c:1134-1159 (row-major view) */
	static const int pick[18] = {0, 7, 14, 6, 12, 13, 21, 28, 35, 27, 33, 34, 3, 10, 17, 9, 15, 16};
	for (int k = 0; k < 18; ++k) R[k] = X[pick[k]];
}

/* diff_tensor.c:1166-1186 Cichocki_2B_R_correction: project out rigid
 * co-translation of the pair. */
void orc_cichocki(const double t[18], double out[18])
{
	for (int k = 0; k < 6; ++k) {
		double d = (t[k] + t[k + 6] - 2 * t[k + 12]) / 4;
		out[k] = d;
		out[k + 6] = d;
		out[k + 12] = (2 * t[k + 12] - t[k] - t[k + 6]) / 4;
	}
}

/* diff_tensor.c:346-479 JO_R_lubrication_correction_matrix.  ACCUMULATES
 * into C.  For a near pair (i>j) the radii are passed as (a_j, a_i); block 00
 * lands on bead j, block 11 on bead i, block 10 on the (i,j)/(j,i) pair. */
void orc_jo_r_lubrication_correction_matrix(const double *a, const double *P, int N, double cutoff, int cichocki,
					    double *C)
{
	size_t n = 3 * (size_t)N;
	for (int j = 0; j < N; ++j)
		for (int i = j + 1; i < N; ++i) {
			const double *r = P + 3 * pair_slot(j, i, N);
			double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
			double dist = sqrt(r2);
			if (!((dist - a[j] - a[i]) / (a[j] + a[i]) <= cutoff)) continue;
			double nf[18], ff[18], res[18];
			if (cichocki) {
				double t[18];
				orc_jo_2b_r_matrix(a[j], a[i], r[0], r[1], r[2], t);
				orc_cichocki(t, nf);
				orc_rpy_2b_r_matrix(a[j], a[i], r[0], r[1], r[2], t);
				orc_cichocki(t, ff);
			} else {
				orc_jo_2b_r_matrix(a[j], a[i], r[0], r[1], r[2], nf);
				orc_rpy_2b_r_matrix(a[j], a[i], r[0], r[1], r[2], ff);
			}
			for (int k = 0; k < 18; ++k) res[k] = nf[k] - ff[k];
			add_block(C, n, j, j, res);
			add_block(C, n, i, i, res + 6);
			add_block(C, n, i, j, res + 12);
		}
}

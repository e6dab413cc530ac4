// Forces and energy of the step's force field on gfx950 (box.py:404-428):
//   F = unit * (external force, zeroed outside the region per axis
//               + Lennard-Jones over all pairs + harmonic bonds + harmonic angles)
//   E = sum of the interaction energies (not scaled by the unit).
// Formulas follow interactions.py:641-821 and bead.py:272-278; pair vectors
// are the minimum-image r_j - r_i of bead.c:44-50.
//
// One thread per mobile bead, deterministic: each bead sums its own pair
// forces (the force of pair (i, j) on i is odd in r_j - r_i, so the
// reference's "F[i] += f, F[j] -= f" contributions are reproduced bit for
// bit and only their order differs), and its bonds and angles through
// per-bead incidence lists built on the host.  A workgroup holds 64 beads;
// its four waves split the partner range and combine in a fixed order.  Pair
// energies are counted on the bead with the larger index (on the mobile bead
// for a mobile-immobile pair), bond and angle energies on their first mobile
// participant; one workgroup then sums the per-bead energies in a fixed order.
//
// Immobile beads (indices N..Nt-1) get no force but act on mobile ones:
// Lennard-Jones partners of every mobile bead, and bond and angle members,
// with the pair vectors built exactly as interactions.py:185-416 builds them
// from the mobile-mobile and mobile-immobile pointer matrices (some are
// negated pointers or differences of two, which the rounding can tell apart).
#include "pbx_common.h"

namespace {

constexpr int BEADS = 64;

__device__ __forceinline__ void lj_pair(double px, double py, double pz, double sigma, double epsilon, double alpha,
                                        int kind, double f[3], double& e) {
    const double dist2 = px * px + py * py + pz * pz;
    const double dist = sqrt(dist2);
    const double s2 = sigma * sigma / dist2;
    double coef;
    if (kind == 3) {  // LJ_6_12_force / _energy (interactions.py:783-821)
        const double s6 = s2 * s2 * s2;
        const double s12 = s6 * s6;
        coef = alpha * epsilon / dist * (6 * s6 - 12 * s12);
        e = alpha * epsilon * (s12 - s6);
    } else if (kind == 1) {  // LJ_6_attractive (:711-745)
        const double s6 = s2 * s2 * s2;
        coef = 6.0 * alpha * epsilon * s6 / dist;
        e = -alpha * epsilon * s6;
    } else {  // LJ_12_repulsive (:747-781)
        const double s12 = s2 * s2 * s2 * s2 * s2 * s2;
        coef = -12.0 * alpha * epsilon * s12 / dist;
        e = alpha * epsilon * s12;
    }
    f[0] = coef * (px / dist);
    f[1] = coef * (py / dist);
    f[2] = coef * (pz / dist);
}

// p = wrap(r_b - r_a) (bead.c:75 and :99); sign -1 gives -wrap(r_b - r_a).
__device__ __forceinline__ void ptr3(const double* pos, int a, int b, double L, double p[3], double sign = 1.0) {
    for (int c = 0; c < 3; ++c) p[c] = sign * pbx_min_image(pos[3 * b + c] - pos[3 * a + c], L);
}

// Pair vectors (p12, p23) of angle (i, j, k) as interactions.py:291-416
// forms them; beads >= N are immobile.  The caller only asks for angles with
// at least one mobile member.
__device__ __forceinline__ void angle_pointers(const double* pos, int i, int j, int k, int N, double L,
                                               double p12[3], double p23[3]) {
    const bool mi = i < N, mj = j < N, mk = k < N;
    if (mi) {  // :295-344, the bead listing the angle moves
        ptr3(pos, i, j, L, p12);
        if (mj) {
            ptr3(pos, j, k, L, p23);
        } else if (mk) {
            ptr3(pos, k, j, L, p23, -1.0);  // -rik[k][j]
        } else {
            double pik[3];  // rik[i][k] - p12
            ptr3(pos, i, k, L, pik);
            for (int c = 0; c < 3; ++c) p23[c] = pik[c] - p12[c];
        }
    } else if (mj) {  // :349-384, an immobile bead lists it
        ptr3(pos, j, i, L, p12, -1.0);  // -rik[j][i]
        ptr3(pos, j, k, L, p23);
    } else {
        double pik[3];
        ptr3(pos, k, j, L, p23, -1.0);  // -rik[k][j]
        ptr3(pos, k, i, L, pik, -1.0);  // -rik[k][i]
        for (int c = 0; c < 3; ++c) p12[c] = pik[c] - p23[c];
    }
}

__global__ __launch_bounds__(256) void forces_kernel(PbxForceArgs a, const double* pos, double* F, double* e_part) {
    __shared__ double part[4][BEADS][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = blockIdx.x * BEADS + lane;
    const bool live = b < a.N;
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    if (a.lj && live) {
        const double rb[3] = {pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]};
        const double hb = a.hc[b], eb = a.eps[b];
        // partners: every other mobile bead, then the immobile ones
        // (interactions.py:251-289)
        const int chunk = (a.Nt + 3) / 4;
        const int j0 = w * chunk, j1 = min(a.Nt, j0 + chunk);
        for (int j = j0; j < j1; ++j) {
            if (j == b) continue;
            const double px = pbx_min_image(pos[3 * j] - rb[0], a.L);
            const double py = pbx_min_image(pos[3 * j + 1] - rb[1], a.L);
            const double pz = pbx_min_image(pos[3 * j + 2] - rb[2], a.L);
            double f[3], ee;
            lj_pair(px, py, pz, hb + a.hc[j], sqrt(eb * a.eps[j]), a.alpha, a.lj, f, ee);
            fx += f[0];
            fy += f[1];
            fz += f[2];
            if (j < b || j >= a.N) e += ee;
        }
    }
    part[w][lane][0] = fx;
    part[w][lane][1] = fy;
    part[w][lane][2] = fz;
    part[w][lane][3] = e;
    __syncthreads();
    if (w != 0 || !live) return;
    // external force (box.py:708-723): Fex unless outside an enabled axis range
    double F3[3] = {a.fex[0], a.fex[1], a.fex[2]};
    for (int c = 0; c < 3; ++c)
        if (a.reg_on[c] && (pos[3 * b + c] < a.reg_lo[c] || pos[3 * b + c] > a.reg_hi[c])) {
            F3[0] = F3[1] = F3[2] = 0.0;
            break;
        }
    double E = 0.0;
    for (int q = 0; q < 4; ++q) {
        F3[0] += part[q][lane][0];
        F3[1] += part[q][lane][1];
        F3[2] += part[q][lane][2];
        E += part[q][lane][3];
    }
    // harmonic bonds (interactions.py:641-665): entry = 2 * bond + role.
    // Listed by a mobile bead i: p = wrap(r_j - r_i) (:189-230); by an
    // immobile one: p = -rik[j][i] and only j moves (:232-245).
    for (int t = a.bond_ptr[b]; t < a.bond_ptr[b + 1]; ++t) {
        const int q = a.bond_ent[t] >> 1, role = a.bond_ent[t] & 1;
        const int bi = a.bond_ij[2 * q], bj = a.bond_ij[2 * q + 1];
        double p[3];
        if (bi < a.N)
            ptr3(pos, bi, bj, a.L, p);
        else
            ptr3(pos, bj, bi, a.L, p, -1.0);
        const double deq = a.bond_par[2 * q], k = a.bond_par[2 * q + 1];
        const double dist = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        const double s = role ? -1.0 : 1.0;
        for (int c = 0; c < 3; ++c) F3[c] += s * (p[c] / dist * k * (dist - deq));
        if (role == 0 || bi >= a.N) E += 0.5 * k * (dist - deq) * (dist - deq);
    }
    // harmonic angles (interactions.py:667-709): entry = 4 * angle + role
    for (int t = a.ang_ptr[b]; t < a.ang_ptr[b + 1]; ++t) {
        const int q = a.ang_ent[t] >> 2, role = a.ang_ent[t] & 3;
        const int i = a.ang_ijk[3 * q], j = a.ang_ijk[3 * q + 1], k3 = a.ang_ijk[3 * q + 2];
        double p12[3], p23[3];
        angle_pointers(pos, i, j, k3, a.N, a.L, p12, p23);
        // the first mobile member carries the energy
        const int first = i < a.N ? 0 : j < a.N ? 1 : 2;
        const double d2_12 = p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2];
        const double d2_23 = p23[0] * p23[0] + p23[1] * p23[1] + p23[2] * p23[2];
        const double d12 = sqrt(d2_12), d23 = sqrt(d2_23);
        const double dot = p12[0] * p23[0] + p12[1] * p23[1] + p12[2] * p23[2];
        double th = acos(-dot / d12 / d23) * (180.0 / M_PI);  // angle_pbc, degrees
        th *= M_PI / 180.0;
        const double th_eq = a.ang_par[2 * q] * (M_PI / 180.0), kc = a.ang_par[2 * q + 1];
        const double ct = cos(th);
        const double pre = kc * (th - th_eq) / sin(th);
        for (int c = 0; c < 3; ++c) {
            double sc;
            if (role == 0)
                sc = (p23[c] / d23 + ct * p12[c] / d12) / d12;
            else if (role == 1)
                sc = (p12[c] - p23[c]) / d12 / d23 - ct * (p12[c] / d2_12 - p23[c] / d2_23);
            else
                sc = -p12[c] / d12 / d23 - ct * p23[c] / d2_23;
            F3[c] += pre * sc;
        }
        if (role == first) E += 0.5 * kc * (th - th_eq) * (th - th_eq);
    }
    for (int c = 0; c < 3; ++c) F[3 * b + c] = a.unit == 1.0 ? F3[c] : a.unit * F3[c];
    e_part[b] = E;
}

// E = sum of e_part in a fixed order (one workgroup)
__global__ __launch_bounds__(256) void energy_sum_kernel(const double* e_part, int N, double* E) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) acc += e_part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *E = s[0];
}

}  // namespace

int pbx_launch_forces(const PbxForceArgs& a, const double* pos, double* F, double* e_part, double* E,
                      hipStream_t s) {
    PBX_LAUNCH(forces_kernel, dim3((a.N + BEADS - 1) / BEADS), dim3(256), 0, s, a, pos, F, e_part);
    PBX_LAUNCH(energy_sum_kernel, dim3(1), dim3(256), 0, s, e_part, a.N, E);
    PBX_HIP(hipGetLastError());
    return PBX_OK;
}

// Native N(0,1) stream for the move attempts: the sequence that
// numpy.random.RandomState(seed).normal(0.0, 1.0, k) returns (box.py:393-399
// draws it 3N at a time; box.py:621-629 seeds it with an integer).
//
//  * MT19937 (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998) with the
//    init_genrand seeding: key[0] = seed,
//    key[i] = 1812433253 (key[i-1] ^ (key[i-1] >> 30)) + i  (mod 2^32);
//  * uniform doubles on [0, 1) from two outputs, 53 bits:
//    ((a >> 5) 2^26 + (b >> 6)) / 2^53;
//  * Marsaglia's polar method (Marsaglia & Bray, SIAM Review 6(3), 1964):
//    x1, x2 = 2u - 1 until 0 < r2 = x1^2 + x2^2 < 1, f = sqrt(-2 ln r2 / r2);
//    f x2 is returned and f x1 kept for the following call, as numpy's
//    legacy generator does, so the stream is the same however it is split
//    into calls.
//
// Host code only; floating-point contraction is off so each product and sum
// rounds separately.
#include <math.h>
#include <stdint.h>

#include "pbx_noise.h"

#pragma clang fp contract(off)

namespace {

constexpr int MT_N = 624;
constexpr int MT_M = 397;
constexpr uint32_t MATRIX_A = 0x9908b0dfu;
constexpr uint32_t UPPER = 0x80000000u;
constexpr uint32_t LOWER = 0x7fffffffu;

void regenerate(PbxMT* g) {
    uint32_t* mt = g->mt;
    int k = 0;
    for (; k < MT_N - MT_M; ++k) {
        uint32_t y = (mt[k] & UPPER) | (mt[k + 1] & LOWER);
        mt[k] = mt[k + MT_M] ^ (y >> 1) ^ ((y & 1u) ? MATRIX_A : 0u);
    }
    for (; k < MT_N - 1; ++k) {
        uint32_t y = (mt[k] & UPPER) | (mt[k + 1] & LOWER);
        mt[k] = mt[k + (MT_M - MT_N)] ^ (y >> 1) ^ ((y & 1u) ? MATRIX_A : 0u);
    }
    uint32_t y = (mt[MT_N - 1] & UPPER) | (mt[0] & LOWER);
    mt[MT_N - 1] = mt[MT_M - 1] ^ (y >> 1) ^ ((y & 1u) ? MATRIX_A : 0u);
    g->pos = 0;
}

inline uint32_t next32(PbxMT* g) {
    if (g->pos >= MT_N) regenerate(g);
    uint32_t y = g->mt[g->pos++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

inline double next_double(PbxMT* g) {
    const int32_t a = (int32_t)(next32(g) >> 5), b = (int32_t)(next32(g) >> 6);
    return (a * 67108864.0 + b) / 9007199254740992.0;
}

inline double gauss(PbxMT* g) {
    if (g->has_gauss) {
        g->has_gauss = 0;
        const double t = g->gauss;
        g->gauss = 0.0;
        return t;
    }
    double x1, x2, r2;
    do {
        x1 = 2.0 * next_double(g) - 1.0;
        x2 = 2.0 * next_double(g) - 1.0;
        r2 = x1 * x1 + x2 * x2;
    } while (r2 >= 1.0 || r2 == 0.0);
    const double f = sqrt(-2.0 * log(r2) / r2);
    g->gauss = f * x1;
    g->has_gauss = 1;
    return f * x2;
}

}  // namespace

void pbx_mt_seed(PbxMT* g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < MT_N; ++i) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->pos = MT_N;
    g->has_gauss = 0;
    g->gauss = 0.0;
}

// normal(loc = 0, scale = 1): loc + scale * gauss (exact for these values)
void pbx_mt_normals(PbxMT* g, double* out, long long count) {
    for (long long i = 0; i < count; ++i) out[i] = 0.0 + 1.0 * gauss(g);
}

extern "C" int pbx_rng_normals(unsigned int seed, long long skip, double* out, long long count) {
    if (skip < 0 || count < 0 || (count > 0 && !out)) return -1;
    PbxMT g;
    pbx_mt_seed(&g, seed);
    double sink;
    for (long long i = 0; i < skip; ++i) pbx_mt_normals(&g, &sink, 1);
    pbx_mt_normals(&g, out, count);
    return 0;
}

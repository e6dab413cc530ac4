// MT19937 + polar Gaussian state (pbx_noise.hip).
#pragma once
#include <stdint.h>

struct PbxMT {
    uint32_t mt[624];
    int pos;
    int has_gauss;
    double gauss;
};

void pbx_mt_seed(PbxMT* g, uint32_t seed);
void pbx_mt_normals(PbxMT* g, double* out, long long count);

// fgss_bounds.h -- float32 similarity bounds shared by host and device code.
//
// These are the reference's Jaccard bound expressions
// (include/fgssjoin/similarity.cuh:9-82), kept in the same operation order so
// that every float32 result is bit-identical:
//   min_size(l)      = (float)l * t
//   max_size(l)      = (float)l / t
//   min_overlap(a,b) = (t / (1 + t)) * (float)(a + b)
//   max_prefix(l)    = (unsigned)((float)l - ceil(min_size(l)) + 1)
//   mid_prefix(l)    = (unsigned)((float)l - ceil(min_overlap(l,l)) + 1)
//   similarity       = (float)inter / (float)(a + b - inter)
// All code including this header is compiled with -ffp-contract=off and without
// fast-math, and device division stays the correctly rounded IEEE divide.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FGSS_HD __host__ __device__ __forceinline__

FGSS_HD float fgss_min_size(uint32_t l, float t) { return (float)l * t; }
FGSS_HD float fgss_max_size(uint32_t l, float t) { return (float)l / t; }
FGSS_HD float fgss_min_overlap(uint32_t a, uint32_t b, float t)
{
    float r = t / (1.0f + t);
    return r * (float)(a + b);
}
FGSS_HD uint32_t fgss_max_prefix(uint32_t l, float t)
{
    return (uint32_t)((float)l - ceilf(fgss_min_size(l, t)) + 1.0f);
}
FGSS_HD uint32_t fgss_mid_prefix(uint32_t l, float t)
{
    return (uint32_t)((float)l - ceilf(fgss_min_overlap(l, l, t)) + 1.0f);
}
FGSS_HD float fgss_similarity(uint32_t a, uint32_t b, uint32_t inter)
{
    return (float)inter / (float)(a + b - inter);
}
// floor(min_size) and floor(max_size) as unsigned (filtering.cu:90-92, :114-115)
FGSS_HD uint32_t fgss_floor_min_size(uint32_t l, float t) { return (uint32_t)floorf(fgss_min_size(l, t)); }
FGSS_HD uint32_t fgss_floor_max_size(uint32_t l, float t) { return (uint32_t)floorf(fgss_max_size(l, t)); }

// The 4-way size filter of filtering.cu:117-120, true when the pair survives.
FGSS_HD bool fgss_size_ok(uint32_t lq, uint32_t ls, float t)
{
    return !(ls < fgss_floor_min_size(lq, t) || ls > fgss_floor_max_size(lq, t) ||
             lq < fgss_floor_min_size(ls, t) || lq > fgss_floor_max_size(ls, t));
}

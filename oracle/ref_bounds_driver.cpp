// Driver around the reference's float32 bound expressions
// (include/fgssjoin/similarity.cuh:9-82), compiled host-only by hipcc (which
// understands its __host__ __device__ qualifiers) into oracle/_ref/libref_bounds.so.
// TEST INFRASTRUCTURE ONLY: it pins the oracle's restatement of those expressions.
#include <stdint.h>

#include "similarity.cuh"

extern "C" float ref_min_size(uint32_t l, float t) { return jac_min_size(l, t); }
extern "C" float ref_max_size(uint32_t l, float t) { return jac_max_size(l, t); }
extern "C" float ref_min_overlap(uint32_t a, uint32_t b, float t) { return jac_min_overlap(a, b, t); }
extern "C" uint32_t ref_max_prefix(uint32_t l, float t) { return jac_max_prefix(l, t); }
extern "C" uint32_t ref_mid_prefix(uint32_t l, float t) { return jac_mid_prefix(l, t); }
extern "C" float ref_similarity(uint32_t a, uint32_t b, uint32_t i) { return jac_similarity(a, b, i); }

/* edenhip_diag.h -- self-tests of device arithmetic used by libedenhip's kernels.
 *
 * No counterpart in the reference: these entry points check, on the GPU, arithmetic that
 * the kernels rely on to reproduce the reference's fp32 results bit for bit. */
#ifndef EDENHIP_DIAG_H
#define EDENHIP_DIAG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compare the kernels' VCC-free division with the compiler's fp32 division on `count`
 * hashed operand pairs (half with exponents near the fast path's range) and on every pair
 * of the `n_special` given values. *mismatches gets the number of pairs whose results
 * differ in any bit (NaNs compare equal). Returns 0, or -1 for bad arguments, -2 for a
 * device error. */
int edenhip_diag_divide(uint64_t seed, int64_t count, const float *special, int32_t n_special, int64_t *mismatches);

#ifdef __cplusplus
}
#endif
#endif /* EDENHIP_DIAG_H */

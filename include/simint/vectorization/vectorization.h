/* Memory-layout macros of the simint API (skel/simint/vectorization/vectorization.h).
 *
 * The GPU build keeps the host structure-of-arrays unpadded: SIMINT_SIMD_LEN
 * is 1, so SIMINT_SIMD_ROUND(x) == x and shell pairs carry no padding
 * primitives.  Device copies are re-laid-out by the library. */
#pragma once

#include <stdlib.h>

#define SIMINT_SIMD_LEN 1
#define SIMINT_SIMD_ALIGN_DBL (SIMINT_SIMD_LEN * 8)
#define SIMINT_SIMD_ALIGN_INT (SIMINT_SIMD_LEN * sizeof(int))
#define SIMINT_SIMD_ALIGN SIMINT_SIMD_ALIGN_DBL

#define SIMINT_ASSUME_ALIGN_DBL(x)
#define SIMINT_ASSUME_ALIGN_INT(x)

#define SIMINT_ALLOC(x) malloc((x))
#define SIMINT_FREE(x) free((x))

/* round a count of elements up to the SIMD boundary */
#define SIMINT_SIMD_ROUND(x) (((x) + ((SIMINT_SIMD_LEN - 1))) & (~(SIMINT_SIMD_LEN - 1)))

#define SIMINT_ALIGN_ARRAY_DBL __attribute__((aligned(SIMINT_SIMD_ALIGN_DBL)))
#define SIMINT_ALIGN_ARRAY_INT __attribute__((aligned(SIMINT_SIMD_ALIGN_INT)))

/* shell pairs per batch in the reference's CPU kernels */
#define SIMINT_NSHELL_SIMD (2 * SIMINT_SIMD_LEN)

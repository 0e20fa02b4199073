/* One-electron integrals (skel/simint/simint_oneelectron.h:12-25): overlap,
 * kinetic energy and the nuclear attraction of point charges for one pair
 * of shells, ncart(a->am) x ncart(b->am) values in Psi4 order.  Same
 * signatures and meaning as the reference; each call runs on the GPU
 * (simint_hip_compute_osoei) and returns with the integrals in `integrals`.
 * Return value: 1 on success, or a negative SIMINT_ERROR_* code
 * (SIMINT_ERROR_UNSUPPORTED above angular momentum 5). */
#pragma once

#include "simint/shell/shell.h"
#include "simint/ostei/ostei_config.h"

#ifdef __cplusplus
#include "simint/cpp_restrict.hpp"
extern "C" {
#endif

int simint_compute_overlap(struct simint_shell * a,
                           struct simint_shell * b,
                           double * restrict integrals);

int simint_compute_ke(struct simint_shell * a,
                      struct simint_shell * b,
                      double * restrict integrals);

/* -sum_C Z_C <a|1/|r - C||b> over the ncenter charges Z at (x, y, z) */
int simint_compute_potential(int ncenter,
                             double * Z, double * x, double * y, double * z,
                             struct simint_shell const * sh1,
                             struct simint_shell const * sh2,
                             double * restrict integrals);

#ifdef __cplusplus
}
#endif

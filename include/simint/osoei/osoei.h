/* The one-electron functions behind simint_oneelectron.h
 * (skel/simint/osoei/osoei.h:27-36), same signatures.  The reference's
 * host helpers simint_osoei_overlap_terms / _ke_terms (osoei.h:11-23, the
 * per-primitive 1D recurrence tables) are not exported: here the tables
 * live in the GPU kernel's local data share. */
#pragma once

#include "simint/shell/shell.h"

#ifdef __cplusplus
#include "simint/cpp_restrict.hpp"
extern "C" {
#endif

int simint_compute_osoei_overlap(struct simint_shell const * sh1,
                                 struct simint_shell const * sh2,
                                 double * restrict integrals);

int simint_compute_osoei_ke(struct simint_shell const * sh1,
                            struct simint_shell const * sh2,
                            double * restrict integrals);

int simint_compute_osoei_potential(int ncenter,
                                   double * Z, double * x, double * y, double * z,
                                   struct simint_shell const * sh1,
                                   struct simint_shell const * sh2,
                                   double * restrict integrals);

#ifdef __cplusplus
}
#endif

/* Class-dispatch entry points (skel/simint/ostei/ostei.h:11-56).  Same
 * semantics as simint_compute_eri / simint_compute_eri_deriv. */
#pragma once

#include "simint/shell/shell.h"

#ifdef __cplusplus
#include "simint/cpp_restrict.hpp"
extern "C" {
#endif

/* Pointer to a class function of the reference's dispatch table
 * (skel/simint/ostei/ostei.h:11-15): shell pairs passed by value, screening
 * tolerance, workspace, output.  Kept for source compatibility; this library
 * dispatches to device kernels internally and exports no table of these. */
typedef int (*simint_osteifunc)(struct simint_multi_shellpair const,
                                struct simint_multi_shellpair const,
                                double,
                                double * restrict,
                                double * restrict);

int simint_compute_ostei(struct simint_multi_shellpair const * P,
                         struct simint_multi_shellpair const * Q,
                         double screen_tol,
                         double * restrict work,
                         double * restrict integrals);

int simint_compute_ostei_deriv(int deriv,
                               struct simint_multi_shellpair const * P,
                               struct simint_multi_shellpair const * Q,
                               double screen_tol,
                               double * restrict work,
                               double * restrict integrals);

#ifdef __cplusplus
}
#endif

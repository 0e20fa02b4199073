/* Two-electron repulsion integrals (skel/simint/simint_eri.h:22-62).
 *
 * simint_compute_eri(P, Q, screen_tol, work, integrals)
 *   Computes (ab|cd) for every bra shell pair in P (the first
 *   P->nshell12_clip) against every ket shell pair in Q, on the GPU, and
 *   returns when the results are in `integrals` (host memory):
 *     integrals[(ab * Q->nshell12_clip + cd) * ncart + ((a*nb + b)*nc + c)*nd + d]
 *   with Cartesian components in Psi4 order.  Returns the number of shell
 *   quartets computed, or -1 when the whole batch is screened out
 *   (P->screen_max * Q->screen_max < screen_tol^2); the caller then treats
 *   the integrals as zero, as with the reference.  Other negative values are
 *   errors (SIMINT_ERROR_*).  `work` is accepted for compatibility and not
 *   used.  Primitive pairs whose screening values satisfy
 *   P->screen[i] * Q->screen[j] < screen_tol^2 are skipped, one primitive
 *   quartet at a time.  If screen_tol > 0 but the pairs carry no screening
 *   values (created with SIMINT_SCREEN_NONE), no primitive is skipped.
 *
 * simint_compute_eri_deriv(deriv, ...)
 *   deriv = 0 is simint_compute_eri; deriv = 1 gives first derivatives with
 *   respect to the 12 centre coordinates, 12 doubles per Cartesian quartet
 *   (OSTEI_HRR_Writer.cpp:145-146).
 *
 * Thread safety: compute calls may run concurrently from several host
 * threads (each thread gets its own HIP stream and device buffers). */
#pragma once

#include <stddef.h>
#include "simint/shell/shell.h"
#include "simint/ostei/ostei_config.h"

#ifdef __cplusplus
#include "simint/cpp_restrict.hpp"
extern "C" {
#endif

#define SIMINT_ERROR_NO_DEVICE    (-2)
#define SIMINT_ERROR_UNSUPPORTED  (-3)
#define SIMINT_ERROR_HIP          (-4)
#define SIMINT_ERROR_ARGUMENT     (-5)

int simint_compute_eri(struct simint_multi_shellpair const * P,
                       struct simint_multi_shellpair const * Q,
                       double screen_tol,
                       double * restrict work,
                       double * restrict integrals);

int simint_compute_eri_deriv(int deriv,
                             struct simint_multi_shellpair const * P,
                             struct simint_multi_shellpair const * Q,
                             double screen_tol,
                             double * restrict work,
                             double * restrict integrals);

size_t simint_eri_worksize(int derorder, int maxam);
size_t simint_eri_workmem(int derorder, int maxam);

#ifdef __cplusplus
}
#endif

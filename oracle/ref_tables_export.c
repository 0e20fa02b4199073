/* Exports the reference's long-range Boys factors (a static table in a
 * header) next to the two table sources compiled alongside, so the oracle's
 * own tables can be compared with the reference's.  Test infrastructure. */
#include "simint/boys/boys_longfac.h"

const double *simint_ref_boys_longfac(void) { return boys_longfac; }
int simint_ref_boys_longfac_n(void) { return BOYS_LONGFAC_MAXN + 1; }

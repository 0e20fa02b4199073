/* Shell-pair screening (skel/simint/shell/shell_screen.h:9-58).
 * Values are (ab|ab)-type bounds without the square root, compared by the
 * compute entry points against screen_tol^2.  The Schwarz variants compute
 * the (ab|ab) integrals on the GPU. */
#pragma once

#include "simint/shell/shell.h"

#ifdef __cplusplus
#include "simint/cpp_restrict.hpp"
extern "C" {
#endif

double simint_shellscreen(struct simint_shell const * A,
                          struct simint_shell const * B,
                          int screen_method);

double simint_primscreen(struct simint_shell const * A,
                         struct simint_shell const * B,
                         double * out,
                         int screen_method);

double simint_shellscreen_schwarz(struct simint_shell const * A,
                                  struct simint_shell const * B);

double simint_shellscreen_fastschwarz(struct simint_shell const * A,
                                      struct simint_shell const * B);

double simint_primscreen_schwarz(struct simint_shell const * A,
                                 struct simint_shell const * B,
                                 double * out);

double simint_primscreen_fastschwarz(struct simint_shell const * A,
                                     struct simint_shell const * B,
                                     double * out);

#ifdef __cplusplus
}
#endif

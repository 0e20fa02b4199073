/* simint drop-in C API for the MI355X ERI engine.
 *
 * Same public surface as simint's convenience header
 * (skel/simint/simint.h:1-7): library init/finalize, the shell and
 * shell-pair API, the one-electron entry points and the
 * two-electron-integral entry points. */
#pragma once

#include "simint/ostei/ostei_config.h"
#include "simint/simint_init.h"
#include "simint/simint_oneelectron.h"
#include "simint/simint_eri.h"

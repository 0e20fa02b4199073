/* simint drop-in C API for the MI355X ERI engine.
 *
 * Same public surface as simint's convenience header
 * (skel/simint/simint.h:1-7): library init/finalize, the shell and
 * shell-pair API, and the two-electron-integral entry points.  The
 * one-electron entry points (skel/simint/simint_oneelectron.h) are not part
 * of this library. */
#pragma once

#include "simint/ostei/ostei_config.h"
#include "simint/simint_init.h"
#include "simint/simint_eri.h"

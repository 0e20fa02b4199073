// Class dispatch: (la lb | lc ld) -> kernel launcher.  The tables are
// generated (gen/dispatch.cpp): one lane-per-quartet and one wave-per-quartet
// launcher per class where built, and the family in use for each class.
#pragma once
#include <hip/hip_runtime.h>
#include "device_types.h"

typedef hipError_t (*simint_launch_fn)(const EriArgs *, hipStream_t);

typedef hipError_t (*simint_list_launch_fn)(const EriListArgs *, hipStream_t);

simint_launch_fn simint_dispatch(int la, int lb, int lc, int ld);
// staging doubles per lane the selected kernel needs in EriArgs::scratch
// (split lane-per-quartet family), 0 for the others
int simint_dispatch_staging(int la, int lb, int lc, int ld);
// quartet-list launcher of the class's selected family (or the other family
// when the selected one has no list kernel); null if the class is not built
simint_list_launch_fn simint_dispatch_list(int la, int lb, int lc, int ld);
// FP32-VRR/HRR prefix kernel of the class (lane-per-quartet family only);
// null where the class has none or runs the wave-per-quartet family
simint_launch_fn simint_dispatch_mixed(int la, int lb, int lc, int ld);

/* Library initialisation (skel/simint/simint_init.h:21,37).
 *
 * simint_init() selects and initialises the GPU (HIP device 0, or the device
 * named by SIMINT_DEVICE), uploads the Boys-function table and builds the
 * class dispatch table.  It may be called several times; calls after the
 * first do nothing.  Not thread-safe, like the reference.  If no GPU is
 * available the compute entry points return SIMINT_ERROR_NO_DEVICE. */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

void simint_init(void);
void simint_finalize(void);

#ifdef __cplusplus
}
#endif

// Services of the host runtime (runtime.hip) shared with the other host
// sources of the library: device set-up, the calling thread's stream and
// per-thread staging buffers, and the thread's last-error text.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace simint_rt {

// set up the device once (Boys grid, derivative tables); false if there is
// no usable GPU
bool ensure_device();
// the device copy of the 366 x 32 Boys grid
const double *boys_grid();
// the stream the calling thread's calls queue on
hipStream_t stream();
// Wait for the calling thread's earlier work, then make sure its pinned host
// staging buffer holds `host_bytes` and its device buffers `in_bytes` and
// `out_bytes`; returns their addresses (valid until the thread's next call).
hipError_t thread_buffers(size_t host_bytes, size_t in_bytes, size_t out_bytes, void **host, void **din,
                          void **dout);
void set_error(const char *msg);
void set_hip_error(const char *what, hipError_t e);

}  // namespace simint_rt

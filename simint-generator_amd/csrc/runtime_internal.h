// Services of the host runtime (runtime.hip) shared with the other host
// sources of the library: device set-up, the calling thread's stream and
// per-thread staging buffers, and the thread's last-error text.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "device_types.h"
#include "simint/simint_hip.h"

struct simint_multi_shellpair;

// Shell pairs of one (am1, am2) resident in device memory: the linear and
// tiled views the class kernels read, and the 2a / 2b weighted prefactors of
// the derivative terms.
struct simint_hip_pairs {
    int am1, am2, n;
    bool has_screen;
    double screen_max;
    void *dev = nullptr;
    PairsLinear lin;
    PairsTiled til;
    int ntile;
    size_t bytes = 0;                            // device memory held
    const double *wa = nullptr, *wb = nullptr;   // linear prefac * alpha2, prefac * beta2
    const double *twa = nullptr, *twb = nullptr; // the same in the tiled view
};

// A shell set resident in device memory (one-electron integrals, device
// pair construction): per shell AM, primitive count, first primitive and
// centre; per primitive exponent and contraction coefficient.
struct simint_hip_shells {
    int n = 0;
    int maxam = 0;
    void *dev = nullptr;
    const int *am = nullptr, *nprim = nullptr, *poff = nullptr;
    const double *xyz = nullptr;     // 3 per shell
    const double *alpha = nullptr, *coef = nullptr;
    int *host_am = nullptr, *host_nprim = nullptr;   // host copies (owned)
};

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
// Move one (la|lb) block of one-electron integrals from Psi4 order into the
// drop-in calls' current Cartesian order (out must not overlap in); *moved
// tells whether anything was written to out.
hipError_t to_drop_in_order(int la, int lb, const double *in, double *out, hipStream_t s, bool *moved);
void set_error(const char *msg);
void set_hip_error(const char *what, hipError_t e);
// Largest |(bb|bb)| diagonal element of every shell pair of a host pair set,
// into out (host, one per pair); 0 or a negative SIMINT_ERROR_*.  Declared
// again in shell_screen.cpp, which is compiled as plain C++.
int host_pair_bounds(const simint_multi_shellpair *P, double *out);
// jk.hip: queue the J/K digestion of the integrals the class kernel of class
// am wrote for `items` (ket tile lanes mapped to pairs by col; the items of
// bra i of the launch start at tstart[i], device, nbra entries) on stream s
hipError_t launch_jk_digest(const int am[4], const WorkItem *items, long long nitems, const long long *tstart,
                            int nbra, const long long *col, const double *ints, const simint_hip_jk *jk,
                            hipStream_t s);

}  // namespace simint_rt

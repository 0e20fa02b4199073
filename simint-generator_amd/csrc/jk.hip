// Coulomb and exchange digestion of shell-quartet integrals (the consumer a
// Fock build puts after the integral kernels; the reference returns the
// integrals only).
//
// A unique shell quartet (AB|CD) stands for its images under the eight
// index permutations.  Its stabiliser has s = s_AB s_CD s_(AB)(CD) elements
// (each factor 2 when the two sides are the same shell / pair), so summing
// all eight images with weight f = 1/s counts every ordered quartet of
// functions once.  For a symmetric density the eight images collapse to
//     J'_ab += 4 f v D_cd      J'_cd += 4 f v D_ab
//     K'_ac += 2 f v D_bd      K'_ad += 2 f v D_bc
//     K'_bc += 2 f v D_ad      K'_bd += 2 f v D_ac
// and J = (J' + J'^T)/2, K = (K' + K'^T)/2 (jk_symmetrize).
//
// One wavefront digests one work item of the class launch (a bra pair
// against up to 64 kets of a tile, the same items the class kernel ran):
// each lane reads its quartet's integrals where the kernel wrote them.  The
// bra-side J sums are reduced across the wave before one atomic add per
// element; the ket-side J and the four exchange blocks are added per lane.
#include <hip/hip_runtime.h>

#include "simint/simint_eri.h"
#include "simint/simint_hip.h"
#include "device_types.h"
#include "runtime_internal.h"

namespace {

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct JkDev {
    int nbf, same;
    const int *bra_fun, *ket_fun;
    const double *D;
    double *J, *K;
};

__global__ void __launch_bounds__(256) jk_digest(const WorkItem *items, long long nitems, const long long *col,
                                                 const double *ints, int na, int nb, int nc, int nd, JkDev jk)
{
    const long long w = (long long)blockIdx.x * 4 + threadIdx.x / 64;
    if (w >= nitems) return;   // uniform over the wave
    const int lane = threadIdx.x & 63;
    const WorkItem it = items[w];
    const int n = jk.nbf;
    const int fa = jk.bra_fun[2 * it.bra], fb = jk.bra_fun[2 * it.bra + 1];
    if (fa < 0 || fb < 0 || fa + na > n || fb + nb > n) return;
    bool act = lane >= it.lane_lo && lane < it.lane_hi;
    const long long k = act ? col[(long long)it.tile * 64 + lane] : -1;
    act = act && k >= 0;
    int fc = 0, fd = 0;
    if (act) {
        fc = jk.ket_fun[2 * k];
        fd = jk.ket_fun[2 * k + 1];
        act = fc >= 0 && fd >= 0 && fc + nc <= n && fd + nd <= n;
    }
    const int ncd = nc * nd;
    const long long stride = (long long)na * nb * ncd;
    const double *q = ints + it.obase + (act ? k : 0) * stride;
    double f = act ? 1.0 : 0.0;
    if (fa == fb) f *= 0.5;
    if (fc == fd) f *= 0.5;
    if (jk.same && k == it.bra) f *= 0.5;
    const double *D = jk.D;
    if (jk.J) {
        for (int a = 0; a < na; a++)
            for (int b = 0; b < nb; b++) {
                double s = 0.0;
                if (act) {
                    const double *r = q + (a * nb + b) * ncd;
                    for (int c = 0; c < nc; c++)
                        for (int d = 0; d < nd; d++) s += r[c * nd + d] * D[(long long)(fc + c) * n + fd + d];
                }
                s = wave_sum(f * s);
                if (lane == 0) unsafeAtomicAdd(jk.J + (long long)(fa + a) * n + fb + b, 4.0 * s);
            }
        if (act)
            for (int c = 0; c < nc; c++)
                for (int d = 0; d < nd; d++) {
                    double t = 0.0;
                    for (int a = 0; a < na; a++)
                        for (int b = 0; b < nb; b++)
                            t += q[(a * nb + b) * ncd + c * nd + d] * D[(long long)(fa + a) * n + fb + b];
                    unsafeAtomicAdd(jk.J + (long long)(fc + c) * n + fd + d, 4.0 * f * t);
                }
    }
    if (jk.K && act) {
        const double g = 2.0 * f;
        for (int a = 0; a < na; a++)
            for (int c = 0; c < nc; c++) {   // K_ac += v D_bd
                double t = 0.0;
                for (int b = 0; b < nb; b++)
                    for (int d = 0; d < nd; d++)
                        t += q[(a * nb + b) * ncd + c * nd + d] * D[(long long)(fb + b) * n + fd + d];
                unsafeAtomicAdd(jk.K + (long long)(fa + a) * n + fc + c, g * t);
            }
        for (int a = 0; a < na; a++)
            for (int d = 0; d < nd; d++) {   // K_ad += v D_bc
                double t = 0.0;
                for (int b = 0; b < nb; b++)
                    for (int c = 0; c < nc; c++)
                        t += q[(a * nb + b) * ncd + c * nd + d] * D[(long long)(fb + b) * n + fc + c];
                unsafeAtomicAdd(jk.K + (long long)(fa + a) * n + fd + d, g * t);
            }
        for (int b = 0; b < nb; b++)
            for (int c = 0; c < nc; c++) {   // K_bc += v D_ad
                double t = 0.0;
                for (int a = 0; a < na; a++)
                    for (int d = 0; d < nd; d++)
                        t += q[(a * nb + b) * ncd + c * nd + d] * D[(long long)(fa + a) * n + fd + d];
                unsafeAtomicAdd(jk.K + (long long)(fb + b) * n + fc + c, g * t);
            }
        for (int b = 0; b < nb; b++)
            for (int d = 0; d < nd; d++) {   // K_bd += v D_ac
                double t = 0.0;
                for (int a = 0; a < na; a++)
                    for (int c = 0; c < nc; c++)
                        t += q[(a * nb + b) * ncd + c * nd + d] * D[(long long)(fa + a) * n + fc + c];
                unsafeAtomicAdd(jk.K + (long long)(fb + b) * n + fd + d, g * t);
            }
    }
}

// M <- (M + M^T)/2: thread (i, j) with i < j averages the pair
__global__ void __launch_bounds__(256) symmetrize(int n, double *M)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    if (i >= j) return;
    const double v = 0.5 * (M[(long long)i * n + j] + M[(long long)j * n + i]);
    M[(long long)i * n + j] = v;
    M[(long long)j * n + i] = v;
}

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

}  // namespace

namespace simint_rt {

hipError_t launch_jk_digest(const int am[4], const WorkItem *items, long long nitems, const long long *col,
                            const double *ints, const simint_hip_jk *jk, hipStream_t s)
{
    JkDev d{jk->nbf, jk->same, jk->bra_fun, jk->ket_fun, jk->density, jk->J, jk->K};
    const long long nblk = (nitems + 3) / 4;
    hipLaunchKernelGGL(jk_digest, dim3((unsigned)nblk), dim3(256), 0, s, items, nitems, col, ints, ncart(am[0]),
                       ncart(am[1]), ncart(am[2]), ncart(am[3]), d);
    return hipGetLastError();
}

}  // namespace simint_rt

extern "C" {

int simint_hip_jk_symmetrize(int nbf, double *d_J, double *d_K)
{
    if (nbf <= 0) return SIMINT_ERROR_ARGUMENT;
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const hipStream_t s = simint_rt::stream();
    const long long nn = (long long)nbf * nbf;
    const unsigned nblk = (unsigned)((nn + 255) / 256);
    for (double *M : {d_J, d_K}) {
        if (!M) continue;
        hipLaunchKernelGGL(symmetrize, dim3(nblk), dim3(256), 0, s, nbf, M);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) { simint_rt::set_hip_error("symmetrize launch", e); return SIMINT_ERROR_HIP; }
    }
    return 0;
}

}  // extern "C"

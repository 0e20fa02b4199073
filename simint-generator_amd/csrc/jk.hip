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
// The exchange blocks whose row is a bra function (K_ac, K_ad) are summed
// over b (and K_bc, K_bd over a) in registers before their atomic adds, and
// where the bra's rows fit the local data share one block takes all of a
// bra's items and sums those rows over all its kets there first
// (jk_digest_rows).
#include <hip/hip_runtime.h>

#include <atomic>

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

// The same digestion with the ket shells' and the second bra shell's
// component counts fixed at compile time: each lane reads every row
// (a, b, :, :) of its quartet once and applies all six updates from
// registers (the density values it needs are loaded once per item, per a or
// per row), instead of one pass over the quartet per update.  The second
// ket shell's components are taken DC at a time (DC divides ND, chosen so
// the accumulators stay in registers), each chunk reading its columns of
// every row.  rows != nullptr (ROWS): the exchange updates whose row is a
// bra function (K_a., K_b.) go to that block's row accumulator in the local
// data share -- row a of the first bra shell at rows[a * nbf], row b of the
// second at rows[(na + b) * nbf] -- and K_b. is added there after every
// row (a, b); without it they go to global memory, K_b. summed over a first.
template <int NB, int NC, int ND, int DC, bool ROWS>
__device__ __forceinline__ void digest_item(const WorkItem &it, int lane, const long long *col, const double *ints,
                                            int na, const JkDev &jk, double *rows)
{
    static_assert(ND % DC == 0, "chunks of the second ket shell divide it");
    constexpr int NCD = NC * ND;
    constexpr int KB = ROWS ? 1 : NB;   // K_b. accumulators held across a
    const int n = jk.nbf;
    const int fa = jk.bra_fun[2 * it.bra], fb = jk.bra_fun[2 * it.bra + 1];
    if (fa < 0 || fb < 0 || fa + na > n || fb + NB > n) return;
    bool act = lane >= it.lane_lo && lane < it.lane_hi;
    const long long k = act ? col[(long long)it.tile * 64 + lane] : -1;
    act = act && k >= 0;
    int fc = 0, fd = 0;
    if (act) {
        fc = jk.ket_fun[2 * k];
        fd = jk.ket_fun[2 * k + 1];
        act = fc >= 0 && fd >= 0 && fc + NC <= n && fd + ND <= n;
    }
    // idle lanes read a quartet of the item (its first lane's, which the
    // class kernel wrote) and density element 0, and add nothing (weight 0)
    const long long kq = act ? k : col[(long long)it.tile * 64 + it.lane_lo];
    const double *q = ints + it.obase + kq * (long long)(na * NB * NCD);
    const double *__restrict__ D = jk.D;
    auto Dl = [&](long long idx) { return D[act ? idx : 0]; };
    // exchange row update: row r of the bra's (na + NB) rows, column c
    auto addK = [&](int r, int c, double v) {
        if constexpr (ROWS) {
            // the local data share, named as such so the add is a ds_add_f64
            __attribute__((address_space(3))) double *lds = (__attribute__((address_space(3))) double *)rows;
            atomicAdd((double *)(lds + (long long)r * n + c), v);
        } else {
            unsafeAtomicAdd(jk.K + (long long)(r < na ? fa + r : fb + r - na) * n + c, v);
        }
    };
    double f = act ? 1.0 : 0.0;
    if (fa == fb) f *= 0.5;
    if (fc == fd) f *= 0.5;
    if (jk.same && k == it.bra) f *= 0.5;
    const double g = 2.0 * f;
    for (int d0 = 0; d0 < ND; d0 += DC) {
        double dcd[NC * DC], jcd[NC * DC], kbc[KB][NC], kbd[KB][DC];
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int d = 0; d < DC; d++) {
                dcd[c * DC + d] = Dl((long long)(fc + c) * n + fd + d0 + d);
                jcd[c * DC + d] = 0.0;
            }
#pragma unroll
        for (int b = 0; b < KB; b++) {
#pragma unroll
            for (int c = 0; c < NC; c++) kbc[b][c] = 0.0;
#pragma unroll
            for (int d = 0; d < DC; d++) kbd[b][d] = 0.0;
        }
        for (int a = 0; a < na; a++) {
            double dad[DC], dac[NC], kac[NC], kad[DC];
#pragma unroll
            for (int d = 0; d < DC; d++) {
                dad[d] = Dl((long long)(fa + a) * n + fd + d0 + d);
                kad[d] = 0.0;
            }
#pragma unroll
            for (int c = 0; c < NC; c++) {
                dac[c] = Dl((long long)(fa + a) * n + fc + c);
                kac[c] = 0.0;
            }
            // one row (a, b, :, d0..d0+DC) of the quartet
            auto row = [&](int b) {
                const int kb = ROWS ? 0 : b;
                if constexpr (ROWS) {
#pragma unroll
                    for (int c = 0; c < NC; c++) kbc[0][c] = 0.0;
#pragma unroll
                    for (int d = 0; d < DC; d++) kbd[0][d] = 0.0;
                }
                double dbd[DC], dbc[NC];
#pragma unroll
                for (int d = 0; d < DC; d++) dbd[d] = Dl((long long)(fb + b) * n + fd + d0 + d);
#pragma unroll
                for (int c = 0; c < NC; c++) dbc[c] = Dl((long long)(fb + b) * n + fc + c);
                const double *r = q + (a * NB + b) * NCD + d0;
                const double dab = D[(long long)(fa + a) * n + fb + b];
                double jab = 0.0;
#pragma unroll
                for (int c = 0; c < NC; c++)
#pragma unroll
                    for (int d = 0; d < DC; d++) {
                        const double v = act ? r[c * ND + d] : 0.0;   // r is readable either way
                        jab = fma(v, dcd[c * DC + d], jab);
                        jcd[c * DC + d] = fma(v, dab, jcd[c * DC + d]);
                        kac[c] = fma(v, dbd[d], kac[c]);
                        kad[d] = fma(v, dbc[c], kad[d]);
                        kbc[kb][c] = fma(v, dad[d], kbc[kb][c]);
                        kbd[kb][d] = fma(v, dac[c], kbd[kb][d]);
                    }
                if (jk.J) {
                    jab = wave_sum(f * jab);
                    if (lane == 0) unsafeAtomicAdd(jk.J + (long long)(fa + a) * n + fb + b, 4.0 * jab);
                }
                if constexpr (ROWS) {
                    if (jk.K && act) {
#pragma unroll
                        for (int c = 0; c < NC; c++) addK(na + b, fc + c, g * kbc[0][c]);
#pragma unroll
                        for (int d = 0; d < DC; d++) addK(na + b, fd + d0 + d, g * kbd[0][d]);
                    }
                }
            };
            if constexpr (ROWS) {
                // nothing is indexed by b: a loop, so one row's loads are in
                // flight at a time
#pragma unroll 1
                for (int b = 0; b < NB; b++) row(b);
            } else {
#pragma unroll
                for (int b = 0; b < NB; b++) row(b);
            }
            if (jk.K && act) {
#pragma unroll
                for (int c = 0; c < NC; c++) addK(a, fc + c, g * kac[c]);
#pragma unroll
                for (int d = 0; d < DC; d++) addK(a, fd + d0 + d, g * kad[d]);
            }
        }
        if (act && jk.J) {
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int d = 0; d < DC; d++)
                    unsafeAtomicAdd(jk.J + (long long)(fc + c) * n + fd + d0 + d, 4.0 * f * jcd[c * DC + d]);
        }
        if constexpr (!ROWS) {
            if (act && jk.K) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
#pragma unroll
                    for (int c = 0; c < NC; c++) addK(na + b, fc + c, g * kbc[b][c]);
#pragma unroll
                    for (int d = 0; d < DC; d++) addK(na + b, fd + d0 + d, g * kbd[b][d]);
                }
            }
        }
    }
}

// one wavefront per work item
template <int NB, int NC, int ND, int DC>
__global__ void __launch_bounds__(256) jk_digest_t(const WorkItem *items, long long nitems, const long long *col,
                                                   const double *ints, int na, JkDev jk)
{
    const long long w = (long long)blockIdx.x * 4 + threadIdx.x / 64;
    if (w >= nitems) return;   // uniform over the wave
    digest_item<NB, NC, ND, DC, false>(items[w], threadIdx.x & 63, col, ints, na, jk, nullptr);
}

// One block per bra pair of the launch: its four wavefronts take the bra's
// work items in turn, and the exchange rows of the bra's functions are
// summed over all its kets in the local data share ((na + NB) x nbf
// doubles, dynamic) and added to K once per touched element at the end --
// never more global atomics than one per quartet and element, usually far
// fewer (every ket of the bra adds to the same rows).  Launched only with
// an exchange matrix to fill.
template <int NB, int NC, int ND, int DC>
__global__ void __launch_bounds__(256) jk_digest_rows(const WorkItem *items, long long nitems,
                                                      const long long *tstart, int nbra, const long long *col,
                                                      const double *ints, int na, JkDev jk)
{
    extern __shared__ double rows[];
    const int bi = blockIdx.x;
    const long long i0 = tstart[bi], i1 = bi + 1 < nbra ? tstart[bi + 1] : nitems;
    if (i0 >= i1) return;   // uniform over the block
    const int n = jk.nbf;
    const int R = na + NB;
    for (long long e = threadIdx.x; e < (long long)R * n; e += 256) rows[e] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (long long w = i0 + threadIdx.x / 64; w < i1; w += 4)
        digest_item<NB, NC, ND, DC, true>(items[w], lane, col, ints, na, jk, rows);
    __syncthreads();
    const int bra = items[i0].bra;
    const int fa = jk.bra_fun[2 * bra], fb = jk.bra_fun[2 * bra + 1];
    if (fa < 0 || fb < 0 || fa + na > n || fb + NB > n) return;
    for (long long e = threadIdx.x; e < (long long)R * n; e += 256) {
        const double v = rows[e];
        if (v == 0.0) continue;
        const int r = (int)(e / n), c = (int)(e - (long long)r * n);
        unsafeAtomicAdd(jk.K + (long long)(r < na ? fa + r : fb + r - na) * n + c, v);
    }
}

constexpr int nc_of(int l) { return (l + 1) * (l + 2) / 2; }

// doubles per lane the specialised kernel holds for a chunk of dc
// components of the second ket shell: density and Coulomb sums of the ket
// block, the per-a and per-row vectors, and without the row accumulator the
// K_b. sums over a plus (a quarter of) the unrolled rows' loads in flight.
// Kept within kRegDoubles (gfx950 has 256 VGPRs per lane; the rest hold
// addresses and loads): every instance then runs without scratch.
constexpr int kRegDoubles = 96;
constexpr int reg_doubles(int nb, int nc, int dc, bool rows)
{
    return 2 * nc * dc + 4 * (nc + dc) + (rows ? 0 : nb * (nc + dc) + nb * nc * dc / 4);
}

// components of the second ket shell per chunk: the largest divisor of nd
// that fits; 0: none does (the generic kernel)
constexpr int pick_dc(int nb, int nc, int nd, bool rows)
{
    for (int dc = nd; dc >= 1; --dc)
        if (nd % dc == 0 && reg_doubles(nb, nc, dc, rows) <= kRegDoubles) return dc;
    return 0;
}

// exchange row accumulators of the per-bra kernel: at most this many bytes
// of the local data share per block (and no more than the device allows,
// 160 KB on gfx950).  A block above 80 KB holds its CU's share alone
// whatever its size, so the whole of it is offered.
constexpr long long kRowBytes = 160 * 1024;

std::atomic<int> g_rows_enabled{1};   // simint_hip_set_jk_rows

long long row_bytes_limit()
{
    static std::atomic<long long> lim{-1};
    long long v = lim.load(std::memory_order_relaxed);
    if (v < 0) {
        int dev = 0, a = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            a = 65536;
        v = a < kRowBytes ? a : kRowBytes;
        lim.store(v, std::memory_order_relaxed);
    }
    return g_rows_enabled.load(std::memory_order_relaxed) ? v : 0;
}

struct Launch {
    const WorkItem *items;
    long long nitems;
    const long long *tstart;
    int nbra;
    const long long *col;
    const double *ints;
    int na;
};

template <int LB, int LC, int LD>
hipError_t launch_t(const Launch &L, const JkDev &d, hipStream_t s)
{
    constexpr int NB = nc_of(LB), NC = nc_of(LC), ND = nc_of(LD);
    const long long rbytes = 8LL * (L.na + NB) * d.nbf;
    constexpr int DCR = pick_dc(NB, NC, ND, true), DCG = pick_dc(NB, NC, ND, false);
    const WorkItem *items = L.items;
    const long long nitems = L.nitems;
    const long long *col = L.col;
    const double *ints = L.ints;
    const int na = L.na;
    const unsigned nwave_blocks = (unsigned)((nitems + 3) / 4);
    bool done = false;
    if constexpr (DCR > 0) {
        if (d.K && rbytes <= row_bytes_limit()) {
            hipLaunchKernelGGL((jk_digest_rows<NB, NC, ND, DCR>), dim3((unsigned)L.nbra), dim3(256), (size_t)rbytes,
                               s, items, nitems, L.tstart, L.nbra, col, ints, na, d);
            done = true;
        }
    }
    if constexpr (DCG > 0) {
        if (!done) {
            hipLaunchKernelGGL((jk_digest_t<NB, NC, ND, DCG>), dim3(nwave_blocks), dim3(256), 0, s, items, nitems,
                               col, ints, na, d);
            done = true;
        }
    }
    if (!done)
        hipLaunchKernelGGL(jk_digest, dim3(nwave_blocks), dim3(256), 0, s, items, nitems, col, ints, na, NB, NC, ND,
                           d);
    return hipGetLastError();
}

template <int LB, int LC>
hipError_t launch_d(int ld, const Launch &L, const JkDev &d, hipStream_t s)
{
    switch (ld) {
    case 0: return launch_t<LB, LC, 0>(L, d, s);
    case 1: return launch_t<LB, LC, 1>(L, d, s);
    case 2: return launch_t<LB, LC, 2>(L, d, s);
    default: return launch_t<LB, LC, 3>(L, d, s);
    }
}

template <int LB>
hipError_t launch_c(int lc, int ld, const Launch &L, const JkDev &d, hipStream_t s)
{
    switch (lc) {
    case 0: return launch_d<LB, 0>(ld, L, d, s);
    case 1: return launch_d<LB, 1>(ld, L, d, s);
    case 2: return launch_d<LB, 2>(ld, L, d, s);
    default: return launch_d<LB, 3>(ld, L, d, s);
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

hipError_t launch_jk_digest(const int am[4], const WorkItem *items, long long nitems, const long long *tstart,
                            int nbra, const long long *col, const double *ints, const simint_hip_jk *jk,
                            hipStream_t s)
{
    JkDev d{jk->nbf, jk->same, jk->bra_fun, jk->ket_fun, jk->density, jk->J, jk->K};
    if (am[1] > 3 || am[2] > 3 || am[3] > 3) {
        hipLaunchKernelGGL(jk_digest, dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, s, items, nitems, col, ints,
                           ncart(am[0]), ncart(am[1]), ncart(am[2]), ncart(am[3]), d);
        return hipGetLastError();
    }
    const Launch L{items, nitems, tstart, nbra, col, ints, ncart(am[0])};
    switch (am[1]) {
    case 0: return launch_c<0>(am[2], am[3], L, d, s);
    case 1: return launch_c<1>(am[2], am[3], L, d, s);
    case 2: return launch_c<2>(am[2], am[3], L, d, s);
    default: return launch_c<3>(am[2], am[3], L, d, s);
    }
}

}  // namespace simint_rt

extern "C" {

int simint_hip_set_jk_rows(int enable)
{
    return g_rows_enabled.exchange(enable ? 1 : 0);
}

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

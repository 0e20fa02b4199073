// Device-side front end of a screened quartet set: shell pairs filled on the
// GPU from a device shell set (the Gaussian-product quantities of
// skel/simint/shell/shell.c:425-582, the fast-Schwarz values of
// skel/simint/shell/shell_screen.c:226-270 and the descending sort by them,
// shell.c:37-71), the Schwarz bound of every pair from its diagonal quartets,
// the order of a pair set by (primitive count, bound), and the number of kets
// of a bound-sorted group that survive the shell-quartet screen for every
// bra.  A caller whose shells already sit in HBM sets up its quartet lists
// without host-side pair data; only the per-bra counts come back.
//
// The fill runs the host fill's arithmetic in the same order with
// multiply-add contraction off, so every field but the ones behind exp() /
// pow() (prefactor, screening value) is bit-identical to the host fill, and
// those agree to the last bit or two.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "simint/constants.h"
#include "simint/simint.h"
#include "simint/simint_hip.h"
#include "dispatch.h"
#include "runtime_internal.h"

namespace {

constexpr int kMaxPairPrim = 1024;   // primitive pairs of one shell pair the device sort handles

struct FillArgs {
    const int *am, *nprim, *poff;
    const double *xyz, *alpha, *coef;
    const int *pairs;      // 2 per pair
    int npair;
    int screen;            // 0: none, 2: fast Schwarz
    const int *ppoff;      // first primitive pair of each pair
    // linear view
    double *alpha_, *inv, *prefac, *x, *y, *z, *pax, *pay, *paz, *pbx, *pby, *pbz, *scr, *a2, *b2, *wa, *wb;
    int *np;
    double *abx, *aby, *abz;
};

__device__ void swap_prim(const FillArgs &f, int i, int j)
{
    double *fld[14] = {f.x, f.y, f.z, f.pax, f.pay, f.paz, f.pbx, f.pby, f.pbz, f.alpha_, f.prefac, f.scr, f.a2, f.b2};
#pragma unroll
    for (int k = 0; k < 14; k++) {
        const double t = fld[k][i];
        fld[k][i] = fld[k][j];
        fld[k][j] = t;
    }
}

// The host's descending quicksort on the screening value (the same
// partitioning, so ties end in the same order), iteratively over a small
// stack of ranges.
__device__ void sort_desc(const FillArgs &f, int left, int right)
{
    int stack[64];
    int sp = 0;
    stack[sp++] = left;
    stack[sp++] = right;
    while (sp > 0) {
        const int r = stack[--sp], l = stack[--sp];
        int i = l, j = r;
        const double mid = f.scr[(i + j) / 2];
        while (i <= j) {
            while (f.scr[i] > mid) i++;
            while (f.scr[j] < mid) j--;
            if (i <= j) {
                swap_prim(f, i, j);
                i++;
                j--;
            }
        }
        // the two parts are disjoint, so the order they are sorted in does
        // not change the result: push the larger first so the stack stays
        // within 2 log2(n) + 2 entries (n <= kMaxPairPrim)
        const bool rr = i < r, ll = j > l;
        if (rr && ll && (r - i) < (j - l)) {
            stack[sp++] = l; stack[sp++] = j;
            stack[sp++] = i; stack[sp++] = r;
        } else {
            if (rr) { stack[sp++] = i; stack[sp++] = r; }
            if (ll) { stack[sp++] = l; stack[sp++] = j; }
        }
    }
}

// one thread per shell pair (shell.c:456-576 with SIMINT_SIMD_LEN 1)
__global__ void __launch_bounds__(128) pair_fill(const FillArgs f)
{
#pragma clang fp contract(off)
    const int ab = blockIdx.x * 128 + threadIdx.x;
    if (ab >= f.npair) return;
    const int A = f.pairs[2 * ab], B = f.pairs[2 * ab + 1];
    const int same = A == B;
    const double Ax = f.xyz[3 * A], Ay = f.xyz[3 * A + 1], Az = f.xyz[3 * A + 2];
    const double Bx = f.xyz[3 * B], By = f.xyz[3 * B + 1], Bz = f.xyz[3 * B + 2];
    const double Xab_x = Ax - Bx, Xab_y = Ay - By, Xab_z = Az - Bz;
    const double Xab = Xab_x * Xab_x + Xab_y * Xab_y + Xab_z * Xab_z;
    const int na = f.nprim[A], nb = f.nprim[B];
    const double *aa = f.alpha + f.poff[A], *ab_ = f.alpha + f.poff[B];
    const double *ca = f.coef + f.poff[A], *cb = f.coef + f.poff[B];
    int idx = f.ppoff[ab];
    const int first = idx;
    for (int i = 0; i < na; ++i) {
        const double ai = aa[i];
        const double AxAa = Ax * ai, AyAa = Ay * ai, AzAa = Az * ai;
        const int jend = same ? i + 1 : nb;
        for (int j = 0; j < jend; ++j) {
            const double aj = ab_[j];
            const double ab_sum = ai + aj;
            const double ab_mul = ai * aj;
            double pf = ca[i] * cb[j] * exp(-Xab * ab_mul / ab_sum) * SQRT_TWO_PI_52 / ab_sum;
            if (same && i != j) pf *= 2.0;
            f.prefac[idx] = pf;
            const double px = (AxAa + aj * Bx) / ab_sum;
            const double py = (AyAa + aj * By) / ab_sum;
            const double pz = (AzAa + aj * Bz) / ab_sum;
            f.x[idx] = px;
            f.y[idx] = py;
            f.z[idx] = pz;
            f.pax[idx] = px - Ax;
            f.pay[idx] = py - Ay;
            f.paz[idx] = pz - Az;
            f.pbx[idx] = px - Bx;
            f.pby[idx] = py - By;
            f.pbz[idx] = pz - Bz;
            f.alpha_[idx] = ab_sum;
            if (same && i != j) {
                f.a2[idx] = ai + aj;
                f.b2[idx] = ai + aj;
            } else {
                f.a2[idx] = 2.0 * ai;
                f.b2[idx] = 2.0 * aj;
            }
            if (f.screen == SIMINT_SCREEN_FASTSCHWARZ) {
                // the s-type (ij|ij) of the primitive pair (shell_screen.c:226-270)
                const double oop = 1.0 / ab_sum;
                const double rho = ab_mul / ab_sum;
                double G = SQRT_TWO_TIMES_PI_52 * pow(oop, 2.5) * exp(-2.0 * rho * Xab);
                G *= ca[i] * ca[i] * cb[j] * cb[j];
                f.scr[idx] = G;
            } else {
                f.scr[idx] = 0.0;
            }
            ++idx;
        }
    }
    if (f.screen && idx - first > 1) sort_desc(f, first, idx - 1);
    for (int k = first; k < idx; ++k) {
        f.inv[k] = 1.0 / f.alpha_[k];
        f.wa[k] = f.prefac[k] * f.a2[k];
        f.wb[k] = f.prefac[k] * f.b2[k];
    }
    f.np[ab] = idx - first;
    f.abx[ab] = Xab_x;
    f.aby[ab] = Xab_y;
    f.abz[ab] = Xab_z;
}

struct TileArgs {
    PairsLinear lin;
    const double *lwa, *lwb;
    int npair, ntile;
    const long long *toff;
    const int *tnp;
    double *alpha, *inv, *prefac, *x, *y, *z, *pax, *pay, *paz, *pbx, *pby, *pbz, *scr, *a2, *b2, *wa, *wb;
    double *abx, *aby, *abz;
    long long *col;
};

// the tiled view: primitive j of lane l of tile t at toff[t] + 64 j + l;
// padding rows carry alpha = 1 (and 1/alpha, 2a, 2b = 1) and zeros, so they
// contribute nothing
__global__ void __launch_bounds__(64) tile_fill(const TileArgs a)
{
    const int t = blockIdx.x, l = threadIdx.x;
    const int p = t * SIMINT_TILE + l;
    const bool live = p < a.npair;
    const long long tl = (long long)t * SIMINT_TILE + l;
    a.col[tl] = live ? p : -1;
    a.abx[tl] = live ? a.lin.abx[p] : 0.0;
    a.aby[tl] = live ? a.lin.aby[p] : 0.0;
    a.abz[tl] = live ? a.lin.abz[p] : 0.0;
    const int n = live ? a.lin.np[p] : 0;
    const int o = live ? a.lin.poff[p] : 0;
    for (int j = 0; j < a.tnp[t]; ++j) {
        const long long k = a.toff[t] + (long long)j * SIMINT_TILE + l;
        if (j < n) {
            const int s = o + j;
            a.alpha[k] = a.lin.alpha[s];
            a.inv[k] = a.lin.inv_alpha[s];
            a.prefac[k] = a.lin.prefac[s];
            a.x[k] = a.lin.x[s];
            a.y[k] = a.lin.y[s];
            a.z[k] = a.lin.z[s];
            a.pax[k] = a.lin.pax[s];
            a.pay[k] = a.lin.pay[s];
            a.paz[k] = a.lin.paz[s];
            a.pbx[k] = a.lin.pbx[s];
            a.pby[k] = a.lin.pby[s];
            a.pbz[k] = a.lin.pbz[s];
            a.scr[k] = a.lin.screen[s];
            a.a2[k] = a.lin.alpha2[s];
            a.b2[k] = a.lin.beta2[s];
            a.wa[k] = a.lwa[s];
            a.wb[k] = a.lwb[s];
        } else {
            a.alpha[k] = 1.0;
            a.inv[k] = 1.0;
            a.prefac[k] = 0.0;
            a.x[k] = a.y[k] = a.z[k] = 0.0;
            a.pax[k] = a.pay[k] = a.paz[k] = 0.0;
            a.pbx[k] = a.pby[k] = a.pbz[k] = 0.0;
            a.scr[k] = 0.0;
            a.a2[k] = 1.0;
            a.b2[k] = 1.0;
            a.wa[k] = 0.0;
            a.wb[k] = 0.0;
        }
    }
}

// max |(bb|bb)| over the diagonal elements of each pair's diagonal block
__global__ void __launch_bounds__(256) diag_max(const double *blocks, int npair, int nab, double *bound)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npair) return;
    const double *q = blocks + (long long)p * nab * nab;
    double m = 0.0;
    for (int k = 0; k < nab; k++) m = fmax(m, fabs(q[(long long)k * nab + k]));
    bound[p] = m;
}

// kets of a bound-sorted (descending) group surviving bound_b bound_k >=
// tol^2 for bra b: a binary search; within one group only k <= b counts
__global__ void __launch_bounds__(256) screen_counts(int nbra, const double *bb, int nket, const double *kb,
                                                     int same, double tol2, int *counts)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nbra) return;
    int n = 0;
    if (bb[b] > 0.0) {
        const double need = tol2 / bb[b];
        int lo = 0, hi = nket;   // kb[k] >= need for k < lo; kb[k] < need for k >= hi
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (kb[mid] >= need) lo = mid + 1;
            else hi = mid;
        }
        n = lo;
    }
    if (same) n = min(n, b + 1);
    counts[b] = n;
}

__global__ void __launch_bounds__(256) iota_kernel(int n, int *v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ void __launch_bounds__(256) gather_int(int n, const int *idx, const int *src, int *dst)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

inline unsigned blocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" {

int simint_hip_pairs_from_shells(const simint_hip_shells *S, int npair, const int *pairs, int screen_method,
                                 simint_hip_pairs **out)
{
    if (!S || !out || npair < 0 || (npair > 0 && !pairs)) return SIMINT_ERROR_ARGUMENT;
    if (screen_method != SIMINT_SCREEN_NONE && screen_method != SIMINT_SCREEN_FASTSCHWARZ) {
        simint_rt::set_error("device pair fill: screening by fast Schwarz (2) or none (0)");
        return SIMINT_ERROR_UNSUPPORTED;
    }
    // the integer layout (validation, counts, offsets, tiles) on the host;
    // every floating-point quantity on the device
    int am1 = 0, am2 = 0;
    std::vector<int> ppoff(npair + 1, 0);
    for (int i = 0; i < npair; i++) {
        const int A = pairs[2 * i], B = pairs[2 * i + 1];
        if (A < 0 || B < 0 || A >= S->n || B >= S->n) {
            simint_rt::set_error("device pair fill: shell index out of range");
            return SIMINT_ERROR_ARGUMENT;
        }
        if (i == 0) {
            am1 = S->host_am[A];
            am2 = S->host_am[B];
        } else if (S->host_am[A] != am1 || S->host_am[B] != am2) {
            simint_rt::set_error("device pair fill: pairs of more than one (am1, am2)");
            return SIMINT_ERROR_ARGUMENT;
        }
        const int na = S->host_nprim[A], nb = S->host_nprim[B];
        const int np = A == B ? na * (na + 1) / 2 : na * nb;
        if (np > kMaxPairPrim) {
            simint_rt::set_error("device pair fill: too many primitive pairs in one shell pair");
            return SIMINT_ERROR_UNSUPPORTED;
        }
        ppoff[i + 1] = ppoff[i] + np;
    }
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const int nprim = ppoff[npair];
    const int ntile = (npair + SIMINT_TILE - 1) / SIMINT_TILE;
    std::vector<int> tnp(std::max(ntile, 1), 0);
    std::vector<long long> toff(ntile + 1, 0);
    for (int t = 0; t < ntile; t++) {
        for (int l = 0; l < SIMINT_TILE && t * SIMINT_TILE + l < npair; l++) {
            const int p = t * SIMINT_TILE + l;
            tnp[t] = std::max(tnp[t], ppoff[p + 1] - ppoff[p]);
        }
        toff[t + 1] = toff[t] + (long long)tnp[t] * SIMINT_TILE;
    }
    const long long tiled = toff[ntile];

    // one allocation: linear view (17 fields per primitive pair), per pair
    // offsets, counts and A-B, shell indices, tiled view (17 fields per row
    // slot), tile offsets and rows, per tile lane A-B and column
    const size_t nl = (size_t)std::max(nprim, 1), nt = (size_t)std::max(tiled, 1LL);
    const size_t nlane = (size_t)std::max(ntile, 1) * SIMINT_TILE, npx = (size_t)std::max(npair, 1);
    size_t bytes = 0;
    auto take = [&](size_t n) {
        const size_t off = bytes;
        bytes = (bytes + n + 255) & ~(size_t)255;
        return off;
    };
    const size_t o_lin = take(sizeof(double) * 17 * nl);
    const size_t o_poff = take(sizeof(int) * (npair + 1));
    const size_t o_np = take(sizeof(int) * npx);
    const size_t o_ab = take(sizeof(double) * 3 * npx);
    const size_t o_pairs = take(sizeof(int) * 2 * npx);
    const size_t o_til = take(sizeof(double) * 17 * nt);
    const size_t o_toff = take(sizeof(long long) * (ntile + 1));
    const size_t o_tnp = take(sizeof(int) * tnp.size());
    const size_t o_tab = take(sizeof(double) * 3 * nlane);
    const size_t o_col = take(sizeof(long long) * nlane);
    const hipStream_t s = simint_rt::stream();
    hipError_t e;
    void *d = nullptr;
    if ((e = hipMalloc(&d, bytes)) != hipSuccess) { simint_rt::set_hip_error("hipMalloc(pairs)", e); return SIMINT_ERROR_HIP; }
    char *base = static_cast<char *>(d);
    auto fail = [&](const char *what, hipError_t err) {
        hipStreamSynchronize(s);
        hipFree(d);
        simint_rt::set_hip_error(what, err);
        return SIMINT_ERROR_HIP;
    };
    if ((e = hipMemcpyAsync(base + o_poff, ppoff.data(), sizeof(int) * (npair + 1), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (npair > 0 && (e = hipMemcpyAsync(base + o_pairs, pairs, sizeof(int) * 2 * npair, hipMemcpyHostToDevice, s)) !=
                          hipSuccess) ||
        (e = hipMemcpyAsync(base + o_toff, toff.data(), sizeof(long long) * (ntile + 1), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(base + o_tnp, tnp.data(), sizeof(int) * tnp.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
        return fail("hipMemcpyAsync(pair layout)", e);
    double *L = reinterpret_cast<double *>(base + o_lin);
    double *T = reinterpret_cast<double *>(base + o_til);
    FillArgs f;
    f.am = S->am;
    f.nprim = S->nprim;
    f.poff = S->poff;
    f.xyz = S->xyz;
    f.alpha = S->alpha;
    f.coef = S->coef;
    f.pairs = reinterpret_cast<const int *>(base + o_pairs);
    f.npair = npair;
    f.screen = screen_method;
    f.ppoff = reinterpret_cast<const int *>(base + o_poff);
    double *fl[17];
    for (int k = 0; k < 17; k++) fl[k] = L + (size_t)k * nl;
    f.alpha_ = fl[0]; f.inv = fl[1]; f.prefac = fl[2]; f.x = fl[3]; f.y = fl[4]; f.z = fl[5];
    f.pax = fl[6]; f.pay = fl[7]; f.paz = fl[8]; f.pbx = fl[9]; f.pby = fl[10]; f.pbz = fl[11];
    f.scr = fl[12]; f.a2 = fl[13]; f.b2 = fl[14]; f.wa = fl[15]; f.wb = fl[16];
    f.np = reinterpret_cast<int *>(base + o_np);
    double *d_ab = reinterpret_cast<double *>(base + o_ab);
    f.abx = d_ab;
    f.aby = d_ab + npx;
    f.abz = d_ab + 2 * npx;
    if (npair > 0) {
        hipLaunchKernelGGL(pair_fill, dim3(blocks(npair, 128)), dim3(128), 0, s, f);
        if ((e = hipGetLastError()) != hipSuccess) return fail("pair fill launch", e);
    }
    simint_hip_pairs *p = new simint_hip_pairs;
    p->am1 = am1;
    p->am2 = am2;
    p->n = npair;
    p->has_screen = screen_method != SIMINT_SCREEN_NONE;
    // as the host fill: the shell-level maximum starts at 1e99 (shell.c:434)
    p->screen_max = 1.e99;
    p->dev = d;
    p->bytes = bytes;
    PairsLinear &lin = p->lin;
    lin.alpha = f.alpha_; lin.inv_alpha = f.inv; lin.prefac = f.prefac;
    lin.x = f.x; lin.y = f.y; lin.z = f.z;
    lin.pax = f.pax; lin.pay = f.pay; lin.paz = f.paz;
    lin.pbx = f.pbx; lin.pby = f.pby; lin.pbz = f.pbz;
    lin.screen = f.scr; lin.alpha2 = f.a2; lin.beta2 = f.b2;
    lin.poff = f.ppoff; lin.np = f.np;
    lin.abx = f.abx; lin.aby = f.aby; lin.abz = f.abz;
    p->wa = f.wa;
    p->wb = f.wb;
    TileArgs ta;
    ta.lin = lin;
    ta.lwa = f.wa;
    ta.lwb = f.wb;
    ta.npair = npair;
    ta.ntile = ntile;
    ta.toff = reinterpret_cast<const long long *>(base + o_toff);
    ta.tnp = reinterpret_cast<const int *>(base + o_tnp);
    double *tf[17];
    for (int k = 0; k < 17; k++) tf[k] = T + (size_t)k * nt;
    ta.alpha = tf[0]; ta.inv = tf[1]; ta.prefac = tf[2]; ta.x = tf[3]; ta.y = tf[4]; ta.z = tf[5];
    ta.pax = tf[6]; ta.pay = tf[7]; ta.paz = tf[8]; ta.pbx = tf[9]; ta.pby = tf[10]; ta.pbz = tf[11];
    ta.scr = tf[12]; ta.a2 = tf[13]; ta.b2 = tf[14]; ta.wa = tf[15]; ta.wb = tf[16];
    double *d_tab = reinterpret_cast<double *>(base + o_tab);
    ta.abx = d_tab;
    ta.aby = d_tab + nlane;
    ta.abz = d_tab + 2 * nlane;
    ta.col = reinterpret_cast<long long *>(base + o_col);
    if (ntile > 0) {
        hipLaunchKernelGGL(tile_fill, dim3(ntile), dim3(64), 0, s, ta);
        if ((e = hipGetLastError()) != hipSuccess) {
            delete p;
            return fail("tile fill launch", e);
        }
    }
    PairsTiled &til = p->til;
    til.alpha = ta.alpha; til.inv_alpha = ta.inv; til.prefac = ta.prefac;
    til.x = ta.x; til.y = ta.y; til.z = ta.z;
    til.pax = ta.pax; til.pay = ta.pay; til.paz = ta.paz;
    til.pbx = ta.pbx; til.pby = ta.pby; til.pbz = ta.pbz;
    til.screen = ta.scr; til.alpha2 = ta.a2; til.beta2 = ta.b2;
    til.toff = ta.toff; til.tnp = ta.tnp;
    til.abx = ta.abx; til.aby = ta.aby; til.abz = ta.abz;
    til.col = ta.col;
    p->twa = ta.wa;
    p->twb = ta.wb;
    p->ntile = ntile;
    // the host layout arrays are released when this returns: wait for the copies
    if ((e = hipStreamSynchronize(s)) != hipSuccess) {
        delete p;
        return fail("hipStreamSynchronize", e);
    }
    *out = p;
    return 0;
}

long long simint_hip_pair_bounds(const simint_hip_pairs *P, double *d_bound)
{
    if (!P || (P->n > 0 && !d_bound)) return SIMINT_ERROR_ARGUMENT;
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    if (P->n == 0) return 0;
    const long long nab = (long long)((P->am1 + 1) * (P->am1 + 2) / 2) * ((P->am2 + 1) * (P->am2 + 2) / 2);
    double *blocks_ = nullptr;
    hipError_t e = hipMalloc(&blocks_, sizeof(double) * nab * nab * P->n);
    if (e != hipSuccess) { simint_rt::set_hip_error("hipMalloc(diagonal blocks)", e); return SIMINT_ERROR_HIP; }
    const long long r = simint_hip_compute_eri_diag(P, blocks_);
    const hipStream_t s = simint_rt::stream();
    if (r >= 0) {
        hipLaunchKernelGGL(diag_max, dim3(blocks(P->n, 256)), dim3(256), 0, s, blocks_, P->n, (int)nab, d_bound);
        e = hipGetLastError();
    }
    hipStreamSynchronize(s);
    hipFree(blocks_);
    if (r < 0) return r;
    if (e != hipSuccess) { simint_rt::set_hip_error("bound launch", e); return SIMINT_ERROR_HIP; }
    return P->n;
}

int simint_hip_order_pairs(const simint_hip_pairs *P, const double *d_bound, int *d_order)
{
    if (!P || (P->n > 0 && (!d_bound || !d_order))) return SIMINT_ERROR_ARGUMENT;
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const int n = P->n;
    if (n == 0) return 0;
    const hipStream_t s = simint_rt::stream();
    // two stable descending radix sorts: by bound, then by primitive count
    double *kb = nullptr;
    int *ki = nullptr, *v0 = nullptr, *v1 = nullptr, *kn = nullptr;
    size_t t1 = 0, t2 = 0;
    hipcub::DeviceRadixSort::SortPairsDescending(nullptr, t1, kb, kb, v0, v1, n, 0, 64, s);
    hipcub::DeviceRadixSort::SortPairsDescending(nullptr, t2, ki, ki, v0, v1, n, 0, 32, s);
    const size_t tb = std::max(t1, t2);
    void *w = nullptr;
    hipError_t e = hipMalloc(&w, tb + sizeof(double) * n + sizeof(int) * 4 * n + 1024);
    if (e != hipSuccess) { simint_rt::set_hip_error("hipMalloc(sort)", e); return SIMINT_ERROR_HIP; }
    char *c = static_cast<char *>(w);
    void *ct = c;
    c += (tb + 255) & ~(size_t)255;
    kb = reinterpret_cast<double *>(c);
    c += ((sizeof(double) * n) + 255) & ~(size_t)255;
    v0 = reinterpret_cast<int *>(c);
    v1 = v0 + n;
    ki = v1 + n;
    kn = ki + n;
    size_t a1 = t1, a2 = t2;
    hipLaunchKernelGGL(iota_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, n, v0);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairsDescending(ct, a1, d_bound, kb, v0, v1, n, 0, 64, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gather_int, dim3(blocks(n, 256)), dim3(256), 0, s, n, v1, P->lin.np, ki);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairsDescending(ct, a2, ki, kn, v1, d_order, n, 0, 32, s);
    hipStreamSynchronize(s);
    hipFree(w);
    if (e != hipSuccess) { simint_rt::set_hip_error("pair sort", e); return SIMINT_ERROR_HIP; }
    return 0;
}

int simint_hip_screen_counts(int nbra, const double *d_bra_bound, int nket, const double *d_ket_bound, int same,
                             double screen_tol, int *d_counts)
{
    if (nbra < 0 || nket < 0 || (nbra > 0 && (!d_bra_bound || !d_counts)) || (nket > 0 && !d_ket_bound))
        return SIMINT_ERROR_ARGUMENT;
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    if (nbra == 0) return 0;
    hipLaunchKernelGGL(screen_counts, dim3(blocks(nbra, 256)), dim3(256), 0, simint_rt::stream(), nbra, d_bra_bound,
                       nket, d_ket_bound, same ? 1 : 0, screen_tol * screen_tol, d_counts);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { simint_rt::set_hip_error("screen count launch", e); return SIMINT_ERROR_HIP; }
    return 0;
}

long long simint_hip_pairs_export(const simint_hip_pairs *P, double *fields, int *nprim12, double *ab)
{
    if (!P) return SIMINT_ERROR_ARGUMENT;
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const hipStream_t s = simint_rt::stream();
    hipError_t e;
    int nprim = 0;
    if (P->n > 0) {
        // the last pair's first primitive and count give the total
        int last[2] = {0, 0};
        if ((e = hipMemcpyAsync(&last[0], P->lin.poff + P->n - 1, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(&last[1], P->lin.np + P->n - 1, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess) {
            simint_rt::set_hip_error("pair export", e);
            return SIMINT_ERROR_HIP;
        }
        nprim = last[0] + last[1];
    }
    if (!fields && !nprim12 && !ab) return nprim;
    const double *src[14] = {P->lin.x, P->lin.y, P->lin.z, P->lin.pax, P->lin.pay, P->lin.paz, P->lin.pbx,
                             P->lin.pby, P->lin.pbz, P->lin.alpha, P->lin.prefac, P->lin.alpha2, P->lin.beta2,
                             P->lin.screen};
    e = hipSuccess;
    for (int k = 0; fields && k < 14 && e == hipSuccess && nprim > 0; k++)
        e = hipMemcpyAsync(fields + (size_t)k * nprim, src[k], sizeof(double) * nprim, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && nprim12 && P->n > 0)
        e = hipMemcpyAsync(nprim12, P->lin.np, sizeof(int) * P->n, hipMemcpyDeviceToHost, s);
    const double *abs_[3] = {P->lin.abx, P->lin.aby, P->lin.abz};
    for (int k = 0; ab && k < 3 && e == hipSuccess && P->n > 0; k++)
        e = hipMemcpyAsync(ab + (size_t)k * P->n, abs_[k], sizeof(double) * P->n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { simint_rt::set_hip_error("pair export", e); return SIMINT_ERROR_HIP; }
    return nprim;
}

}  // extern "C"

namespace simint_rt {

int host_pair_bounds(const simint_multi_shellpair *P, double *out)
{
    simint_hip_pairs *dp = nullptr;
    int rc = simint_hip_pairs_create(P, &dp);
    if (rc < 0) return rc;
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(double) * (P->nshell12 > 0 ? P->nshell12 : 1));
    if (e != hipSuccess) {
        simint_hip_pairs_free(dp);
        set_hip_error("hipMalloc(pair bounds)", e);
        return SIMINT_ERROR_HIP;
    }
    const long long r = simint_hip_pair_bounds(dp, d);   // synchronises the stream
    if (r >= 0 && P->nshell12 > 0) {
        e = hipMemcpyAsync(out, d, sizeof(double) * P->nshell12, hipMemcpyDeviceToHost, stream());
        if (e == hipSuccess) e = hipStreamSynchronize(stream());
    }
    hipFree(d);
    simint_hip_pairs_free(dp);
    if (r < 0) return (int)r;
    if (e != hipSuccess) { set_hip_error("pair bounds copy", e); return SIMINT_ERROR_HIP; }
    return 0;
}

}  // namespace simint_rt

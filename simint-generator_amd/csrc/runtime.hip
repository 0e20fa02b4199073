// Host runtime of the MI355X ERI engine: device set-up, shell-pair upload,
// work lists, class dispatch, and the simint_compute_eri drop-in.
//
// simint_compute_eri keeps the reference's dispatch structure
// (skel/simint/ostei/ostei.c:10-42): a shell-level screen, then the class
// kernel for (P.am1 P.am2 | Q.am1 Q.am2).  The class kernel runs on the GPU;
// there is no CPU path.  If the library finds no GPU every compute call
// returns SIMINT_ERROR_NO_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>

#include "simint/simint.h"
#include "simint/simint_hip.h"
#include "simint/ostei/ostei.h"
#include "device_types.h"
#include "dispatch.h"
#include "runtime_internal.h"

namespace {

const double kBoysGrid[366 * 32] = {
#include "gen/boys_grid.inc"
};

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// first-derivative assembly table of one class: for each Cartesian k of
// the class and each explicit centre slot j (the three centres other than
// the skipped one, in A..D order) and direction x, the offsets of the
// (a+1) and (a-1) values in their class blocks and the factor a_x
// One entry of a derivative assembly table, packed into 32 bits (the
// assembly reads nine per Cartesian): bits 0-14 the (a+1) index, 15-29 the
// (a-1) index, 30-31 the exponent the (a-1) term is weighted by.  Indices
// stay below 2^15: the largest (a+1) block, (gf|ff), has 15000 values.
typedef unsigned int D1Entry;
__host__ __device__ inline D1Entry d1_pack(int ip, int im, int ex)
{
    return (unsigned)ip | ((unsigned)im << 15) | ((unsigned)ex << 30);
}
constexpr int D1_N = SIMINT_OSTEI_DERIV1_MAXAM + 1;
static_assert((D1_N + 1) * (D1_N + 2) / 2 * (D1_N * (D1_N + 1) / 2) * (D1_N * (D1_N + 1) / 2) *
                      (D1_N * (D1_N + 1) / 2) < (1 << 15),
              "derivative assembly indices must fit the 15-bit fields of D1Entry");
// the (a+1) classes of a derivative reach one AM above the derivative maximum
constexpr int kTermMaxAM = SIMINT_OSTEI_DERIV1_MAXAM + 1;

struct DeviceState {
    std::once_flag once;
    bool ok = false;
    std::atomic<bool> ready{false};   // set once the device is fully set up
    int id = 0;
    double *boys = nullptr;
    double *scratch = nullptr;        // device address of the fixed scratch area
    const D1Entry *d1tab[D1_N][D1_N][D1_N][D1_N] = {};
};
DeviceState g_dev;

// Fixed scratch area of simint_hip_compute_eri_deriv1_list_noalloc: the
// (a+1) / (a-1) class blocks of a derivative list live here instead of in a
// caller-supplied work buffer.  The device keeps it for the life of the
// process; every call on the device shares it.
constexpr size_t kNoAllocScratchDoubles = SIMINT_HIP_NOALLOC_SCRATCH_DOUBLES;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

thread_local std::string t_err;

void set_err(const char *what, hipError_t e)
{
    char buf[512];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    t_err = buf;
}

void device_init()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        t_err = "no HIP device";
        return;
    }
    int id = 0;
    if (const char *s = getenv("SIMINT_DEVICE")) id = atoi(s);
    if (id < 0 || id >= n) id = 0;
    if (hipSetDevice(id) != hipSuccess) {
        t_err = "hipSetDevice failed";
        return;
    }
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(kBoysGrid));
    if (e != hipSuccess) { set_err("hipMalloc(boys)", e); return; }
    e = hipMemcpy(d, kBoysGrid, sizeof(kBoysGrid), hipMemcpyHostToDevice);
    if (e != hipSuccess) { set_err("hipMemcpy(boys)", e); return; }
    // derivative assembly tables of every class up to the derivative maximum
    std::vector<D1Entry> all;
    std::vector<size_t> start;
    auto ncart_h = [](int l) { return (l + 1) * (l + 2) / 2; };
    auto cidx = [](int ix, int iy, int iz) {
        const int l = ix + iy + iz;
        return (l - ix) * (l - ix + 1) / 2 + (l - ix - iy);
    };
    for (int a = 0; a < D1_N; a++)
        for (int b = 0; b < D1_N; b++)
            for (int c = 0; c < D1_N; c++)
                for (int dd = 0; dd < D1_N; dd++) {
                    const int am[4] = {a, b, c, dd};
                    const int mx = std::max(std::max(a, b), std::max(c, dd));
                    int skip = 0;
                    while (am[skip] != mx) skip++;
                    const int n[4] = {ncart_h(a), ncart_h(b), ncart_h(c), ncart_h(dd)};
                    start.push_back(all.size());
                    for (int k = 0; k < n[0] * n[1] * n[2] * n[3]; k++) {
                        int comp[4], r = k;
                        for (int q = 3; q >= 0; q--) { comp[q] = r % n[q]; r /= n[q]; }
                        for (int cen = 0; cen < 4; cen++) {
                            if (cen == skip) continue;
                            const int l = am[cen];
                            int e[3] = {0, 0, 0}, idx = 0;
                            for (int ix = l; ix >= 0; ix--)
                                for (int iy = l - ix; iy >= 0; iy--, idx++)
                                    if (idx == comp[cen]) { e[0] = ix; e[1] = iy; e[2] = l - ix - iy; }
                            for (int x = 0; x < 3; x++) {
                                int dp[4] = {n[0], n[1], n[2], n[3]}, cp[4] = {comp[0], comp[1], comp[2], comp[3]};
                                int ep[3] = {e[0], e[1], e[2]};
                                ep[x] += 1;
                                dp[cen] = ncart_h(l + 1);
                                cp[cen] = cidx(ep[0], ep[1], ep[2]);
                                const int ip = ((cp[0] * dp[1] + cp[1]) * dp[2] + cp[2]) * dp[3] + cp[3];
                                int im = 0;
                                if (e[x] > 0) {
                                    int dm[4] = {n[0], n[1], n[2], n[3]}, cm[4] = {comp[0], comp[1], comp[2], comp[3]};
                                    int em[3] = {e[0], e[1], e[2]};
                                    em[x] -= 1;
                                    dm[cen] = ncart_h(l - 1);
                                    cm[cen] = cidx(em[0], em[1], em[2]);
                                    im = ((cm[0] * dm[1] + cm[1]) * dm[2] + cm[2]) * dm[3] + cm[3];
                                }
                                all.push_back(d1_pack(ip, im, e[x]));
                            }
                        }
                    }
                }
    D1Entry *dt = nullptr;
    e = hipMalloc(&dt, sizeof(D1Entry) * std::max(all.size(), (size_t)1));
    if (e != hipSuccess) { set_err("hipMalloc(deriv tables)", e); return; }
    e = hipMemcpy(dt, all.data(), sizeof(D1Entry) * all.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { set_err("hipMemcpy(deriv tables)", e); return; }
    int k = 0;
    for (int a = 0; a < D1_N; a++)
        for (int b = 0; b < D1_N; b++)
            for (int c = 0; c < D1_N; c++)
                for (int dd = 0; dd < D1_N; dd++) g_dev.d1tab[a][b][c][dd] = dt + start[k++];
    g_dev.id = id;
    g_dev.boys = d;
    // the host name `scratch` only stands for the device array: look up its
    // device address once, with this device current, and use that
    void *sp = nullptr;
    e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess) { set_err("hipGetSymbolAddress(scratch)", e); return; }
    g_dev.scratch = static_cast<double *>(sp);
    g_dev.ok = true;
    g_dev.ready.store(true, std::memory_order_release);
}

bool ensure_device()
{
    std::call_once(g_dev.once, device_init);
    if (g_dev.ok) hipSetDevice(g_dev.id);
    return g_dev.ok;
}

// device memory grown on demand, never shrunk while the thread lives
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= cap) return hipSuccess;
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max(n, (size_t)1 << 20);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
};

struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= cap) return hipSuccess;
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max(n, (size_t)1 << 20);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) cap = want;
        return e;
    }
};

// Per-thread stream and buffers.  They are never released: freeing HIP
// resources from thread-exit or static destructors can race the runtime's
// own teardown, and the process exit returns the memory anyway.
struct ThreadCtx {
    hipStream_t s = nullptr;       // the stream calls of this thread queue on
    hipStream_t own = nullptr;     // this thread's own stream (s unless redirected)
    DevBuf in, out, items;
    DevBuf split;      // staging of the split lane-per-quartet kernels
    PinBuf stage;
    ThreadCtx()
    {
        hipStreamCreateWithFlags(&own, hipStreamNonBlocking);
        s = own;
    }
};

ThreadCtx &ctx()
{
    thread_local ThreadCtx *c = nullptr;
    if (!c) c = new ThreadCtx();
    return *c;
}

// ---------------------------------------------------------------- packing
// A host-side image of device arrays, laid out in one buffer (256-byte
// aligned sub-arrays) so one copy moves everything.
struct Image {
    std::vector<char> buf;
    size_t add(size_t bytes)
    {
        size_t off = (buf.size() + 255) & ~(size_t)255;
        buf.resize(off + std::max(bytes, (size_t)8));
        return off;
    }
    template <class T> T *at(size_t off) { return reinterpret_cast<T *>(buf.data() + off); }
};

// prefac_a2 / prefac_b2 (when present) hold prefac * alpha2 and prefac *
// beta2: with them an ordinary class kernel of class (a+1 b|..) yields the
// 2a-weighted contraction a first derivative needs
// (OSTEIDeriv1_Writer.cpp:745-746, :800-801 contract with these factors).
struct LinearOffs {
    size_t alpha, inv, prefac, x, y, z, pax, pay, paz, pbx, pby, pbz, screen, a2, b2, poff, np, abx, aby, abz;
    size_t prefac_a2 = 0, prefac_b2 = 0;
};
struct TiledOffs {
    size_t alpha, inv, prefac, x, y, z, pax, pay, paz, pbx, pby, pbz, screen, a2, b2, toff, tnp, abx, aby, abz, col;
    size_t prefac_a2 = 0, prefac_b2 = 0;
    int ntile;
};

// bra view: pairs [0, n) of P, primitives contiguous as in the host struct
LinearOffs pack_linear(Image &im, const simint_multi_shellpair *P, int n, bool weighted = false)
{
    LinearOffs o;
    long long nprim = 0;
    for (int i = 0; i < n; i++) nprim += P->nprim12[i];
    const size_t db = sizeof(double) * (size_t)nprim;
    o.alpha = im.add(db); o.inv = im.add(db); o.prefac = im.add(db);
    o.x = im.add(db); o.y = im.add(db); o.z = im.add(db);
    o.pax = im.add(db); o.pay = im.add(db); o.paz = im.add(db);
    o.pbx = im.add(db); o.pby = im.add(db); o.pbz = im.add(db);
    o.screen = im.add(db); o.a2 = im.add(db); o.b2 = im.add(db);
    o.poff = im.add(sizeof(int) * (size_t)n); o.np = im.add(sizeof(int) * (size_t)n);
    o.abx = im.add(sizeof(double) * (size_t)n); o.aby = im.add(sizeof(double) * (size_t)n);
    o.abz = im.add(sizeof(double) * (size_t)n);
    const struct { size_t off; const double *src; } cp[] = {
        {o.alpha, P->alpha}, {o.prefac, P->prefac}, {o.x, P->x}, {o.y, P->y}, {o.z, P->z},
        {o.pax, P->PA_x}, {o.pay, P->PA_y}, {o.paz, P->PA_z}, {o.pbx, P->PB_x}, {o.pby, P->PB_y},
        {o.pbz, P->PB_z}, {o.a2, P->alpha2}, {o.b2, P->beta2}};
    for (auto &c : cp) memcpy(im.at<double>(c.off), c.src, db);
    if (P->screen) memcpy(im.at<double>(o.screen), P->screen, db);
    else memset(im.at<double>(o.screen), 0, db);
    double *inv = im.at<double>(o.inv);
    const double *al = P->alpha;
    for (long long k = 0; k < nprim; k++) inv[k] = 1.0 / al[k];
    if (weighted) {
        o.prefac_a2 = im.add(db);
        o.prefac_b2 = im.add(db);
        double *wa = im.at<double>(o.prefac_a2), *wb = im.at<double>(o.prefac_b2);
        for (long long k = 0; k < nprim; k++) {
            wa[k] = P->prefac[k] * P->alpha2[k];
            wb[k] = P->prefac[k] * P->beta2[k];
        }
    }
    int *poff = im.at<int>(o.poff), *np = im.at<int>(o.np);
    int acc = 0;
    for (int i = 0; i < n; i++) {
        poff[i] = acc;
        np[i] = P->nprim12[i];
        acc += P->nprim12[i];
        im.at<double>(o.abx)[i] = P->AB_x[i];
        im.at<double>(o.aby)[i] = P->AB_y[i];
        im.at<double>(o.abz)[i] = P->AB_z[i];
    }
    return o;
}

// ket view: pairs [0, n) of Q in tiles of 64.  `order` lists the pairs in
// tile order; col[] maps a tile lane back to its position in `colmap`.
TiledOffs pack_tiled(Image &im, const simint_multi_shellpair *Q, const std::vector<int> &order,
                     const std::vector<long long> &colmap, bool weighted = false)
{
    TiledOffs o;
    const int n = (int)order.size();
    const int ntile = (n + SIMINT_TILE - 1) / SIMINT_TILE;
    o.ntile = ntile;
    std::vector<long long> pstart(Q->nshell12 + 1, 0);
    for (int i = 0; i < Q->nshell12; i++) pstart[i + 1] = pstart[i] + Q->nprim12[i];
    std::vector<int> tnp(ntile, 0);
    std::vector<long long> toff(ntile, 0);
    long long total = 0;
    for (int t = 0; t < ntile; t++) {
        int mx = 0;
        for (int l = 0; l < SIMINT_TILE && t * SIMINT_TILE + l < n; l++)
            mx = std::max(mx, Q->nprim12[order[t * SIMINT_TILE + l]]);
        tnp[t] = mx;
        toff[t] = total;
        total += (long long)mx * SIMINT_TILE;
    }
    const size_t db = sizeof(double) * (size_t)std::max(total, 1LL);
    o.alpha = im.add(db); o.inv = im.add(db); o.prefac = im.add(db);
    o.x = im.add(db); o.y = im.add(db); o.z = im.add(db);
    o.pax = im.add(db); o.pay = im.add(db); o.paz = im.add(db);
    o.pbx = im.add(db); o.pby = im.add(db); o.pbz = im.add(db);
    o.screen = im.add(db); o.a2 = im.add(db); o.b2 = im.add(db);
    o.toff = im.add(sizeof(long long) * (size_t)ntile);
    o.tnp = im.add(sizeof(int) * (size_t)ntile);
    const size_t nl = (size_t)ntile * SIMINT_TILE;
    o.abx = im.add(sizeof(double) * nl); o.aby = im.add(sizeof(double) * nl); o.abz = im.add(sizeof(double) * nl);
    o.col = im.add(sizeof(long long) * nl);
    double *f[15] = {im.at<double>(o.alpha), im.at<double>(o.inv), im.at<double>(o.prefac), im.at<double>(o.x),
                     im.at<double>(o.y), im.at<double>(o.z), im.at<double>(o.pax), im.at<double>(o.pay),
                     im.at<double>(o.paz), im.at<double>(o.pbx), im.at<double>(o.pby), im.at<double>(o.pbz),
                     im.at<double>(o.screen), im.at<double>(o.a2), im.at<double>(o.b2)};
    const double *src[15] = {Q->alpha, nullptr, Q->prefac, Q->x, Q->y, Q->z, Q->PA_x, Q->PA_y, Q->PA_z,
                             Q->PB_x, Q->PB_y, Q->PB_z, Q->screen, Q->alpha2, Q->beta2};
    // padding: alpha = 1, everything else 0 (zero contribution)
    for (long long k = 0; k < total; k++) {
        for (int a = 0; a < 15; a++) f[a][k] = 0.0;
        f[0][k] = 1.0;
        f[1][k] = 1.0;
        f[13][k] = 1.0;
        f[14][k] = 1.0;
    }
    double *wa = nullptr, *wb = nullptr;
    if (weighted) {
        o.prefac_a2 = im.add(db);
        o.prefac_b2 = im.add(db);
        wa = im.at<double>(o.prefac_a2);
        wb = im.at<double>(o.prefac_b2);
        memset(wa, 0, db);
        memset(wb, 0, db);
        // f[] pointers may have moved with the buffer
        double *g[15] = {im.at<double>(o.alpha), im.at<double>(o.inv), im.at<double>(o.prefac), im.at<double>(o.x),
                         im.at<double>(o.y), im.at<double>(o.z), im.at<double>(o.pax), im.at<double>(o.pay),
                         im.at<double>(o.paz), im.at<double>(o.pbx), im.at<double>(o.pby), im.at<double>(o.pbz),
                         im.at<double>(o.screen), im.at<double>(o.a2), im.at<double>(o.b2)};
        for (int a = 0; a < 15; a++) f[a] = g[a];
    }
    long long *colp = im.at<long long>(o.col);
    for (int t = 0; t < ntile; t++) {
        im.at<long long>(o.toff)[t] = toff[t];
        im.at<int>(o.tnp)[t] = tnp[t];
        for (int l = 0; l < SIMINT_TILE; l++) {
            const size_t tl = (size_t)t * SIMINT_TILE + l;
            const int pos = t * SIMINT_TILE + l;
            if (pos >= n) {
                colp[tl] = -1;
                im.at<double>(o.abx)[tl] = im.at<double>(o.aby)[tl] = im.at<double>(o.abz)[tl] = 0.0;
                continue;
            }
            const int q = order[pos];
            colp[tl] = colmap[pos];
            im.at<double>(o.abx)[tl] = Q->AB_x[q];
            im.at<double>(o.aby)[tl] = Q->AB_y[q];
            im.at<double>(o.abz)[tl] = Q->AB_z[q];
            for (int j = 0; j < Q->nprim12[q]; j++) {
                const long long k = toff[t] + (long long)j * SIMINT_TILE + l;
                const long long s = pstart[q] + j;
                for (int a = 0; a < 15; a++) {
                    if (a == 1) f[a][k] = 1.0 / Q->alpha[s];
                    else if (src[a]) f[a][k] = src[a][s];
                    else f[a][k] = 0.0;
                }
                if (weighted) {
                    wa[k] = Q->prefac[s] * Q->alpha2[s];
                    wb[k] = Q->prefac[s] * Q->beta2[s];
                }
            }
        }
    }
    return o;
}

PairsLinear linear_ptrs(char *base, const LinearOffs &o)
{
    PairsLinear p;
    auto D = [&](size_t off) { return reinterpret_cast<const double *>(base + off); };
    p.alpha = D(o.alpha); p.inv_alpha = D(o.inv); p.prefac = D(o.prefac);
    p.x = D(o.x); p.y = D(o.y); p.z = D(o.z);
    p.pax = D(o.pax); p.pay = D(o.pay); p.paz = D(o.paz);
    p.pbx = D(o.pbx); p.pby = D(o.pby); p.pbz = D(o.pbz);
    p.screen = D(o.screen); p.alpha2 = D(o.a2); p.beta2 = D(o.b2);
    p.poff = reinterpret_cast<const int *>(base + o.poff);
    p.np = reinterpret_cast<const int *>(base + o.np);
    p.abx = D(o.abx); p.aby = D(o.aby); p.abz = D(o.abz);
    return p;
}

PairsTiled tiled_ptrs(char *base, const TiledOffs &o)
{
    PairsTiled p;
    auto D = [&](size_t off) { return reinterpret_cast<const double *>(base + off); };
    p.alpha = D(o.alpha); p.inv_alpha = D(o.inv); p.prefac = D(o.prefac);
    p.x = D(o.x); p.y = D(o.y); p.z = D(o.z);
    p.pax = D(o.pax); p.pay = D(o.pay); p.paz = D(o.paz);
    p.pbx = D(o.pbx); p.pby = D(o.pby); p.pbz = D(o.pbz);
    p.screen = D(o.screen); p.alpha2 = D(o.a2); p.beta2 = D(o.b2);
    p.toff = reinterpret_cast<const long long *>(base + o.toff);
    p.tnp = reinterpret_cast<const int *>(base + o.tnp);
    p.abx = D(o.abx); p.aby = D(o.aby); p.abz = D(o.abz);
    p.col = reinterpret_cast<const long long *>(base + o.col);
    return p;
}

int validate(const simint_multi_shellpair *P)
{
    if (!P) return SIMINT_ERROR_ARGUMENT;
    if (P->am1 < 0 || P->am2 < 0 || P->am1 > SIMINT_OSTEI_MAXAM || P->am2 > SIMINT_OSTEI_MAXAM)
        return SIMINT_ERROR_UNSUPPORTED;
    if (P->nshell12_clip < 0 || P->nshell12_clip > P->nshell12) return SIMINT_ERROR_ARGUMENT;
    return 0;
}

// A first-derivative class needs its (a+1) classes and an assembly table:
// both exist up to SIMINT_OSTEI_DERIV1_MAXAM on every centre.  Checked before
// anything is queued, so an unsupported class leaves the stream untouched.
bool deriv1_built(int a, int b, int c, int d)
{
    return std::max(std::max(a, b), std::max(c, d)) <= SIMINT_OSTEI_DERIV1_MAXAM &&
           std::min(std::min(a, b), std::min(c, d)) >= 0;
}

// ----------------------------------------------------- first derivatives
// d/dX_i (ab|cd) for the three explicit centres from the 2a-weighted (a+1)
// class and the (a-1) class, and the fourth centre (the first with the
// highest AM) as minus their sum -- the assembly of
// OSTEIDeriv1_Writer.cpp:334-537.  Output: 12 doubles per Cartesian,
// [Ax Ay Az Bx ... Dz].
struct Deriv1Args {
    int am[4];
    const double *plus[4];     // per explicit centre: (a+1) class, contiguous per quartet
    const double *minus[4];    // (a-1) class (null where a = 0)
    int skip;
    long long nquartet;
    double *out;
    const long long *qout;     // per quartet output offset, or null: contiguous
    // filled by launch_deriv1_assemble
    long long nk = 0;          // Cartesians of the class
    const D1Entry *tab = nullptr;
    int cen[3] = {0, 0, 0};    // explicit centres in slot order
    long long np[3] = {0, 0, 0}, nm[3] = {0, 0, 0};   // (a+1), (a-1) block sizes per slot
    const double *ps[3] = {nullptr, nullptr, nullptr};   // plus / minus per slot (the kernel
    const double *ms[3] = {nullptr, nullptr, nullptr};   // indexes these by the unrolled slot only)
};

__global__ void __launch_bounds__(256) deriv1_assemble(const Deriv1Args d)
{
    // the 12 values of the block's 256 Cartesians are gathered in LDS and
    // written as one contiguous run when the output is dense
    __shared__ double stage[256 * 12];
    const long long total = d.nquartet * d.nk;
    const long long gid0 = (long long)blockIdx.x * 256;
    const long long gid = gid0 + threadIdx.x;
    const bool live = gid < total;
    double v12[12];
    // (quartet, Cartesian) of this thread: a 32-bit division where it fits
    long long q = 0;
    int k = 0;
    if (live) {
        if (total < (1LL << 31)) {
            const unsigned qq = (unsigned)gid / (unsigned)d.nk;
            q = qq;
            k = (int)((unsigned)gid - qq * (unsigned)d.nk);
        } else {
            q = gid / d.nk;
            k = (int)(gid - q * d.nk);
        }
    }
    if (live) {
        const D1Entry *__restrict__ t = d.tab + (long long)k * 9;
        double s3[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double *__restrict__ P = d.ps[j] + q * d.np[j];
            const double *__restrict__ M = d.ms[j] ? d.ms[j] + q * d.nm[j] : P;
#pragma unroll
            for (int x = 0; x < 3; x++) {
                const D1Entry e = t[j * 3 + x];
                const int ex = (int)(e >> 30);
                double v = P[e & 0x7fffu];
                if (ex > 0) v -= (double)ex * M[(e >> 15) & 0x7fffu];
                v12[j * 3 + x] = v;
                s3[x] -= v;
            }
        }
#pragma unroll
        for (int x = 0; x < 3; x++) v12[9 + x] = s3[x];
    }
    if (d.qout) {
        if (!live) return;
        double *__restrict__ o = d.out + d.qout[q] + (long long)k * 12;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int x = 0; x < 3; x++) o[d.cen[j] * 3 + x] = v12[j * 3 + x];
#pragma unroll
        for (int x = 0; x < 3; x++) o[d.skip * 3 + x] = v12[9 + x];
        return;
    }
    if (live) {
        double *row = stage + threadIdx.x * 12;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int x = 0; x < 3; x++) row[d.cen[j] * 3 + x] = v12[j * 3 + x];
#pragma unroll
        for (int x = 0; x < 3; x++) row[d.skip * 3 + x] = v12[9 + x];
    }
    __syncthreads();
    const long long nval = (total - gid0 < 256 ? total - gid0 : 256) * 12;
    double *__restrict__ o = d.out + gid0 * 12;
    for (int i = threadIdx.x; i < nval; i += 256) o[i] = stage[i];
}

hipError_t launch_deriv1_assemble(Deriv1Args d, hipStream_t s)
{
    if (d.am[0] >= D1_N || d.am[1] >= D1_N || d.am[2] >= D1_N || d.am[3] >= D1_N) return hipErrorInvalidValue;
    const int n[4] = {ncart(d.am[0]), ncart(d.am[1]), ncart(d.am[2]), ncart(d.am[3])};
    d.nk = (long long)n[0] * n[1] * n[2] * n[3];
    d.tab = g_dev.d1tab[d.am[0]][d.am[1]][d.am[2]][d.am[3]];
    int j = 0;
    for (int c = 0; c < 4; c++) {
        if (c == d.skip) continue;
        int p[4] = {n[0], n[1], n[2], n[3]}, m[4] = {n[0], n[1], n[2], n[3]};
        p[c] = ncart(d.am[c] + 1);
        m[c] = d.am[c] > 0 ? ncart(d.am[c] - 1) : 0;
        d.cen[j] = c;
        d.ps[j] = d.plus[c];
        d.ms[j] = d.minus[c];
        d.np[j] = (long long)p[0] * p[1] * p[2] * p[3];
        d.nm[j] = (long long)m[0] * m[1] * m[2] * m[3];
        j++;
    }
    const long long total = d.nquartet * d.nk;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(deriv1_assemble, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d);
    return hipGetLastError();
}

// Cartesian order of the drop-in calls' output.  GAMESS orders d as
// xx yy zz xy xz yz and f as xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz
// (generator/Ordering_GAMESS.cpp); every other AM is the same as Psi4's.
std::atomic<int> g_ordering{SIMINT_ORDERING_PSI4};

// gamess_pos[l][k]: position in GAMESS order of the Psi4 component k of AM l
__constant__ signed char kGamessPos[4][10] = {
    {0},
    {0, 1, 2},
    // psi4: xx xy xz yy yz zz
    {0, 3, 4, 1, 5, 2},
    // psi4: xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    {0, 3, 4, 5, 9, 7, 1, 6, 8, 2},
};

struct PermArgs {
    int am[4];
    int n[4];
    long long nq;      // quartets
    int width;         // values per Cartesian quartet (12 for first derivatives)
    const double *in;
    double *out;
};

// out = in with every quartet's Cartesian components moved to GAMESS order
__global__ void __launch_bounds__(256) to_gamess(const PermArgs p)
{
    const long long nc = (long long)p.n[0] * p.n[1] * p.n[2] * p.n[3];
    const long long total = p.nq * nc * p.width;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
        const int x = (int)(i % p.width);
        long long r = i / p.width;
        int comp[4];
        for (int k = 3; k >= 0; k--) {
            comp[k] = (int)(r % p.n[k]);
            r /= p.n[k];
        }
        const long long q = r;
        long long t = 0;
        for (int k = 0; k < 4; k++) {
            const int g = p.am[k] <= 3 ? kGamessPos[p.am[k]][comp[k]] : comp[k];
            t = t * p.n[k] + g;
        }
        p.out[(q * nc + t) * p.width + x] = p.in[i];
    }
}

hipError_t launch_to_gamess(const int am[4], long long nq, int width, const double *in, double *out, hipStream_t s)
{
    PermArgs p;
    for (int k = 0; k < 4; k++) {
        p.am[k] = am[k];
        p.n[k] = ncart(am[k]);
    }
    p.nq = nq;
    p.width = width;
    p.in = in;
    p.out = out;
    const long long total = nq * p.n[0] * p.n[1] * p.n[2] * p.n[3] * width;
    if (total == 0) return hipSuccess;
    const long long nb = std::min<long long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(to_gamess, dim3((unsigned)nb), dim3(256), 0, s, p);
    return hipGetLastError();
}

bool gamess_needed(const int am[4])
{
    return g_ordering.load(std::memory_order_relaxed) == SIMINT_ORDERING_GAMESS &&
           (am[0] >= 2 || am[1] >= 2 || am[2] >= 2 || am[3] >= 2);
}

}  // namespace

// ------------------------------------------------------------------- C API
extern "C" {

void simint_init(void) { ensure_device(); }

int simint_hip_set_ordering(int ordering)
{
    if (ordering != SIMINT_ORDERING_PSI4 && ordering != SIMINT_ORDERING_GAMESS) return SIMINT_ERROR_ARGUMENT;
    return g_ordering.exchange(ordering);
}

int simint_hip_ordering(void) { return g_ordering.load(); }

void simint_finalize(void)
{
    // device memory is released when the process ends; per-thread buffers
    // when their thread exits
}

const char *simint_hip_last_error(void) { return t_err.c_str(); }

int simint_compute_ostei(const simint_multi_shellpair *P, const simint_multi_shellpair *Q, double screen_tol,
                         double *restrict work, double *restrict integrals)
{
    return simint_compute_ostei_deriv(0, P, Q, screen_tol, work, integrals);
}

int simint_compute_ostei_deriv(int deriv, const simint_multi_shellpair *P, const simint_multi_shellpair *Q,
                               double screen_tol, double *restrict work, double *restrict integrals)
{
    (void)work;
    int v = validate(P);
    if (v) return v;
    v = validate(Q);
    if (v) return v;
    if (deriv < 0 || deriv > SIMINT_OSTEI_MAXDER) {
        t_err = "derivative order not built";
        return SIMINT_ERROR_UNSUPPORTED;
    }
    if (deriv == 1 && !deriv1_built(P->am1, P->am2, Q->am1, Q->am2)) {
        t_err = "first derivatives not built for this class";
        return SIMINT_ERROR_UNSUPPORTED;
    }
    // shell-level screen (ostei.c:17-20): screen values are not square-rooted
    const double tol2 = screen_tol * screen_tol;
    if (screen_tol > 0.0 && (P->screen_max * Q->screen_max) < tol2) return -1;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const int am[4] = {P->am1, P->am2, Q->am1, Q->am2};
    const int nP = P->nshell12_clip, nQ = Q->nshell12_clip;
    const int nc = ncart(am[0]) * ncart(am[1]) * ncart(am[2]) * ncart(am[3]);

    // class launches: the class itself, or for deriv = 1 the (a+1) class of
    // every centre but the one with the highest AM (2a-weighted) and the
    // (a-1) class where a > 0 (ostei_deriv1_generator.cpp:96-124)
    struct Term { int am[4]; int weight; int centre; int sign; simint_launch_fn fn; int nc; size_t off; };
    std::vector<Term> terms;
    int skip = -1;
    if (deriv == 0) {
        terms.push_back(Term{{am[0], am[1], am[2], am[3]}, 0, -1, 0, nullptr, nc, 0});
    } else {
        const int mx = std::max(std::max(am[0], am[1]), std::max(am[2], am[3]));
        for (int c = 0; c < 4; c++)
            if (am[c] == mx) { skip = c; break; }
        for (int c = 0; c < 4; c++) {
            if (c == skip) continue;
            Term t{{am[0], am[1], am[2], am[3]}, 1 + (c & 1), c, +1, nullptr, 0, 0};
            t.am[c] += 1;
            terms.push_back(t);
            if (am[c] > 0) {
                Term m{{am[0], am[1], am[2], am[3]}, 0, c, -1, nullptr, 0, 0};
                m.am[c] -= 1;
                terms.push_back(m);
            }
        }
    }
    size_t scratch = 0;
    for (auto &t : terms) {
        if (t.am[0] > kTermMaxAM || t.am[1] > kTermMaxAM || t.am[2] > kTermMaxAM || t.am[3] > kTermMaxAM ||
            !(t.fn = simint_dispatch(t.am[0], t.am[1], t.am[2], t.am[3]))) {
            t_err = "class not built";
            return SIMINT_ERROR_UNSUPPORTED;
        }
        t.nc = ncart(t.am[0]) * ncart(t.am[1]) * ncart(t.am[2]) * ncart(t.am[3]);
        t.off = scratch;
        scratch += (size_t)nP * nQ * t.nc;
    }
    if (nP == 0 || nQ == 0) return 0;

    // kets into tiles, grouped by primitive count (stable), so a tile's
    // lanes share one primitive loop
    std::vector<int> order(nQ);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return Q->nprim12[a] > Q->nprim12[b]; });
    std::vector<long long> colmap(order.begin(), order.end());

    Image im;
    const LinearOffs lo = pack_linear(im, P, nP, deriv > 0);
    const TiledOffs to = pack_tiled(im, Q, order, colmap, deriv > 0);
    std::vector<size_t> items_off;
    const long long nit = (long long)nP * to.ntile;
    for (auto &t : terms) {
        const size_t off = im.add(sizeof(WorkItem) * (size_t)nit);
        items_off.push_back(off);
        WorkItem *items = im.at<WorkItem>(off);
        long long k = 0;
        for (int b = 0; b < nP; b++)
            for (int tt = 0; tt < to.ntile; tt++) {
                WorkItem &w = items[k++];
                w.bra = b;
                w.tile = tt;
                w.lane_lo = 0;
                w.lane_hi = std::min(SIMINT_TILE, nQ - tt * SIMINT_TILE);
                w.pad = 0;
                w.obase = (long long)b * nQ * t.nc;
            }
    }

    ThreadCtx &c = ctx();
    hipError_t e;
    // an earlier asynchronous call of this thread (_range, _prefix, _diag)
    // may still read c.stage, c.in or c.split: wait for it before reuse
    if ((e = hipStreamSynchronize(c.s)) != hipSuccess) { set_err("hipStreamSynchronize", e); return SIMINT_ERROR_HIP; }
    const size_t final_n = (size_t)nP * nQ * nc * (deriv ? 12 : 1);
    const bool gamess = gamess_needed(am);
    const size_t out_bytes = sizeof(double) * (scratch + (deriv ? final_n : 0) + (gamess ? final_n : 0));
    if ((e = c.stage.reserve(im.buf.size())) != hipSuccess) { set_err("hipHostMalloc", e); return SIMINT_ERROR_HIP; }
    if ((e = c.in.reserve(im.buf.size())) != hipSuccess) { set_err("hipMalloc", e); return SIMINT_ERROR_HIP; }
    if ((e = c.out.reserve(out_bytes)) != hipSuccess) { set_err("hipMalloc", e); return SIMINT_ERROR_HIP; }
    memcpy(c.stage.p, im.buf.data(), im.buf.size());
    e = hipMemcpyAsync(c.in.p, c.stage.p, im.buf.size(), hipMemcpyHostToDevice, c.s);
    if (e != hipSuccess) { set_err("hipMemcpyAsync(H2D)", e); return SIMINT_ERROR_HIP; }

    size_t nst = 0;
    for (auto &t : terms) nst = std::max(nst, (size_t)simint_dispatch_staging(t.am[0], t.am[1], t.am[2], t.am[3]));
    if (nst && (e = c.split.reserve(sizeof(double) * nst * SIMINT_TILE * (size_t)nit)) != hipSuccess) {
        set_err("hipMalloc", e);
        return SIMINT_ERROR_HIP;
    }
    char *base = static_cast<char *>(c.in.p);
    double *dout = static_cast<double *>(c.out.p);
    for (size_t k = 0; k < terms.size(); k++) {
        const Term &t = terms[k];
        EriArgs a;
        a.scratch = static_cast<double *>(c.split.p);
        a.bra = linear_ptrs(base, lo);
        a.ket = tiled_ptrs(base, to);
        if (t.weight && t.centre == 0) a.bra.prefac = reinterpret_cast<const double *>(base + lo.prefac_a2);
        if (t.weight && t.centre == 1) a.bra.prefac = reinterpret_cast<const double *>(base + lo.prefac_b2);
        if (t.weight && t.centre == 2) a.ket.prefac = reinterpret_cast<const double *>(base + to.prefac_a2);
        if (t.weight && t.centre == 3) a.ket.prefac = reinterpret_cast<const double *>(base + to.prefac_b2);
        a.items = reinterpret_cast<const WorkItem *>(base + items_off[k]);
        a.nitems = nit;
        a.out = dout + t.off;
        a.boys_grid = g_dev.boys;
        a.tol2 = (screen_tol > 0.0 && P->screen && Q->screen) ? tol2 : 0.0;
        a.ket_screen_max = Q->screen_max;
        a.stride = t.nc;
        if ((e = t.fn(&a, c.s)) != hipSuccess) {
            set_err("class kernel launch", e);
            hipStreamSynchronize(c.s);   // the call returns with nothing of it still queued
            return SIMINT_ERROR_HIP;
        }
    }
    const double *result = dout;
    if (deriv == 1) {
        Deriv1Args d;
        for (int q = 0; q < 4; q++) {
            d.am[q] = am[q];
            d.plus[q] = nullptr;
            d.minus[q] = nullptr;
        }
        for (const Term &t : terms) {
            if (t.sign > 0) d.plus[t.centre] = dout + t.off;
            else d.minus[t.centre] = dout + t.off;
        }
        d.skip = skip;
        d.nquartet = (long long)nP * nQ;
        d.out = dout + scratch;
        d.qout = nullptr;
        if ((e = launch_deriv1_assemble(d, c.s)) != hipSuccess) {
            set_err("deriv1 assembly", e);
            hipStreamSynchronize(c.s);
            return SIMINT_ERROR_HIP;
        }
        result = d.out;
    }
    if (gamess) {
        double *perm = dout + scratch + (deriv ? final_n : 0);
        if ((e = launch_to_gamess(am, (long long)nP * nQ, deriv ? 12 : 1, result, perm, c.s)) != hipSuccess) {
            set_err("ordering permutation", e);
            hipStreamSynchronize(c.s);
            return SIMINT_ERROR_HIP;
        }
        result = perm;
    }
    e = hipMemcpyAsync(integrals, result, sizeof(double) * final_n, hipMemcpyDeviceToHost, c.s);
    if (e != hipSuccess) {
        set_err("hipMemcpyAsync(D2H)", e);
        hipStreamSynchronize(c.s);
        return SIMINT_ERROR_HIP;
    }
    e = hipStreamSynchronize(c.s);
    if (e != hipSuccess) { set_err("hipStreamSynchronize", e); return SIMINT_ERROR_HIP; }
    return nP * nQ;
}

int simint_compute_eri(const simint_multi_shellpair *P, const simint_multi_shellpair *Q, double screen_tol,
                       double *restrict work, double *restrict integrals)
{
    return simint_compute_ostei(P, Q, screen_tol, work, integrals);
}

int simint_compute_eri_deriv(int deriv, const simint_multi_shellpair *P, const simint_multi_shellpair *Q,
                             double screen_tol, double *restrict work, double *restrict integrals)
{
    return simint_compute_ostei_deriv(deriv, P, Q, screen_tol, work, integrals);
}

size_t simint_eri_worksize(int derorder, int maxam) { return simint_ostei_worksize(derorder, maxam); }
size_t simint_eri_workmem(int derorder, int maxam) { return simint_ostei_workmem(derorder, maxam); }

// --------------------------------------------------------- device-resident
int simint_hip_pairs_create(const simint_multi_shellpair *P, simint_hip_pairs **out)
{
    int v = validate(P);
    if (v) return v;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const int n = P->nshell12_clip;
    Image im;
    const LinearOffs lo = pack_linear(im, P, n, true);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::vector<long long> colmap(order.begin(), order.end());
    const TiledOffs to = pack_tiled(im, P, order, colmap, true);
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, im.buf.size());
    if (e != hipSuccess) { set_err("hipMalloc(pairs)", e); return SIMINT_ERROR_HIP; }
    e = hipMemcpy(d, im.buf.data(), im.buf.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { hipFree(d); set_err("hipMemcpy(pairs)", e); return SIMINT_ERROR_HIP; }
    simint_hip_pairs *p = new simint_hip_pairs;
    p->am1 = P->am1;
    p->am2 = P->am2;
    p->n = n;
    p->screen_max = P->screen_max;
    p->has_screen = P->screen != nullptr;
    p->dev = d;
    p->bytes = im.buf.size();
    p->lin = linear_ptrs(static_cast<char *>(d), lo);
    p->til = tiled_ptrs(static_cast<char *>(d), to);
    p->ntile = to.ntile;
    p->wa = reinterpret_cast<const double *>(static_cast<char *>(d) + lo.prefac_a2);
    p->wb = reinterpret_cast<const double *>(static_cast<char *>(d) + lo.prefac_b2);
    p->twa = reinterpret_cast<const double *>(static_cast<char *>(d) + to.prefac_a2);
    p->twb = reinterpret_cast<const double *>(static_cast<char *>(d) + to.prefac_b2);
    *out = p;
    return 0;
}

void simint_hip_pairs_free(simint_hip_pairs *p)
{
    if (!p) return;
    if (p->dev) hipFree(p->dev);
    delete p;
}

int simint_hip_pairs_count(const simint_hip_pairs *p) { return p ? p->n : 0; }

size_t simint_hip_pairs_bytes(const simint_hip_pairs *p) { return p ? p->bytes : 0; }

}  // extern "C"

namespace {

// items for bra b = b0 + i: its ket range [lo, hi) (lo = ket_lo[i] or 0,
// hi = nket[i]) in tiles of 64; ket k is written at out_off[i] + (k - lo) * stride
__global__ void expand_prefix_items(int b0, int nbra, const int *ket_lo, const int *nket,
                                    const long long *tile_start, const long long *out_off, long long stride,
                                    WorkItem *items)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbra) return;
    const int lo = ket_lo ? ket_lo[i] : 0;
    const int hi = nket[i];
    const long long s = tile_start[i];
    const int t0 = lo / SIMINT_TILE;
    const int nt = hi > lo ? (hi + SIMINT_TILE - 1) / SIMINT_TILE - t0 : 0;
    for (int t = 0; t < nt; t++) {
        WorkItem w;
        w.bra = b0 + i;
        w.tile = t0 + t;
        w.lane_lo = max(lo - (t0 + t) * SIMINT_TILE, 0);
        w.lane_hi = min(SIMINT_TILE, hi - (t0 + t) * SIMINT_TILE);
        w.pad = 0;
        w.obase = out_off[i] - (long long)lo * stride;
        items[s + t] = w;
    }
}

}  // namespace

namespace {

// One range launch of class `am` over the pairs' primitive data (which do not
// depend on the AM the class kernel is compiled for); with `weight` 0..3 the
// 2a-weighted prefactors of that centre replace the plain ones.
struct RangeTerm {
    int am[4];
    simint_launch_fn fn;
    int weight;   // -1: none
};

// items_out / nitems_out (optional): the work items of the first term, left
// in the thread's item buffer for kernels queued after the class launch
long long range_launch(const simint_hip_pairs *P, const simint_hip_pairs *Q, int b0, int nbra, const int *ket_lo,
                       const int *ket_hi, const long long *const *out_offs, const RangeTerm *terms, int nterm,
                       double screen_tol, double *const *d_outs, const WorkItem **items_out = nullptr,
                       long long *nitems_out = nullptr, const long long **tstart_out = nullptr)
{
    if (nitems_out) *nitems_out = 0;
    if (nbra == 0) return 0;
    std::vector<long long> tstart(nbra + 1, 0);
    long long nq = 0;
    for (int i = 0; i < nbra; i++) {
        const int lo = ket_lo ? ket_lo[i] : 0, hi = ket_hi[i];
        if (lo < 0 || hi > Q->n) return SIMINT_ERROR_ARGUMENT;
        const long long nt = hi > lo ? (hi + SIMINT_TILE - 1) / SIMINT_TILE - lo / SIMINT_TILE : 0;
        tstart[i + 1] = tstart[i] + nt;
        nq += hi > lo ? hi - lo : 0;
    }
    const long long nitems = tstart[nbra];
    if (nitems == 0) return 0;
    ThreadCtx &c = ctx();
    Image im;
    const size_t o_lo = im.add(sizeof(int) * nbra);
    const size_t o_nk = im.add(sizeof(int) * nbra);
    const size_t o_ts = im.add(sizeof(long long) * nbra);
    std::vector<size_t> o_oo(nterm);
    for (int t = 0; t < nterm; t++) o_oo[t] = im.add(sizeof(long long) * nbra);
    if (ket_lo) memcpy(im.at<int>(o_lo), ket_lo, sizeof(int) * nbra);
    memcpy(im.at<int>(o_nk), ket_hi, sizeof(int) * nbra);
    memcpy(im.at<long long>(o_ts), tstart.data(), sizeof(long long) * nbra);
    for (int t = 0; t < nterm; t++) memcpy(im.at<long long>(o_oo[t]), out_offs[t], sizeof(long long) * nbra);
    hipError_t e;
    // the previous call's kernels may still read c.in / c.items: wait for them
    if ((e = hipStreamSynchronize(c.s)) != hipSuccess) { set_err("hipStreamSynchronize", e); return SIMINT_ERROR_HIP; }
    if ((e = c.stage.reserve(im.buf.size())) != hipSuccess) { set_err("hipHostMalloc", e); return SIMINT_ERROR_HIP; }
    if ((e = c.in.reserve(im.buf.size())) != hipSuccess) { set_err("hipMalloc", e); return SIMINT_ERROR_HIP; }
    if ((e = c.items.reserve(sizeof(WorkItem) * (size_t)nitems * nterm)) != hipSuccess) {
        set_err("hipMalloc", e);
        return SIMINT_ERROR_HIP;
    }
    size_t nst = 0;
    for (int t = 0; t < nterm; t++)
        nst = std::max(nst, (size_t)simint_dispatch_staging(terms[t].am[0], terms[t].am[1], terms[t].am[2],
                                                            terms[t].am[3]));
    if (nst && (e = c.split.reserve(sizeof(double) * nst * SIMINT_TILE * (size_t)nitems)) != hipSuccess) {
        set_err("hipMalloc", e);
        return SIMINT_ERROR_HIP;
    }
    memcpy(c.stage.p, im.buf.data(), im.buf.size());
    if ((e = hipMemcpyAsync(c.in.p, c.stage.p, im.buf.size(), hipMemcpyHostToDevice, c.s)) != hipSuccess) {
        set_err("hipMemcpyAsync", e);
        return SIMINT_ERROR_HIP;
    }
    char *base = static_cast<char *>(c.in.p);
    for (int t = 0; t < nterm; t++) {
        const RangeTerm &T = terms[t];
        WorkItem *items = static_cast<WorkItem *>(c.items.p) + (size_t)nitems * t;
        const long long stride = (long long)ncart(T.am[0]) * ncart(T.am[1]) * ncart(T.am[2]) * ncart(T.am[3]);
        hipLaunchKernelGGL(expand_prefix_items, dim3((nbra + 255) / 256), dim3(256), 0, c.s, b0, nbra,
                           ket_lo ? reinterpret_cast<const int *>(base + o_lo) : (const int *)nullptr,
                           reinterpret_cast<const int *>(base + o_nk),
                           reinterpret_cast<const long long *>(base + o_ts),
                           reinterpret_cast<const long long *>(base + o_oo[t]), stride, items);
        if ((e = hipGetLastError()) != hipSuccess) { set_err("expand launch", e); return SIMINT_ERROR_HIP; }
        EriArgs a;
        a.scratch = static_cast<double *>(c.split.p);
        a.bra = P->lin;
        a.ket = Q->til;
        if (T.weight == 0) a.bra.prefac = P->wa;
        if (T.weight == 1) a.bra.prefac = P->wb;
        if (T.weight == 2) a.ket.prefac = Q->twa;
        if (T.weight == 3) a.ket.prefac = Q->twb;
        a.items = items;
        a.nitems = nitems;
        a.out = d_outs[t];
        a.boys_grid = g_dev.boys;
        a.tol2 = (screen_tol > 0.0 && P->has_screen && Q->has_screen) ? screen_tol * screen_tol : 0.0;
        a.ket_screen_max = Q->screen_max;
        a.stride = stride;
        if ((e = T.fn(&a, c.s)) != hipSuccess) { set_err("class kernel launch", e); return SIMINT_ERROR_HIP; }
    }
    if (items_out) *items_out = static_cast<const WorkItem *>(c.items.p);
    if (nitems_out) *nitems_out = nitems;
    if (tstart_out) *tstart_out = reinterpret_cast<const long long *>(base + o_ts);
    return nq;
}

}  // namespace

extern "C" {

long long simint_hip_compute_eri_range(const simint_hip_pairs *P, const simint_hip_pairs *Q, int b0, int nbra,
                                       const int *ket_lo, const int *ket_hi, const long long *out_off,
                                       double screen_tol, int precision, double *d_out)
{
    if (!P || !Q || b0 < 0 || nbra < 0 || b0 + nbra > P->n || (precision != 0 && precision != 1))
        return SIMINT_ERROR_ARGUMENT;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    RangeTerm t{{P->am1, P->am2, Q->am1, Q->am2}, nullptr, -1};
    if (precision == 1) t.fn = simint_dispatch_mixed(P->am1, P->am2, Q->am1, Q->am2);
    if (!t.fn) t.fn = simint_dispatch(P->am1, P->am2, Q->am1, Q->am2);
    if (!t.fn) { t_err = "class not built"; return SIMINT_ERROR_UNSUPPORTED; }
    return range_launch(P, Q, b0, nbra, ket_lo, ket_hi, &out_off, &t, 1, screen_tol, &d_out);
}

long long simint_hip_compute_jk_range(const simint_hip_pairs *P, const simint_hip_pairs *Q, int b0, int nbra,
                                      const int *ket_lo, const int *ket_hi, const long long *out_off,
                                      double screen_tol, int precision, double *d_out, const simint_hip_jk *jk)
{
    if (!jk || jk->nbf <= 0 || !jk->density || !jk->bra_fun || !jk->ket_fun) return SIMINT_ERROR_ARGUMENT;
    if (!P || !Q || b0 < 0 || nbra < 0 || b0 + nbra > P->n || (precision != 0 && precision != 1))
        return SIMINT_ERROR_ARGUMENT;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    RangeTerm t{{P->am1, P->am2, Q->am1, Q->am2}, nullptr, -1};
    if (precision == 1) t.fn = simint_dispatch_mixed(P->am1, P->am2, Q->am1, Q->am2);
    if (!t.fn) t.fn = simint_dispatch(P->am1, P->am2, Q->am1, Q->am2);
    if (!t.fn) { t_err = "class not built"; return SIMINT_ERROR_UNSUPPORTED; }
    const WorkItem *items = nullptr;
    const long long *tstart = nullptr;
    long long nitems = 0;
    const long long nq = range_launch(P, Q, b0, nbra, ket_lo, ket_hi, &out_off, &t, 1, screen_tol, &d_out, &items,
                                      &nitems, &tstart);
    if (nq <= 0 || nitems == 0 || (!jk->J && !jk->K)) return nq;
    const hipError_t e = simint_rt::launch_jk_digest(t.am, items, nitems, tstart, nbra, Q->til.col, d_out, jk,
                                                     ctx().s);
    if (e != hipSuccess) { set_err("digest launch", e); return SIMINT_ERROR_HIP; }
    return nq;
}

long long simint_hip_compute_eri_prefix(const simint_hip_pairs *P, const simint_hip_pairs *Q, int b0, int nbra,
                                        const int *nket, const long long *out_off, double screen_tol,
                                        double *d_out)
{
    if (!P || !Q || b0 < 0 || nbra < 0 || b0 + nbra > P->n) return SIMINT_ERROR_ARGUMENT;
    for (int i = 0; i < nbra; i++)
        if (nket[i] < 0 || nket[i] > Q->n) return SIMINT_ERROR_ARGUMENT;
    return simint_hip_compute_eri_range(P, Q, b0, nbra, nullptr, nket, out_off, screen_tol, 0, d_out);
}

long long simint_hip_compute_eri_diag(const simint_hip_pairs *P, double *d_out)
{
    if (!P) return SIMINT_ERROR_ARGUMENT;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const simint_launch_fn fn = simint_dispatch(P->am1, P->am2, P->am1, P->am2);
    if (!fn) { t_err = "class not built"; return SIMINT_ERROR_UNSUPPORTED; }
    if (P->n == 0) return 0;
    ThreadCtx &c = ctx();
    std::vector<WorkItem> items(P->n);
    for (int b = 0; b < P->n; b++) {
        items[b].bra = b;
        items[b].tile = b / SIMINT_TILE;
        items[b].lane_lo = b % SIMINT_TILE;
        items[b].lane_hi = b % SIMINT_TILE + 1;
        items[b].obase = 0;
    }
    hipError_t e;
    if ((e = hipStreamSynchronize(c.s)) != hipSuccess) { set_err("hipStreamSynchronize", e); return SIMINT_ERROR_HIP; }
    const size_t bytes = sizeof(WorkItem) * items.size();
    if ((e = c.stage.reserve(bytes)) != hipSuccess) { set_err("hipHostMalloc", e); return SIMINT_ERROR_HIP; }
    if ((e = c.items.reserve(bytes)) != hipSuccess) { set_err("hipMalloc", e); return SIMINT_ERROR_HIP; }
    const size_t nst = (size_t)simint_dispatch_staging(P->am1, P->am2, P->am1, P->am2);
    if (nst && (e = c.split.reserve(sizeof(double) * nst * SIMINT_TILE * items.size())) != hipSuccess) {
        set_err("hipMalloc", e);
        return SIMINT_ERROR_HIP;
    }
    memcpy(c.stage.p, items.data(), bytes);
    if ((e = hipMemcpyAsync(c.items.p, c.stage.p, bytes, hipMemcpyHostToDevice, c.s)) != hipSuccess) {
        set_err("hipMemcpyAsync", e);
        return SIMINT_ERROR_HIP;
    }
    EriArgs a;
    a.scratch = static_cast<double *>(c.split.p);
    a.bra = P->lin;
    a.ket = P->til;
    a.items = static_cast<const WorkItem *>(c.items.p);
    a.nitems = P->n;
    a.out = d_out;
    a.boys_grid = g_dev.boys;
    a.tol2 = 0.0;
    a.ket_screen_max = P->screen_max;
    a.stride = ncart(P->am1) * ncart(P->am2) * ncart(P->am1) * ncart(P->am2);
    if ((e = fn(&a, c.s)) != hipSuccess) { set_err("class kernel launch", e); return SIMINT_ERROR_HIP; }
    return P->n;
}

}  // extern "C"

// ------------------------------------------------------------ quartet lists
namespace {

struct DerivPlan {
    int am[4];
    int skip;
    struct T { int am[4]; int centre, sign, weight; long long nc; size_t off; } t[8];
    int nt = 0;
    size_t scratch = 0;   // doubles per quartet
};

DerivPlan deriv1_plan(const int am[4])
{
    DerivPlan d;
    for (int c = 0; c < 4; c++) d.am[c] = am[c];
    const int mx = std::max(std::max(am[0], am[1]), std::max(am[2], am[3]));
    d.skip = 0;
    for (int c = 0; c < 4; c++)
        if (am[c] == mx) { d.skip = c; break; }
    for (int c = 0; c < 4; c++) {
        if (c == d.skip) continue;
        for (int sgn = +1; sgn >= -1; sgn -= 2) {
            if (sgn < 0 && am[c] == 0) continue;
            DerivPlan::T t;
            for (int k = 0; k < 4; k++) t.am[k] = am[k];
            t.am[c] += sgn;
            t.centre = c;
            t.sign = sgn;
            t.weight = sgn > 0;
            t.nc = (long long)ncart(t.am[0]) * ncart(t.am[1]) * ncart(t.am[2]) * ncart(t.am[3]);
            t.off = d.scratch;
            d.scratch += (size_t)t.nc;
            d.t[d.nt++] = t;
        }
    }
    return d;
}

EriListArgs list_args(const simint_hip_pairs *P, const simint_hip_pairs *Q, long long n, const int *qbra,
                      const int *qket, double screen_tol)
{
    EriListArgs a;
    a.bra = P->lin;
    a.ket = Q->lin;
    a.qbra = qbra;
    a.qket = qket;
    a.qout = nullptr;
    a.ostride = 0;
    a.nquartet = n;
    a.out = nullptr;
    a.boys_grid = g_dev.boys;
    a.tol2 = (screen_tol > 0.0 && P->has_screen && Q->has_screen) ? screen_tol * screen_tol : 0.0;
    a.ket_screen_max = Q->screen_max;
    a.split = 1;
    return a;
}

thread_local int t_list_split = 0;   // 0: by list length

// Lanes per quartet for a lane-per-quartet list launch of n quartets: enough
// that the launch has about 65536 lanes (a short list of one class otherwise
// occupies a few dozen wavefronts), at most 16.
int list_split(long long n)
{
    if (t_list_split > 0) return t_list_split;
    int s = 1;
    while (s < 16 && n * s < 65536) s *= 2;
    return s;
}

}  // namespace

extern "C" {

long long simint_hip_compute_eri_list(const simint_hip_pairs *P, const simint_hip_pairs *Q, long long nquartet,
                                      const int *d_qbra, const int *d_qket, const long long *d_qout,
                                      double screen_tol, double *d_out)
{
    if (!P || !Q || nquartet < 0 || (nquartet > 0 && (!d_qbra || !d_qket || !d_out))) return SIMINT_ERROR_ARGUMENT;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const simint_list_launch_fn fn = simint_dispatch_list(P->am1, P->am2, Q->am1, Q->am2);
    if (!fn) { t_err = "class not built"; return SIMINT_ERROR_UNSUPPORTED; }
    if (nquartet == 0) return 0;
    EriListArgs a = list_args(P, Q, nquartet, d_qbra, d_qket, screen_tol);
    a.qout = d_qout;
    a.ostride = (long long)ncart(P->am1) * ncart(P->am2) * ncart(Q->am1) * ncart(Q->am2);
    a.out = d_out;
    a.split = list_split(nquartet);
    hipError_t e = fn(&a, ctx().s);
    if (e != hipSuccess) { set_err("class kernel launch", e); return SIMINT_ERROR_HIP; }
    return nquartet;
}

long long simint_hip_deriv1_worksize(const simint_hip_pairs *P, const simint_hip_pairs *Q, long long nquartet)
{
    if (!P || !Q || nquartet < 0) return SIMINT_ERROR_ARGUMENT;
    const int am[4] = {P->am1, P->am2, Q->am1, Q->am2};
    return (long long)deriv1_plan(am).scratch * nquartet;
}

}  // extern "C"

namespace {

// the derivative list with its class blocks in d_work (a caller's buffer or
// the fixed scratch area); the device is set up
long long deriv1_list(const simint_hip_pairs *P, const simint_hip_pairs *Q, long long nquartet, const int *d_qbra,
                      const int *d_qket, const long long *d_qout, double screen_tol, double *d_work, double *d_out)
{
    if (!deriv1_built(P->am1, P->am2, Q->am1, Q->am2)) {
        t_err = "first derivatives not built for this class";
        return SIMINT_ERROR_UNSUPPORTED;
    }
    const int am[4] = {P->am1, P->am2, Q->am1, Q->am2};
    const DerivPlan dp = deriv1_plan(am);
    simint_list_launch_fn fns[8];
    for (int k = 0; k < dp.nt; k++) {
        const DerivPlan::T &t = dp.t[k];
        fns[k] = nullptr;
        if (t.am[0] <= kTermMaxAM && t.am[1] <= kTermMaxAM && t.am[2] <= kTermMaxAM && t.am[3] <= kTermMaxAM)
            fns[k] = simint_dispatch_list(t.am[0], t.am[1], t.am[2], t.am[3]);
        if (!fns[k]) { t_err = "class not built"; return SIMINT_ERROR_UNSUPPORTED; }
    }
    if (nquartet == 0) return 0;
    ThreadCtx &c = ctx();
    hipError_t e;
    Deriv1Args d;
    for (int q = 0; q < 4; q++) {
        d.am[q] = am[q];
        d.plus[q] = nullptr;
        d.minus[q] = nullptr;
    }
    for (int k = 0; k < dp.nt; k++) {
        const DerivPlan::T &t = dp.t[k];
        EriListArgs a = list_args(P, Q, nquartet, d_qbra, d_qket, screen_tol);
        if (t.weight) {
            if (t.centre == 0) a.bra.prefac = P->wa;
            if (t.centre == 1) a.bra.prefac = P->wb;
            if (t.centre == 2) a.ket.prefac = Q->wa;
            if (t.centre == 3) a.ket.prefac = Q->wb;
        }
        a.out = d_work + t.off * nquartet;
        a.ostride = t.nc;
        if ((e = fns[k](&a, c.s)) != hipSuccess) { set_err("class kernel launch", e); return SIMINT_ERROR_HIP; }
        if (t.sign > 0) d.plus[t.centre] = a.out;
        else d.minus[t.centre] = a.out;
    }
    d.skip = dp.skip;
    d.nquartet = nquartet;
    d.out = d_out;
    d.qout = d_qout;
    if ((e = launch_deriv1_assemble(d, c.s)) != hipSuccess) { set_err("deriv1 assembly", e); return SIMINT_ERROR_HIP; }
    return nquartet;
}

}  // namespace

extern "C" {

long long simint_hip_compute_eri_deriv1_list(const simint_hip_pairs *P, const simint_hip_pairs *Q,
                                             long long nquartet, const int *d_qbra, const int *d_qket,
                                             const long long *d_qout, double screen_tol, double *d_work,
                                             double *d_out)
{
    if (!P || !Q || nquartet < 0 || (nquartet > 0 && (!d_qbra || !d_qket || !d_out || !d_work)))
        return SIMINT_ERROR_ARGUMENT;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    return deriv1_list(P, Q, nquartet, d_qbra, d_qket, d_qout, screen_tol, d_work, d_out);
}

long long simint_hip_compute_eri_deriv1_list_noalloc(const simint_hip_pairs *P, const simint_hip_pairs *Q,
                                                     long long nquartet, const int *d_qbra, const int *d_qket,
                                                     const long long *d_qout, double screen_tol, double *d_out)
{
    if (!P || !Q || nquartet < 0 || (nquartet > 0 && (!d_qbra || !d_qket || !d_out))) return SIMINT_ERROR_ARGUMENT;
    // no set-up here (it allocates): simint_init has looked up the array
    if (!g_dev.ready.load(std::memory_order_acquire)) {
        t_err = "call simint_init before simint_hip_compute_eri_deriv1_list_noalloc";
        return SIMINT_ERROR_ARGUMENT;
    }
    if (hipSetDevice(g_dev.id) != hipSuccess) return SIMINT_ERROR_NO_DEVICE;
    const int am[4] = {P->am1, P->am2, Q->am1, Q->am2};
    if (!deriv1_built(am[0], am[1], am[2], am[3])) {
        t_err = "first derivatives not built for this class";
        return SIMINT_ERROR_UNSUPPORTED;
    }
    const unsigned long long need = (unsigned long long)deriv1_plan(am).scratch * (unsigned long long)nquartet;
    if (need > kNoAllocScratchDoubles) {
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "derivative list needs %llu doubles of scratch; the fixed area holds %llu (split the list)",
                 need, (unsigned long long)kNoAllocScratchDoubles);
        t_err = buf;
        return SIMINT_ERROR_ARGUMENT;
    }
    return deriv1_list(P, Q, nquartet, d_qbra, d_qket, d_qout, screen_tol, g_dev.scratch, d_out);
}

long long simint_hip_compute_eri_deriv1_range(const simint_hip_pairs *P, const simint_hip_pairs *Q, int b0,
                                              int nbra, const int *ket_lo, const int *ket_hi, double screen_tol,
                                              double *d_work, double *d_out)
{
    if (!P || !Q || b0 < 0 || nbra < 0 || b0 + nbra > P->n || !d_work || !d_out) return SIMINT_ERROR_ARGUMENT;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    if (!deriv1_built(P->am1, P->am2, Q->am1, Q->am2)) {
        t_err = "first derivatives not built for this class";
        return SIMINT_ERROR_UNSUPPORTED;
    }
    const int am[4] = {P->am1, P->am2, Q->am1, Q->am2};
    const DerivPlan dp = deriv1_plan(am);
    RangeTerm terms[8];
    for (int k = 0; k < dp.nt; k++) {
        const DerivPlan::T &t = dp.t[k];
        terms[k] = RangeTerm{{t.am[0], t.am[1], t.am[2], t.am[3]}, nullptr, t.weight ? t.centre : -1};
        if (t.am[0] <= kTermMaxAM && t.am[1] <= kTermMaxAM && t.am[2] <= kTermMaxAM && t.am[3] <= kTermMaxAM)
            terms[k].fn = simint_dispatch(t.am[0], t.am[1], t.am[2], t.am[3]);
        if (!terms[k].fn) { t_err = "class not built"; return SIMINT_ERROR_UNSUPPORTED; }
    }
    // dense quartet numbering in (bra, ket) order
    std::vector<long long> qpre(nbra + 1, 0);
    for (int i = 0; i < nbra; i++) {
        const int lo = ket_lo ? ket_lo[i] : 0, hi = ket_hi[i];
        if (lo < 0 || hi > Q->n) return SIMINT_ERROR_ARGUMENT;
        qpre[i + 1] = qpre[i] + (hi > lo ? hi - lo : 0);
    }
    const long long nq = qpre[nbra];
    if (nq == 0) return 0;
    std::vector<std::vector<long long>> offs(dp.nt, std::vector<long long>(nbra));
    const long long *offp[8];
    double *outs[8];
    for (int k = 0; k < dp.nt; k++) {
        for (int i = 0; i < nbra; i++) offs[k][i] = qpre[i] * dp.t[k].nc;
        offp[k] = offs[k].data();
        outs[k] = d_work + (long long)dp.t[k].off * nq;
    }
    long long r = range_launch(P, Q, b0, nbra, ket_lo, ket_hi, offp, terms, dp.nt, screen_tol, outs);
    if (r < 0) return r;
    Deriv1Args d;
    for (int q = 0; q < 4; q++) {
        d.am[q] = am[q];
        d.plus[q] = nullptr;
        d.minus[q] = nullptr;
    }
    for (int k = 0; k < dp.nt; k++) {
        if (dp.t[k].sign > 0) d.plus[dp.t[k].centre] = outs[k];
        else d.minus[dp.t[k].centre] = outs[k];
    }
    d.skip = dp.skip;
    d.nquartet = nq;
    d.out = d_out;
    d.qout = nullptr;
    hipError_t e = launch_deriv1_assemble(d, ctx().s);
    if (e != hipSuccess) { set_err("deriv1 assembly", e); return SIMINT_ERROR_HIP; }
    return nq;
}

// ------------------------------------------------------------ graph capture
struct simint_hip_graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
};

int simint_hip_capture_begin(void)
{
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    hipError_t e = hipStreamBeginCapture(ctx().s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { set_err("hipStreamBeginCapture", e); return SIMINT_ERROR_HIP; }
    return 0;
}

int simint_hip_capture_end(simint_hip_graph **out)
{
    if (!out) return SIMINT_ERROR_ARGUMENT;
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    simint_hip_graph *g = new simint_hip_graph;
    hipError_t e = hipStreamEndCapture(ctx().s, &g->g);
    if (e != hipSuccess) { delete g; set_err("hipStreamEndCapture", e); return SIMINT_ERROR_HIP; }
    e = hipGraphInstantiate(&g->x, g->g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g->g);
        delete g;
        set_err("hipGraphInstantiate", e);
        return SIMINT_ERROR_HIP;
    }
    *out = g;
    return 0;
}

int simint_hip_graph_launch(simint_hip_graph *g)
{
    if (!g) return SIMINT_ERROR_ARGUMENT;
    hipError_t e = hipGraphLaunch(g->x, ctx().s);
    if (e != hipSuccess) { set_err("hipGraphLaunch", e); return SIMINT_ERROR_HIP; }
    return 0;
}

void simint_hip_graph_free(simint_hip_graph *g)
{
    if (!g) return;
    if (g->x) hipGraphExecDestroy(g->x);
    if (g->g) hipGraphDestroy(g->g);
    delete g;
}

int simint_hip_synchronize(void)
{
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    hipError_t e = hipStreamSynchronize(ctx().s);
    if (e != hipSuccess) { set_err("hipStreamSynchronize", e); return SIMINT_ERROR_HIP; }
    return 0;
}

void *simint_hip_stream(void)
{
    if (!ensure_device()) return nullptr;
    return (void *)ctx().s;
}

int simint_hip_set_list_split(int lanes)
{
    if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16) return SIMINT_ERROR_ARGUMENT;
    const int prev = t_list_split;
    t_list_split = lanes;
    return prev;
}

int simint_hip_set_stream(void *stream)
{
    if (!ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    ThreadCtx &c = ctx();
    c.s = stream ? static_cast<hipStream_t>(stream) : c.own;
    return 0;
}

}  // extern "C"

// ------------------------------------------------ services for other sources
namespace simint_rt {

bool ensure_device() { return ::ensure_device(); }

const double *boys_grid() { return g_dev.boys; }

hipStream_t stream() { return ctx().s; }

hipError_t thread_buffers(size_t host_bytes, size_t in_bytes, size_t out_bytes, void **host, void **din,
                          void **dout)
{
    ThreadCtx &c = ctx();
    hipError_t e;
    if ((e = hipStreamSynchronize(c.s)) != hipSuccess) return e;
    if ((e = c.stage.reserve(host_bytes)) != hipSuccess) return e;
    if ((e = c.in.reserve(in_bytes)) != hipSuccess) return e;
    if ((e = c.out.reserve(out_bytes)) != hipSuccess) return e;
    *host = c.stage.p;
    *din = c.in.p;
    *dout = c.out.p;
    return hipSuccess;
}

hipError_t to_drop_in_order(int la, int lb, const double *in, double *out, hipStream_t s, bool *moved)
{
    const int am[4] = {la, lb, 0, 0};
    *moved = gamess_needed(am);
    if (!*moved) return hipSuccess;
    return launch_to_gamess(am, 1, 1, in, out, s);
}

void set_error(const char *msg) { t_err = msg; }

void set_hip_error(const char *what, hipError_t e) { set_err(what, e); }

}  // namespace simint_rt

/*
 * oracle.c -- CPU restatement of simint's electron-repulsion-integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load the library
 * built from it (liboracle.so).  The product (libsimint.so and its HIP
 * kernels) never links, loads or calls anything here, and the product fails
 * loudly rather than falling back to it.
 *
 * What it restates (reference paths relative to the simint-generator tree):
 *   - Boys function split scheme         skel/simint/boys/boys.h:14-92,
 *                                         boys_taylor.h:14-40, boys_long.h:14-58
 *   - Boys grid (x_i = 0.1 i, i<366, F_0..F_31), computed here from the
 *     positive series F_m(x) = e^-x sum_k (2x)^k / prod_{j<=k}(2m+2j+1) in
 *     long double (the reference tabulates the same function with mpmath,
 *     python/boys/gen_boys_shortgrid.py:32-44)
 *   - shell normalisation                skel/simint/shell/shell.c:248-278
 *   - shell-pair (Gaussian product) fill skel/simint/shell/shell.c:425-582
 *   - fast-Schwarz primitive screen      skel/simint/shell/shell_screen.c:226-270
 *   - Schwarz primitive screen           skel/simint/shell/shell_screen.c:126-223
 *   - per-primitive-quartet setup        generator/ostei/OSTEI_Writer.cpp:595-702
 *   - VRR step selection (Makowski)      generator/ostei/Makowski.hpp:96-241,
 *                                         OSTEI_VRR_Algorithm_Base.cpp:289-469
 *   - VRR arithmetic                     generator/ostei/OSTEI_VRR_Writer.cpp:52-125
 *   - HRR step selection (Makowski)      generator/ostei/Makowski.hpp:26-82,
 *                                         OSTEI_HRR_Algorithm_Base.cpp:160-283
 *   - HRR arithmetic                     generator/ostei/OSTEI_HRR_Writer.cpp:22-99
 *   - primitive screening in the kernel  generator/ostei/OSTEI_Writer.cpp:504-587
 *     (SIMD length 1 semantics: every primitive quartet tested on its own)
 *   - output layout (Psi4 order)         generator/Ordering_Psi4.cpp:16-40,
 *                                         OSTEI_HRR_Writer.cpp:143-146
 *   - first-derivative assembly          generator/ostei/OSTEIDeriv1_Writer.cpp:334-537
 *
 * The VRR is evaluated on demand: each needed (component quartet, m) value is
 * produced by the one Obara-Saika step the reference generator assigns to
 * that component (the first centre in the Makowski order whose angular
 * momentum is non-zero, stepping in the direction of the lowest non-zero
 * exponent, rightmost on ties).  Values are identical to evaluating the
 * reference's step lists; only unneeded (component, m) pairs are skipped.
 */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define OR_MAXL 10          /* max AM on one centre inside the VRR */
#define OR_NG   286         /* number of Cartesians with l <= OR_MAXL: 11*12*13/6 */
#define BOYS_NPT 366
#define BOYS_NCOL 32
#define BOYS_MAXX 36.5

/* -------------------------------------------------------------------------
 * Cartesian bookkeeping (Psi4 order: x from l down, then y from l-x down)
 * generator/Ordering_Psi4.cpp:16-40, skel/test/Common.cpp:20-33
 * ------------------------------------------------------------------------- */
static int g_ex[OR_NG][3];
static int g_l[OR_NG];

static inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static inline int cart_index(int ix, int iy, int iz)
{
    int l = ix + iy + iz;
    return (l - ix) * (l - ix + 1) / 2 + (l - ix - iy);
}
static inline int gid(int ix, int iy, int iz)
{
    int l = ix + iy + iz;
    return l * (l + 1) * (l + 2) / 6 + cart_index(ix, iy, iz);
}

/* -------------------------------------------------------------------------
 * Boys function
 * ------------------------------------------------------------------------- */
static double g_grid[BOYS_NPT][BOYS_NCOL];
static double g_longfac[BOYS_NCOL];

static long double boys_series_ld(int m, long double x)
{
    /* F_m(x) = e^-x * sum_k (2x)^k / ((2m+1)(2m+3)...(2m+2k+1)); all terms
     * positive, so long double keeps ~1e-18 relative accuracy. */
    long double term = 1.0L / (2 * m + 1);
    long double sum = term;
    for (int k = 1; k < 2000; k++) {
        term *= 2.0L * x / (2 * m + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    return expl(-x) * sum;
}

static void init_tables(void)
{
    int id = 0;
    for (int l = 0; l <= OR_MAXL; l++)
        for (int ix = l; ix >= 0; ix--)
            for (int iy = l - ix; iy >= 0; iy--) {
                g_ex[id][0] = ix; g_ex[id][1] = iy; g_ex[id][2] = l - ix - iy;
                g_l[id] = l;
                id++;
            }
    for (int i = 0; i < BOYS_NPT; i++) {
        long double x = (long double)i / 10.0L;
        for (int m = 0; m < BOYS_NCOL; m++)
            g_grid[i][m] = (double)boys_series_ld(m, x);
    }
    /* longfac[n] = sqrt(pi) (2n-1)!! / 2^(n+1)  (skel/simint/boys/boys_longfac.h:13-48) */
    long double f = sqrtl(3.14159265358979323846264338327950288L) / 2.0L;
    for (int n = 0; n < BOYS_NCOL; n++) {
        g_longfac[n] = (double)f;
        f *= (2 * n + 1) / 2.0L;
    }
}

static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static void ensure_init(void) { pthread_once(&g_once, init_tables); }

/* boys_taylor.h:14-40 / :43-67 */
static double boys_taylor_single(double x, int n)
{
    const int idx = (int)(10.0 * (x + 0.05));
    const double xi = (double)idx * 0.1;
    const double dx = xi - x;
    const double *g = &g_grid[idx][n];
    return g[0] + dx * (g[1] + dx * ((1.0 / 2.0) * g[2] + dx * ((1.0 / 6.0) * g[3]
           + dx * ((1.0 / 24.0) * g[4] + dx * ((1.0 / 120.0) * g[5]
           + dx * ((1.0 / 720.0) * g[6] + dx * ((1.0 / 5040.0) * g[7])))))));
}

/* boys_long.h:14-27 (all orders) and :29-35 (single order) */
static void boys_long_all(double *F, double x, int n)
{
    const double x1 = 1.0 / x;
    double x2 = sqrt(x1);
    for (int i = 0; i <= n; i++) {
        F[i] = g_longfac[i] * x2;
        x2 *= x1;
    }
}
static double boys_long_single(double x, int n)
{
    const double p = -(2 * n + 1);
    return g_longfac[n] * sqrt(pow(x, p));
}

/* boys.h:14-92 (scalar lane semantics) */
static void boys_F_split(double *F, double x, int n)
{
    if (n < 4) {
        if (x < BOYS_MAXX)
            for (int i = 0; i <= n; i++) F[i] = boys_taylor_single(x, i);
        else
            boys_long_all(F, x, n);
        return;
    }
    if (x < BOYS_MAXX)
        F[n] = boys_taylor_single(x, n);
    else
        F[n] = boys_long_single(x, n);
    const double x2 = 2.0 * x;
    const double ex = exp(-x);
    for (int n2 = n - 1; n2 >= 0; n2--) {
        const double den = 1.0 / (2.0 * n2 + 1);
        F[n2] = den * fma(x2, F[n2 + 1], ex);
    }
}

/* -------------------------------------------------------------------------
 * Exported table / Boys access (for tests that pin the tables)
 * ------------------------------------------------------------------------- */
void oracle_boys_grid(double *out /* 366*32 */)
{
    ensure_init();
    memcpy(out, g_grid, sizeof(g_grid));
}
void oracle_boys_longfac(double *out /* 32 */)
{
    ensure_init();
    memcpy(out, g_longfac, sizeof(g_longfac));
}
void oracle_boys(double x, int n, double *F)
{
    ensure_init();
    boys_F_split(F, x, n);
}
double oracle_boys_ld(int m, double x)
{
    ensure_init();
    return (double)boys_series_ld(m, (long double)x);
}

/* -------------------------------------------------------------------------
 * Shells
 * ------------------------------------------------------------------------- */
typedef struct {
    int am, nprim;
    double x, y, z;
    const double *alpha;
    const double *coef;
} or_shell;

/* norm_fac[l] = pi^(3/2) (2l-1)!! / 2^l   (skel/simint/shell/shell_constants.c,
 * python/gen_shellconstants.py:36-42) */
static double norm_fac(int l)
{
    long double v = powl(3.14159265358979323846264338327950288L, 1.5L);
    for (int k = 1; k <= 2 * l - 1; k += 2) v *= k;
    for (int k = 0; k < l; k++) v *= 0.5L;
    return (double)v;
}
double oracle_norm_fac(int l) { return norm_fac(l); }

/* shell.c:248-278 */
void oracle_normalize(int n, const int *am, const int *nprim, const int *prim_off,
                      const double *alpha, double *coef)
{
    for (int i = 0; i < n; ++i) {
        const double dam = (double)am[i];
        const double m = dam + 1.5;
        const double m2 = 0.5 * m;
        const double *a = alpha + prim_off[i];
        double *c = coef + prim_off[i];
        double sum = 0.0;
        for (int j = 0; j < nprim[i]; j++)
            for (int k = 0; k < nprim[i]; k++)
                sum += (c[j] * c[k] * pow(a[j] * a[k], m2)) / (pow(a[j] + a[k], m));
        const double norm = 1.0 / sqrt(sum * norm_fac(am[i]));
        for (int j = 0; j < nprim[i]; ++j) c[j] *= norm * pow(a[j], m2);
    }
}

/* -------------------------------------------------------------------------
 * Primitive-pair data for one shell pair (shell.c:425-538)
 * ------------------------------------------------------------------------- */
typedef struct {
    int n;
    double *x, *y, *z, *PAx, *PAy, *PAz, *PBx, *PBy, *PBz, *alpha, *prefac, *alpha2, *beta2, *screen;
    double ABx, ABy, ABz;
    double buf[];
} or_pair;

#define SQRT_TWO_PI_52 5.91496717279561287782347843505365022963403204272156069835237
#define SQRT_TWO_TIMES_PI_52 24.739429451193148050228357586821944304008810699990039583

/* shell_screen.c:226-270 : (ij|ij) of the s-type primitive pair, not square-rooted */
static double fastschwarz_value(const or_shell *A, const or_shell *B, int i, int j)
{
    const double a = A->alpha[i], b = B->alpha[j];
    const double p = a + b;
    const double oop = 1.0 / p;
    const double rho = (a * b) / p;
    const double Rx = A->x - B->x, Ry = A->y - B->y, Rz = A->z - B->z;
    const double R2 = Rx * Rx + Ry * Ry + Rz * Rz;
    double G = SQRT_TWO_TIMES_PI_52 * pow(oop, 2.5) * exp(-2.0 * rho * R2);
    G *= A->coef[i] * A->coef[i] * B->coef[j] * B->coef[j];
    return G;
}

static int eri_quartet_internal(const or_shell *S[4], const or_pair *P, const or_pair *Q,
                                double tol2, int deriv, double *out);

/* shell.c:37-71 -- quicksort of one pair's primitives by screen value,
 * descending; every per-primitive array moves with its key */
static void pair_swap(or_pair *P, int i, int j)
{
    double *f[14] = {P->x, P->y, P->z, P->PAx, P->PAy, P->PAz, P->PBx, P->PBy, P->PBz,
                     P->alpha, P->prefac, P->screen, P->alpha2, P->beta2};
    for (int k = 0; k < 14; k++) { double t = f[k][i]; f[k][i] = f[k][j]; f[k][j] = t; }
}
static void pair_quicksort(or_pair *P, int left, int right)
{
    int i = left, j = right;
    const double mid = P->screen[(i + j) / 2];
    while (i <= j) {
        while (P->screen[i] > mid) i++;
        while (P->screen[j] < mid) j--;
        if (i <= j) { pair_swap(P, i, j); i++; j--; }
    }
    if (i < right) pair_quicksort(P, i, right);
    if (j > left) pair_quicksort(P, left, j);
}

static or_pair *make_pair(const or_shell *A, const or_shell *B, int same, int screen_method)
{
    const int n = same ? A->nprim * (A->nprim + 1) / 2 : A->nprim * B->nprim;
    or_pair *P = (or_pair *)calloc(1, sizeof(or_pair) + sizeof(double) * 14 * (size_t)(n > 0 ? n : 1));
    P->n = n;
    double *b = P->buf;
    P->x = b; b += n; P->y = b; b += n; P->z = b; b += n;
    P->PAx = b; b += n; P->PAy = b; b += n; P->PAz = b; b += n;
    P->PBx = b; b += n; P->PBy = b; b += n; P->PBz = b; b += n;
    P->alpha = b; b += n; P->prefac = b; b += n; P->alpha2 = b; b += n; P->beta2 = b; b += n;
    P->screen = b;
    const double Xab_x = A->x - B->x, Xab_y = A->y - B->y, Xab_z = A->z - B->z;
    const double Xab = Xab_x * Xab_x + Xab_y * Xab_y + Xab_z * Xab_z;
    P->ABx = Xab_x; P->ABy = Xab_y; P->ABz = Xab_z;
    int idx = 0;
    for (int i = 0; i < A->nprim; ++i) {
        const double ai = A->alpha[i];
        const double AxAa = A->x * ai, AyAa = A->y * ai, AzAa = A->z * ai;
        const int jend = same ? i + 1 : B->nprim;
        for (int j = 0; j < jend; ++j) {
            const double aj = B->alpha[j];
            const double ab_sum = ai + aj;
            const double ab_mul = ai * aj;
            P->prefac[idx] = A->coef[i] * B->coef[j] * exp(-Xab * ab_mul / ab_sum) * SQRT_TWO_PI_52 / ab_sum;
            if (same && i != j) P->prefac[idx] *= 2.0;
            P->x[idx] = (AxAa + aj * B->x) / ab_sum;
            P->y[idx] = (AyAa + aj * B->y) / ab_sum;
            P->z[idx] = (AzAa + aj * B->z) / ab_sum;
            P->PAx[idx] = P->x[idx] - A->x; P->PAy[idx] = P->y[idx] - A->y; P->PAz[idx] = P->z[idx] - A->z;
            P->PBx[idx] = P->x[idx] - B->x; P->PBy[idx] = P->y[idx] - B->y; P->PBz[idx] = P->z[idx] - B->z;
            P->alpha[idx] = ab_sum;
            if (same && i != j) { P->alpha2[idx] = ai + aj; P->beta2[idx] = ai + aj; }
            else { P->alpha2[idx] = 2.0 * ai; P->beta2[idx] = 2.0 * aj; }
            if (screen_method == 2) P->screen[idx] = fastschwarz_value(A, B, i, j);
            ++idx;
        }
    }
    if (screen_method == 1) {
        /* shell_screen.c:126-223 -- max diagonal element of (ij|ij) over the
         * uncontracted primitive pair; for a same-shell pair the loop runs
         * over the triangle and the (unnormalised) coefficients are used as is */
        idx = 0;
        for (int i = 0; i < A->nprim; ++i) {
            const int jend = same ? i + 1 : B->nprim;
            for (int j = 0; j < jend; ++j) {
                or_shell a1 = {A->am, 1, A->x, A->y, A->z, A->alpha + i, A->coef + i};
                or_shell b1 = {B->am, 1, B->x, B->y, B->z, B->alpha + j, B->coef + j};
                or_pair *pp = make_pair(&a1, &b1, 0, 0);
                const int n1 = ncart(A->am), n2 = ncart(B->am);
                const int n12 = n1 * n2;
                double *ints = (double *)malloc(sizeof(double) * (size_t)n12 * n12);
                const or_shell *S[4] = {&a1, &b1, &a1, &b1};
                eri_quartet_internal(S, pp, pp, 0.0, 0, ints);
                double mx = 0.0;
                for (int m = 0; m < n1; m++)
                    for (int nn = 0; nn < n2; nn++) {
                        const int k = (m * n2 + nn) * n12 + (m * n2 + nn);
                        const double v = fabs(ints[k]);
                        if (v > mx) mx = v;
                    }
                P->screen[idx++] = mx;
                free(ints);
                free(pp);
            }
        }
    }
    /* shell.c:579-581, :74-79 */
    if (screen_method && n > 1) pair_quicksort(P, 0, n - 1);
    return P;
}

/* -------------------------------------------------------------------------
 * Recurrence plan for one class (la lb | lc ld), built on demand
 * ------------------------------------------------------------------------- */
enum { ST_BASE = -1, ST_I = 0, ST_J = 1, ST_K = 2, ST_L = 3 };

typedef struct {
    int step;       /* ST_* */
    int dir;        /* 0,1,2 */
    int m;
    int src[8];     /* slots, -1 if absent */
    int ijkl[4];    /* integer factors, OSTEI_VRR_Algorithm_Base.cpp:242-285 */
} vrr_node;

typedef struct { uint64_t key; int val; } hslot;
typedef struct { hslot *t; int cap, n; } hmap;

static void hm_init(hmap *h, int cap)
{
    h->cap = cap; h->n = 0;
    h->t = (hslot *)malloc(sizeof(hslot) * (size_t)cap);
    for (int i = 0; i < cap; i++) h->t[i].val = -1;
}
static int hm_get(const hmap *h, uint64_t k)
{
    uint64_t i = (k * 0x9E3779B97F4A7C15ull) >> 20;
    for (;;) {
        i &= (uint64_t)(h->cap - 1);
        if (h->t[i].val < 0) return -1;
        if (h->t[i].key == k) return h->t[i].val;
        i++;
    }
}
static void hm_grow(hmap *h);
static void hm_put(hmap *h, uint64_t k, int v)
{
    if (2 * (h->n + 1) > h->cap) hm_grow(h);
    uint64_t i = (k * 0x9E3779B97F4A7C15ull) >> 20;
    for (;;) {
        i &= (uint64_t)(h->cap - 1);
        if (h->t[i].val < 0) { h->t[i].key = k; h->t[i].val = v; h->n++; return; }
        if (h->t[i].key == k) { h->t[i].val = v; return; }
        i++;
    }
}
static void hm_grow(hmap *h)
{
    hmap n2;
    hm_init(&n2, h->cap * 2);
    for (int i = 0; i < h->cap; i++)
        if (h->t[i].val >= 0) hm_put(&n2, h->t[i].key, h->t[i].val);
    free(h->t);
    *h = n2;
}

typedef struct {
    int q[4];       /* Gaussian ids a,b,c,d */
} quartet;

static inline uint64_t qkey(const int q[4], int m)
{
    return (uint64_t)q[0] | ((uint64_t)q[1] << 9) | ((uint64_t)q[2] << 18) | ((uint64_t)q[3] << 27)
           | ((uint64_t)m << 36);
}

typedef struct {
    int la, lb, lc, ld;
    int L;                   /* la+lb+lc+ld (+1 for derivative plans) */
    int order[4];            /* Makowski centre order (bottom of the list reduced first) */
    int nnode, capnode;
    vrr_node *node;
    hmap map;
} vrr_plan;

/* Makowski.hpp:111-117: lowest non-zero exponent, rightmost on ties */
static int mk_dir(const int e[3])
{
    int best = -1;
    for (int k = 2; k >= 0; k--)
        if (e[k] > 0 && (best < 0 || e[k] < e[best])) best = k;
    return best;
}

static int gid_step(int g, int dir, int delta)
{
    int e[3] = {g_ex[g][0], g_ex[g][1], g_ex[g][2]};
    e[dir] += delta;
    if (e[dir] < 0) return -1;
    return gid(e[0], e[1], e[2]);
}

/* returns slot for (q, m), creating the node (and its sources) if needed */
static int vrr_need(vrr_plan *pl, const int q[4], int m)
{
    const uint64_t k = qkey(q, m);
    int s = hm_get(&pl->map, k);
    if (s >= 0) return s;

    vrr_node nd;
    memset(&nd, 0, sizeof(nd));
    nd.m = m;
    for (int i = 0; i < 8; i++) nd.src[i] = -1;

    int centre = -1;
    for (int i = 0; i < 4; i++)
        if (g_l[q[pl->order[i]]] > 0) { centre = pl->order[i]; break; }

    if (centre < 0) {
        nd.step = ST_BASE;
    } else {
        /* Makowski.hpp:96-241 -- the eight sources of an Obara-Saika step,
         * in the reference's slot order:
         *   I: s2 = a-2,      s4 = a-1,b-1   s6 = a-1,c-1   s7 = a-1,d-1
         *   J: s2 = a-1,b-1   s4 = b-2       s6 = b-1,c-1   s7 = b-1,d-1
         *   K: s2 = c-2       s4 = c-1,d-1   s6 = a-1,c-1   s7 = b-1,c-1
         *   L: s2 = c-1,d-1   s4 = d-2       s6 = a-1,d-1   s7 = b-1,d-1
         * integer factors (OSTEI_VRR_Algorithm_Base.cpp:242-285) follow the
         * slot: ijkl[0] for s2/s3, ijkl[1] for s4/s5, ijkl[2] s6, ijkl[3] s7 */
        nd.step = centre;
        const int dir = mk_dir(g_ex[q[centre]]);
        nd.dir = dir;
        const int partner = (centre == ST_I) ? ST_J : (centre == ST_J) ? ST_I : (centre == ST_K) ? ST_L : ST_K;
        const int o1 = (centre <= ST_J) ? ST_K : ST_I;
        const int o2 = (centre <= ST_J) ? ST_L : ST_J;
        int s0[4], sc2[4], spa[4], s6[4], s7[4];
        int vc2 = 1, vpa = 1, v6 = 1, v7 = 1;
        memcpy(s0, q, sizeof(s0));
        s0[centre] = gid_step(q[centre], dir, -1);
        memcpy(sc2, q, sizeof(sc2));                 /* centre - 2 */
        sc2[centre] = gid_step(q[centre], dir, -2);
        if (sc2[centre] < 0) vc2 = 0;
        memcpy(spa, s0, sizeof(spa));                /* centre - 1, partner - 1 */
        spa[partner] = gid_step(q[partner], dir, -1);
        if (spa[partner] < 0) vpa = 0;
        memcpy(s6, s0, sizeof(s6));                  /* centre - 1, other side - 1 */
        s6[o1] = gid_step(q[o1], dir, -1);
        if (s6[o1] < 0) v6 = 0;
        memcpy(s7, s0, sizeof(s7));
        s7[o2] = gid_step(q[o2], dir, -1);
        if (s7[o2] < 0) v7 = 0;
        const int fc2 = g_ex[q[centre]][dir] - 1;
        const int fpa = g_ex[q[partner]][dir];
        const int centre_first = (centre == ST_I || centre == ST_K);
        nd.src[0] = vrr_need(pl, s0, m);
        nd.src[1] = vrr_need(pl, s0, m + 1);
        if (centre_first) {
            if (vc2) { nd.src[2] = vrr_need(pl, sc2, m); nd.src[3] = vrr_need(pl, sc2, m + 1); }
            if (vpa) { nd.src[4] = vrr_need(pl, spa, m); nd.src[5] = vrr_need(pl, spa, m + 1); }
            nd.ijkl[0] = fc2; nd.ijkl[1] = fpa;
        } else {
            if (vpa) { nd.src[2] = vrr_need(pl, spa, m); nd.src[3] = vrr_need(pl, spa, m + 1); }
            if (vc2) { nd.src[4] = vrr_need(pl, sc2, m); nd.src[5] = vrr_need(pl, sc2, m + 1); }
            nd.ijkl[0] = fpa; nd.ijkl[1] = fc2;
        }
        nd.ijkl[2] = g_ex[q[o1]][dir];
        nd.ijkl[3] = g_ex[q[o2]][dir];
        if (v6) nd.src[6] = vrr_need(pl, s6, m + 1);
        if (v7) nd.src[7] = vrr_need(pl, s7, m + 1);
    }
    if (pl->nnode == pl->capnode) {
        pl->capnode = pl->capnode ? 2 * pl->capnode : 256;
        pl->node = (vrr_node *)realloc(pl->node, sizeof(vrr_node) * (size_t)pl->capnode);
    }
    s = pl->nnode++;
    pl->node[s] = nd;
    hm_put(&pl->map, k, s);
    return s;
}

/* OSTEI_VRR_Algorithm_Base.cpp:449-469: centre order = stable sort of
 * {3,2,1,0} by the maximum AM of that centre over the VRR targets */
static void vrr_order(const int qam_max[4], int order[4])
{
    int o[4] = {3, 2, 1, 0};
    for (int i = 1; i < 4; i++) {       /* insertion sort is stable */
        int v = o[i], j = i - 1;
        while (j >= 0 && qam_max[o[j]] > qam_max[v]) { o[j + 1] = o[j]; j--; }
        o[j + 1] = v;
    }
    memcpy(order, o, sizeof(o));
}

/* -------------------------------------------------------------------------
 * HRR (OSTEI_HRR_Algorithm_Base.cpp:160-283, Makowski.hpp:26-82)
 *   bra: J (move I->J) unless la < lb, then I;  ket: L unless lc < ld, then K
 *   J: (a,b) = (a+1_i, b-1_i) + AB_i (a, b-1_i), i from b
 *   I: (a,b) = (a-1_i, b+1_i) - AB_i (a-1_i, b), i from a
 *   ket the same with CD
 * ------------------------------------------------------------------------- */
typedef struct {
    int la, lb, lc, ld;
    int bra_j, ket_l;                 /* 1 if J / L direction */
    /* top AM ranges for VRR targets: bra (e,0) if bra_j else (0,e); ket likewise */
    int e_lo, e_hi, f_lo, f_hi;
    vrr_plan vrr;
    /* contracted top storage: index of (e,f) components */
    int ntop;
    int *top_slot;                    /* VRR slot at m=0 for each top value */
    int *top_e, *top_f;               /* Gaussian ids of the non-zero centre */
} class_plan;

static int top_index(const class_plan *cp, int ge, int gf)
{
    /* tops are ordered by e (all AMs e_lo..e_hi, Psi4 order) then f */
    const int le = g_l[ge], lf = g_l[gf];
    int ei = 0, fi = 0, nf = 0;
    for (int l = cp->e_lo; l < le; l++) ei += ncart(l);
    ei += cart_index(g_ex[ge][0], g_ex[ge][1], g_ex[ge][2]);
    for (int l = cp->f_lo; l < lf; l++) fi += ncart(l);
    fi += cart_index(g_ex[gf][0], g_ex[gf][1], g_ex[gf][2]);
    for (int l = cp->f_lo; l <= cp->f_hi; l++) nf += ncart(l);
    return ei * nf + fi;
}

static void build_class(class_plan *cp, int la, int lb, int lc, int ld, int deriv)
{
    memset(cp, 0, sizeof(*cp));
    cp->la = la; cp->lb = lb; cp->lc = lc; cp->ld = ld;
    cp->bra_j = (la >= lb);
    cp->ket_l = (lc >= ld);
    /* derivative classes raise one AM by one on any centre: widen the ranges */
    const int dl = deriv ? 1 : 0;
    if (cp->bra_j) { cp->e_lo = la > dl ? la - dl : 0; cp->e_hi = la + lb + dl; }
    else           { cp->e_lo = lb > dl ? lb - dl : 0; cp->e_hi = la + lb + dl; }
    if (cp->ket_l) { cp->f_lo = lc > dl ? lc - dl : 0; cp->f_hi = lc + ld + dl; }
    else           { cp->f_lo = ld > dl ? ld - dl : 0; cp->f_hi = lc + ld + dl; }

    int qmax[4] = {0, 0, 0, 0};
    if (cp->bra_j) qmax[0] = cp->e_hi; else qmax[1] = cp->e_hi;
    if (cp->ket_l) qmax[2] = cp->f_hi; else qmax[3] = cp->f_hi;
    vrr_order(qmax, cp->vrr.order);
    cp->vrr.L = la + lb + lc + ld + dl;
    hm_init(&cp->vrr.map, 1024);

    int ne = 0, nf = 0;
    for (int l = cp->e_lo; l <= cp->e_hi; l++) ne += ncart(l);
    for (int l = cp->f_lo; l <= cp->f_hi; l++) nf += ncart(l);
    cp->ntop = ne * nf;
    cp->top_slot = (int *)malloc(sizeof(int) * (size_t)cp->ntop);
    cp->top_e = (int *)malloc(sizeof(int) * (size_t)cp->ntop);
    cp->top_f = (int *)malloc(sizeof(int) * (size_t)cp->ntop);
    int t = 0;
    for (int le = cp->e_lo; le <= cp->e_hi; le++)
        for (int ie = 0; ie < ncart(le); ie++)
            for (int lf = cp->f_lo; lf <= cp->f_hi; lf++)
                for (int jf = 0; jf < ncart(lf); jf++) {
                    const int ge = le * (le + 1) * (le + 2) / 6 + ie;
                    const int gf = lf * (lf + 1) * (lf + 2) / 6 + jf;
                    int q[4] = {0, 0, 0, 0};
                    if (cp->bra_j) q[0] = ge; else q[1] = ge;
                    if (cp->ket_l) q[2] = gf; else q[3] = gf;
                    cp->top_e[t] = ge; cp->top_f[t] = gf;
                    t++;
                    (void)q;
                }
    /* reorder to the top_index ordering and create the VRR nodes */
    int *slot = cp->top_slot;
    for (int i = 0; i < cp->ntop; i++) {
        const int ge = cp->top_e[i], gf = cp->top_f[i];
        int q[4] = {0, 0, 0, 0};
        if (cp->bra_j) q[0] = ge; else q[1] = ge;
        if (cp->ket_l) q[2] = gf; else q[3] = gf;
        slot[top_index(cp, ge, gf)] = vrr_need(&cp->vrr, q, 0);
    }
}

/* plan cache */
#define NPLAN (OR_MAXL + 1)
static class_plan *g_plan[2][8][8][8][8];
static pthread_mutex_t g_plan_lock = PTHREAD_MUTEX_INITIALIZER;

static const class_plan *get_plan(int la, int lb, int lc, int ld, int deriv)
{
    pthread_mutex_lock(&g_plan_lock);
    class_plan *cp = g_plan[deriv][la][lb][lc][ld];
    if (!cp) {
        cp = (class_plan *)malloc(sizeof(class_plan));
        build_class(cp, la, lb, lc, ld, deriv);
        g_plan[deriv][la][lb][lc][ld] = cp;
    }
    pthread_mutex_unlock(&g_plan_lock);
    return cp;
}

/* VRR evaluation for one primitive quartet (OSTEI_VRR_Writer.cpp:95-125) */
typedef struct {
    double PA[3], PB[3], QC[3], QD[3];
    double aop_PQ[3], aoq_PQ[3];
    double a_over_p, a_over_q;
    double one_over_2p, one_over_2q, one_over_2pq;
} prim_consts;

static void vrr_eval(const vrr_plan *pl, const prim_consts *pc, const double *F, double *v)
{
    for (int s = 0; s < pl->nnode; s++) {
        const vrr_node *nd = &pl->node[s];
        if (nd->step == ST_BASE) { v[s] = F[nd->m]; continue; }
        const int i = nd->dir;
        double t, c0, c1, c2, c3, aover;
        const double *pa, *aop;
        if (nd->step == ST_I || nd->step == ST_J) {
            pa = (nd->step == ST_I) ? pc->PA : pc->PB;
            aop = pc->aop_PQ; aover = pc->a_over_p;
            c0 = nd->ijkl[0] * pc->one_over_2p;
            c1 = nd->ijkl[1] * pc->one_over_2p;
            c2 = nd->ijkl[2] * pc->one_over_2pq;
            c3 = nd->ijkl[3] * pc->one_over_2pq;
        } else {
            pa = (nd->step == ST_K) ? pc->QC : pc->QD;
            aop = pc->aoq_PQ; aover = pc->a_over_q;
            c0 = nd->ijkl[0] * pc->one_over_2q;
            c1 = nd->ijkl[1] * pc->one_over_2q;
            c2 = nd->ijkl[2] * pc->one_over_2pq;
            c3 = nd->ijkl[3] * pc->one_over_2pq;
        }
        t = pa[i] * v[nd->src[0]];
        t = fma(aop[i], v[nd->src[1]], t);
        if (nd->src[2] >= 0) t = fma(c0, fma(aover, v[nd->src[3]], v[nd->src[2]]), t);
        if (nd->src[4] >= 0) t = fma(c1, fma(aover, v[nd->src[5]], v[nd->src[4]]), t);
        if (nd->src[6] >= 0) t = fma(c2, v[nd->src[6]], t);
        if (nd->src[7] >= 0) t = fma(c3, v[nd->src[7]], t);
        v[s] = t;
    }
}

/* -------------------------------------------------------------------------
 * HRR on the contracted top values for one quartet
 * ------------------------------------------------------------------------- */
typedef struct {
    const class_plan *cp;
    const double *top;      /* contracted tops in top_index order */
    double AB[3], CD[3];
    /* memo for bra HRR: value of (a b | f) for current f, keyed by (ga, gb) */
} hrr_ctx;

/* bra doublet value (ga, gb) for ket top component gf */
static double hrr_bra(const hrr_ctx *h, int ga, int gb, int gf, double *memo, char *have)
{
    const class_plan *cp = h->cp;
    const int key = ga * OR_NG + gb;
    if (have[key]) return memo[key];
    double v;
    if (cp->bra_j) {
        if (g_l[gb] == 0) v = h->top[top_index(cp, ga, gf)];
        else {
            const int i = mk_dir(g_ex[gb]);
            const int gb1 = gid_step(gb, i, -1), ga1 = gid_step(ga, i, +1);
            v = fma(h->AB[i], hrr_bra(h, ga, gb1, gf, memo, have), hrr_bra(h, ga1, gb1, gf, memo, have));
        }
    } else {
        if (g_l[ga] == 0) v = h->top[top_index(cp, gb, gf)];
        else {
            const int i = mk_dir(g_ex[ga]);
            const int ga1 = gid_step(ga, i, -1), gb1 = gid_step(gb, i, +1);
            v = fma(-h->AB[i], hrr_bra(h, ga1, gb, gf, memo, have), hrr_bra(h, ga1, gb1, gf, memo, have));
        }
    }
    memo[key] = v; have[key] = 1;
    return v;
}

/* ket doublet value (gc, gd) given a bra-HRR'd table for fixed final (a,b):
 * abf[gf] = (ab|f) for top ket components gf */
static double hrr_ket(const hrr_ctx *h, int gc, int gd, const double *abf, double *memo, char *have)
{
    const class_plan *cp = h->cp;
    const int key = gc * OR_NG + gd;
    if (have[key]) return memo[key];
    double v;
    if (cp->ket_l) {
        if (g_l[gd] == 0) v = abf[gc];
        else {
            const int i = mk_dir(g_ex[gd]);
            const int gd1 = gid_step(gd, i, -1), gc1 = gid_step(gc, i, +1);
            v = fma(h->CD[i], hrr_ket(h, gc, gd1, abf, memo, have), hrr_ket(h, gc1, gd1, abf, memo, have));
        }
    } else {
        if (g_l[gc] == 0) v = abf[gd];
        else {
            const int i = mk_dir(g_ex[gc]);
            const int gc1 = gid_step(gc, i, -1), gd1 = gid_step(gd, i, +1);
            v = fma(-h->CD[i], hrr_ket(h, gc1, gd, abf, memo, have), hrr_ket(h, gc1, gd1, abf, memo, have));
        }
    }
    memo[key] = v; have[key] = 1;
    return v;
}

/* final (ab|cd) for one contracted quartet: out[((a nb + b) nc + c) nd + d] */
static void hrr_all(const class_plan *cp, const double *top, const double AB[3], const double CD[3],
                    int la, int lb, int lc, int ld, double *out)
{
    hrr_ctx h;
    h.cp = cp; h.top = top;
    memcpy(h.AB, AB, sizeof(h.AB)); memcpy(h.CD, CD, sizeof(h.CD));
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int g0a = la * (la + 1) * (la + 2) / 6, g0b = lb * (lb + 1) * (lb + 2) / 6;
    const int g0c = lc * (lc + 1) * (lc + 2) / 6, g0d = ld * (ld + 1) * (ld + 2) / 6;
    double *memo = (double *)malloc(sizeof(double) * OR_NG * OR_NG);
    char *have = (char *)malloc(OR_NG * OR_NG);
    /* abf[g] holds (ab | g 0) or (ab | 0 g) for top ket Gaussians g */
    double *abf = (double *)malloc(sizeof(double) * (size_t)na * nb * OR_NG);
    /* bra HRR for every top ket component */
    for (int lf = cp->f_lo; lf <= cp->f_hi; lf++)
        for (int jf = 0; jf < ncart(lf); jf++) {
            const int gf = lf * (lf + 1) * (lf + 2) / 6 + jf;
            memset(have, 0, OR_NG * OR_NG);
            for (int a = 0; a < na; a++)
                for (int b = 0; b < nb; b++)
                    abf[(size_t)(a * nb + b) * OR_NG + gf] = hrr_bra(&h, g0a + a, g0b + b, gf, memo, have);
        }
    /* ket HRR for every final bra component */
    for (int a = 0; a < na; a++)
        for (int b = 0; b < nb; b++) {
            memset(have, 0, OR_NG * OR_NG);
            const double *row = abf + (size_t)(a * nb + b) * OR_NG;
            for (int c = 0; c < nc; c++)
                for (int d = 0; d < nd; d++)
                    out[((a * nb + b) * nc + c) * nd + d] = hrr_ket(&h, g0c + c, g0d + d, row, memo, have);
        }
    free(abf); free(memo); free(have);
}

/* -------------------------------------------------------------------------
 * Primitive loop + contraction for one contracted quartet
 * (OSTEI_Writer.cpp:470-747 with SIMD length 1)
 * ------------------------------------------------------------------------- */
static void prim_setup(const or_pair *P, int i, const or_pair *Q, int j, prim_consts *pc,
                       double *F_x_out, double *prefac_out)
{
    const double P_alpha = P->alpha[i], Q_alpha = Q->alpha[j];
    const double PQalpha_mul = P_alpha * Q_alpha;
    const double PQalpha_sum = P_alpha + Q_alpha;
    const double one_over_PQalpha_sum = 1.0 / PQalpha_sum;
    double PQ[3];
    PQ[0] = P->x[i] - Q->x[j];
    PQ[1] = P->y[i] - Q->y[j];
    PQ[2] = P->z[i] - Q->z[j];
    double R2 = PQ[0] * PQ[0];
    R2 = fma(PQ[1], PQ[1], R2);
    R2 = fma(PQ[2], PQ[2], R2);
    const double alpha = PQalpha_mul * one_over_PQalpha_sum;
    const double one_over_p = 1.0 / P_alpha;
    const double one_over_q = 1.0 / Q_alpha;
    pc->one_over_2p = 0.5 * one_over_p;
    pc->one_over_2q = 0.5 * one_over_q;
    pc->one_over_2pq = 0.5 * one_over_PQalpha_sum;
    pc->PA[0] = P->PAx[i]; pc->PA[1] = P->PAy[i]; pc->PA[2] = P->PAz[i];
    pc->PB[0] = P->PBx[i]; pc->PB[1] = P->PBy[i]; pc->PB[2] = P->PBz[i];
    pc->QC[0] = Q->PAx[j]; pc->QC[1] = Q->PAy[j]; pc->QC[2] = Q->PAz[j];
    pc->QD[0] = Q->PBx[j]; pc->QD[1] = Q->PBy[j]; pc->QD[2] = Q->PBz[j];
    pc->a_over_p = (-alpha) * one_over_p;
    for (int k = 0; k < 3; k++) pc->aop_PQ[k] = pc->a_over_p * PQ[k];
    double a_over_q = alpha * one_over_q;
    for (int k = 0; k < 3; k++) pc->aoq_PQ[k] = a_over_q * PQ[k];
    pc->a_over_q = -a_over_q;
    *F_x_out = R2 * alpha;
    double prefac = sqrt(one_over_PQalpha_sum);
    *prefac_out = (P->prefac[i] * Q->prefac[j]) * prefac;
}

/* ---- deriv1 (OSTEIDeriv1_Writer.cpp) ------------------------------------
 * d/dA_x (ab|cd) = 2a (a+1_x b|cd) - a_x (a-1_x b|cd), and likewise for B, C, D.
 * The reference skips the centre with the highest AM and forms it from
 * translational invariance (OSTEIDeriv1_Writer.cpp:497-531); the oracle
 * computes the three explicit centres and derives the fourth the same way.
 * The exponent-weighted quartets are contracted with factors alpha2/beta2
 * (2a, 2b of the pair, shell.c:512-526).                                    */

typedef struct {
    int la, lb, lc, ld;
    int skip;                   /* centre formed by invariance */
} deriv_info;

/* OSTEIDeriv1_Writer / ostei_deriv1_generator.cpp:97-107: skip the centre with
 * the highest AM (first such in A,B,C,D order) */
static int deriv_skip_centre(int la, int lb, int lc, int ld)
{
    int am[4] = {la, lb, lc, ld};
    int best = 0;
    for (int i = 1; i < 4; i++)
        if (am[i] > am[best]) best = i;
    return best;
}

/* compute primitive-contracted (e|f) tops of a class plan with an extra
 * factor per primitive quartet */
static void contract_tops(const class_plan *cp, const or_pair *P, const or_pair *Q, double tol2,
                          const double *fac_p, const double *fac_q, double *acc, double *F, double *v)
{
    const int L = cp->vrr.L;
    for (int t = 0; t < cp->ntop; t++) acc[t] = 0.0;
    for (int i = 0; i < P->n; i++) {
        for (int j = 0; j < Q->n; j++) {
            if (tol2 > 0.0 && P->screen[i] * Q->screen[j] < tol2) continue;
            prim_consts pc;
            double F_x, prefac;
            prim_setup(P, i, Q, j, &pc, &F_x, &prefac);
            boys_F_split(F, F_x, L);
            for (int n = 0; n <= L; n++) F[n] *= prefac;
            vrr_eval(&cp->vrr, &pc, F, v);
            double f = 1.0;
            if (fac_p) f *= fac_p[i];
            if (fac_q) f *= fac_q[j];
            if (fac_p || fac_q)
                for (int t = 0; t < cp->ntop; t++) acc[t] = fma(f, v[cp->top_slot[t]], acc[t]);
            else
                for (int t = 0; t < cp->ntop; t++) acc[t] += v[cp->top_slot[t]];
        }
    }
}

static int eri_quartet_internal(const or_shell *S[4], const or_pair *P, const or_pair *Q,
                                double tol2, int deriv, double *out)
{
    const int la = S[0]->am, lb = S[1]->am, lc = S[2]->am, ld = S[3]->am;
    const double AB[3] = {P->ABx, P->ABy, P->ABz};
    const double CD[3] = {Q->ABx, Q->ABy, Q->ABz};
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int nabcd = na * nb * nc * nd;

    if (!deriv) {
        const class_plan *cp = get_plan(la, lb, lc, ld, 0);
        double *acc = (double *)malloc(sizeof(double) * (size_t)cp->ntop);
        double *v = (double *)malloc(sizeof(double) * (size_t)cp->vrr.nnode);
        double F[64];
        contract_tops(cp, P, Q, tol2, NULL, NULL, acc, F, v);
        hrr_all(cp, acc, AB, CD, la, lb, lc, ld, out);
        free(acc); free(v);
        return 1;
    }

    /* deriv1: 12 doubles per Cartesian: [Ax,Ay,Az,Bx,By,Bz,Cx,Cy,Cz,Dx,Dy,Dz]
     * (OSTEI_HRR_Writer.cpp:145-146) */
    const int skip = deriv_skip_centre(la, lb, lc, ld);
    int lam[4] = {la, lb, lc, ld};
    double *dval = (double *)calloc((size_t)nabcd * 12, sizeof(double));
    double F[64];
    for (int cen = 0; cen < 4; cen++) {
        if (cen == skip) continue;
        /* (+1) class with 2*alpha factor, and (-1) class unweighted */
        int lp[4] = {la, lb, lc, ld};
        lp[cen] += 1;
        const class_plan *cpp = get_plan(lp[0], lp[1], lp[2], lp[3], 0);
        double *accp = (double *)malloc(sizeof(double) * (size_t)cpp->ntop);
        double *vp = (double *)malloc(sizeof(double) * (size_t)cpp->vrr.nnode);
        const double *fp = NULL, *fq = NULL;
        if (cen == 0) fp = P->alpha2; else if (cen == 1) fp = P->beta2;
        else if (cen == 2) fq = Q->alpha2; else fq = Q->beta2;
        contract_tops(cpp, P, Q, tol2, fp, fq, accp, F, vp);
        const int nplus = ncart(lp[0]) * ncart(lp[1]) * ncart(lp[2]) * ncart(lp[3]);
        double *plus = (double *)malloc(sizeof(double) * (size_t)nplus);
        hrr_all(cpp, accp, AB, CD, lp[0], lp[1], lp[2], lp[3], plus);
        double *minus = NULL;
        int lm[4] = {la, lb, lc, ld};
        if (lam[cen] > 0) {
            lm[cen] -= 1;
            const class_plan *cpm = get_plan(lm[0], lm[1], lm[2], lm[3], 0);
            double *accm = (double *)malloc(sizeof(double) * (size_t)cpm->ntop);
            double *vm = (double *)malloc(sizeof(double) * (size_t)cpm->vrr.nnode);
            contract_tops(cpm, P, Q, tol2, NULL, NULL, accm, F, vm);
            const int nminus = ncart(lm[0]) * ncart(lm[1]) * ncart(lm[2]) * ncart(lm[3]);
            minus = (double *)malloc(sizeof(double) * (size_t)nminus);
            hrr_all(cpm, accm, AB, CD, lm[0], lm[1], lm[2], lm[3], minus);
            free(accm); free(vm);
        }
        /* assemble (OSTEIDeriv1_Writer.cpp:334-496) */
        const int g0[4] = {la * (la + 1) * (la + 2) / 6, lb * (lb + 1) * (lb + 2) / 6,
                           lc * (lc + 1) * (lc + 2) / 6, ld * (ld + 1) * (ld + 2) / 6};
        const int nn[4] = {na, nb, nc, nd};
        for (int a = 0; a < na; a++)
            for (int b = 0; b < nb; b++)
                for (int c = 0; c < nc; c++)
                    for (int d = 0; d < nd; d++) {
                        const int comp[4] = {a, b, c, d};
                        const int idx = ((a * nb + b) * nc + c) * nd + d;
                        for (int xyz = 0; xyz < 3; xyz++) {
                            /* (+1) index */
                            int e[3] = {g_ex[g0[cen] + comp[cen]][0], g_ex[g0[cen] + comp[cen]][1],
                                        g_ex[g0[cen] + comp[cen]][2]};
                            e[xyz] += 1;
                            int cp4[4] = {a, b, c, d};
                            cp4[cen] = cart_index(e[0], e[1], e[2]);
                            int np[4] = {ncart(lp[0]), ncart(lp[1]), ncart(lp[2]), ncart(lp[3])};
                            const int ip = ((cp4[0] * np[1] + cp4[1]) * np[2] + cp4[2]) * np[3] + cp4[3];
                            double val = plus[ip];
                            e[xyz] -= 2;
                            if (e[xyz] >= 0) {
                                int cm4[4] = {a, b, c, d};
                                cm4[cen] = cart_index(e[0], e[1], e[2]);
                                int nm[4] = {ncart(lm[0]), ncart(lm[1]), ncart(lm[2]), ncart(lm[3])};
                                const int im = ((cm4[0] * nm[1] + cm4[1]) * nm[2] + cm4[2]) * nm[3] + cm4[3];
                                val = fma(-(double)(e[xyz] + 1), minus[im], val);
                            }
                            dval[(size_t)idx * 12 + cen * 3 + xyz] = val;
                        }
                        (void)nn;
                    }
        free(plus); free(minus); free(accp); free(vp);
    }
    /* translational invariance for the skipped centre (:497-531) */
    for (int idx = 0; idx < nabcd; idx++)
        for (int xyz = 0; xyz < 3; xyz++) {
            double s = 0.0;
            for (int cen = 0; cen < 4; cen++)
                if (cen != skip) s -= dval[(size_t)idx * 12 + cen * 3 + xyz];
            dval[(size_t)idx * 12 + skip * 3 + xyz] = s;
        }
    memcpy(out, dval, sizeof(double) * (size_t)nabcd * 12);
    free(dval);
    return 1;
}

/* -------------------------------------------------------------------------
 * Public oracle entry points (ctypes-friendly)
 * ------------------------------------------------------------------------- */
typedef struct {
    int n;
    or_shell *s;
} shellset;

static void load_shells(int nsh, const int *am, const int *nprim, const int *prim_off,
                        const double *xyz, const double *alpha, const double *coef, or_shell *s)
{
    for (int i = 0; i < nsh; i++) {
        s[i].am = am[i]; s[i].nprim = nprim[i];
        s[i].x = xyz[3 * i]; s[i].y = xyz[3 * i + 1]; s[i].z = xyz[3 * i + 2];
        s[i].alpha = alpha + prim_off[i];
        s[i].coef = coef + prim_off[i];
    }
}

/* Screened Schwarz value of a shell pair as stored by simint (sum/max of
 * primitive screen values is not needed here): returns screen_max. */

/*
 * oracle_eri_quartets: for each quartet (i,j,k,l) of shell indices compute
 * the contracted integrals exactly as simint_compute_eri does for a bra pair
 * (i,j) and ket pair (k,l) built by simint_create_multi_shellpair
 * (same shell index == same shell object).  Output for quartet n starts at
 * out + offsets[n].  screen_method is the one used to build the pairs;
 * screen_tol is the argument of simint_compute_eri.  deriv: 0 or 1.
 */
int oracle_eri_quartets(int nsh, const int *am, const int *nprim, const int *prim_off,
                        const double *xyz, const double *alpha, const double *coef,
                        int nq, const int *quartets, const long long *offsets,
                        int screen_method, double screen_tol, int deriv, double *out)
{
    ensure_init();
    or_shell *s = (or_shell *)malloc(sizeof(or_shell) * (size_t)nsh);
    load_shells(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    const double tol2 = screen_tol * screen_tol;
    for (int n = 0; n < nq; n++) {
        const int *qq = quartets + 4 * n;
        const or_shell *S[4] = {&s[qq[0]], &s[qq[1]], &s[qq[2]], &s[qq[3]]};
        or_pair *P = make_pair(S[0], S[1], qq[0] == qq[1], screen_method);
        or_pair *Q = make_pair(S[2], S[3], qq[2] == qq[3], screen_method);
        eri_quartet_internal(S, P, Q, (screen_method && screen_tol > 0.0) ? tol2 : 0.0, deriv,
                             out + offsets[n]);
        free(P); free(Q);
    }
    free(s);
    return 0;
}

/* Shell-pair fill for pair (i,j), same-shell if i==j (shell.c:425-582 without
 * padding).  Fields in out (each n doubles): x y z PAx PAy PAz PBx PBy PBz
 * alpha prefac alpha2 beta2 screen; returns n.  AB into ab[3].          */
int oracle_fill_pair(int nsh, const int *am, const int *nprim, const int *prim_off,
                     const double *xyz, const double *alpha, const double *coef,
                     int i, int j, int screen_method, double *out, double *ab)
{
    ensure_init();
    or_shell *s = (or_shell *)malloc(sizeof(or_shell) * (size_t)nsh);
    load_shells(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    or_pair *P = make_pair(&s[i], &s[j], i == j, screen_method);
    const int n = P->n;
    double *f[14] = {P->x, P->y, P->z, P->PAx, P->PAy, P->PAz, P->PBx, P->PBy, P->PBz,
                     P->alpha, P->prefac, P->alpha2, P->beta2, P->screen};
    for (int k = 0; k < 14; k++) memcpy(out + (size_t)k * n, f[k], sizeof(double) * (size_t)n);
    ab[0] = P->ABx; ab[1] = P->ABy; ab[2] = P->ABz;
    free(P); free(s);
    return n;
}

/* Reference-structure counts for a class (used to pin the recurrence graph
 * against the reference's worksize table, ostei_config.h.in:14-17):
 *   counts[0] = bcont (contracted top values),
 *   counts[1] = number of VRR (component, m) values actually needed,
 *   counts[2] = number of distinct component quartets in the VRR,
 *   counts[3] = prim = sum over VRR AM classes of ncart*(max m + 1)        */
int oracle_class_counts(int la, int lb, int lc, int ld, long long *counts)
{
    ensure_init();
    const class_plan *cp = get_plan(la, lb, lc, ld, 0);
    counts[0] = cp->ntop;
    counts[1] = cp->vrr.nnode;
    /* per AM class: max m */
    int maxm[11][11][11][11];
    memset(maxm, -1, sizeof(maxm));
    hmap comps;
    hm_init(&comps, 1024);
    long long ncomp = 0;
    for (int k = 0; k < cp->vrr.map.cap; k++) {
        if (cp->vrr.map.t[k].val < 0) continue;
        const uint64_t key = cp->vrr.map.t[k].key;
        const int q0 = (int)(key & 511), q1 = (int)((key >> 9) & 511), q2 = (int)((key >> 18) & 511),
                  q3 = (int)((key >> 27) & 511), m = (int)(key >> 36);
        int *mm = &maxm[g_l[q0]][g_l[q1]][g_l[q2]][g_l[q3]];
        if (m > *mm) *mm = m;
        const uint64_t ck = key & ((1ull << 36) - 1);
        if (hm_get(&comps, ck) < 0) { hm_put(&comps, ck, 1); ncomp++; }
    }
    counts[2] = ncomp;
    long long prim = 0;
    for (int a = 0; a <= OR_MAXL; a++)
        for (int b = 0; b <= OR_MAXL; b++)
            for (int c = 0; c <= OR_MAXL; c++)
                for (int d = 0; d <= OR_MAXL; d++)
                    if (maxm[a][b][c][d] >= 0)
                        prim += (long long)ncart(a) * ncart(b) * ncart(c) * ncart(d) * (maxm[a][b][c][d] + 1);
    counts[3] = prim;
    free(comps.t);
    return 0;
}

/* -------------------------------------------------------------------------
 * fast-Schwarz primitive value exposed on its own, so it can be pinned
 * against the exact (ij|ij) of the s-type primitive pair
 * ------------------------------------------------------------------------- */
double oracle_fastschwarz(int nsh, const int *am, const int *nprim, const int *prim_off,
                          const double *xyz, const double *alpha, const double *coef,
                          int sa, int sb, int i, int j)
{
    or_shell *s = (or_shell *)malloc(sizeof(or_shell) * (size_t)nsh);
    load_shells(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    const double v = fastschwarz_value(&s[sa], &s[sb], i, j);
    free(s);
    return v;
}

/* -------------------------------------------------------------------------
 * One-electron integrals (skel/simint/osoei/osoei_{terms,overlap,ke,potential}.c)
 *
 * Cartesian order and recurrence directions are those of the reference's
 * recur_lookup tables (python/gen_recurinfo.py: Psi4 order; a component is
 * built by stepping the direction of its smallest non-zero exponent,
 * rightmost on ties).
 * ------------------------------------------------------------------------- */
#define OEI_MAXL 7
typedef struct { int e[3]; int dir; int dn1[3]; int dn2[3]; } oei_cart;
static oei_cart g_oc[OEI_MAXL + 1][(OEI_MAXL + 1) * (OEI_MAXL + 2) / 2];

static int oei_find(int l, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || l < 0) return -1;
    return (l - x) * (l - x + 1) / 2 + (l - x - y);
}

static void oei_tables(void)
{
    static int done = 0;
    if (done) return;
    for (int l = 0; l <= OEI_MAXL; l++) {
        int n = 0;
        for (int x = l; x >= 0; x--)
            for (int y = l - x; y >= 0; y--, n++) {
                oei_cart *c = &g_oc[l][n];
                c->e[0] = x; c->e[1] = y; c->e[2] = l - x - y;
                c->dir = -1;
                if (l > 0) {
                    int m = 99;
                    for (int d = 0; d < 3; d++)
                        if (c->e[d] > 0 && c->e[d] < m) m = c->e[d];
                    for (int d = 2; d >= 0; d--)
                        if (c->e[d] == m) { c->dir = d; break; }
                }
                for (int d = 0; d < 3; d++) {
                    int e1[3] = {c->e[0], c->e[1], c->e[2]}, e2[3] = {c->e[0], c->e[1], c->e[2]};
                    e1[d] -= 1; e2[d] -= 2;
                    c->dn1[d] = oei_find(l - 1, e1[0], e1[1], e1[2]);
                    c->dn2[d] = oei_find(l - 2, e2[0], e2[1], e2[2]);
                }
            }
    }
    done = 1;
}

/* osoei_terms.c:7-78 -- 1D overlap tables S_ij, i < nam1, j < nam2, per direction */
static void oei_overlap_terms(double a1, const double *A, double a2, const double *B, int nam1, int nam2,
                              double *terms)
{
    const double oop = 1.0 / (a1 + a2);
    const double mu = a1 * a2 * oop;
    const double oo2p = 0.5 * oop;
    for (int d = 0; d < 3; d++) {
        const double AB = A[d] - B[d];
        const double P = (a1 * A[d] + a2 * B[d]) * oop;
        const double PA = P - A[d], PB = P - B[d];
        double *s = terms + d * nam1 * nam2;
#define S_(i, j) s[(i) * nam2 + (j)]
        S_(0, 0) = sqrt(3.14159265358979323846 * oop) * exp(-mu * AB * AB);
        for (int i = 1; i < nam1; i++) {
            S_(i, 0) = PA * S_(i - 1, 0);
            if (i > 1) S_(i, 0) += (i - 1) * oo2p * S_(i - 2, 0);
        }
        for (int j = 1; j < nam2; j++) {
            S_(0, j) = PB * S_(0, j - 1);
            if (j > 1) S_(0, j) += (j - 1) * oo2p * S_(0, j - 2);
        }
        for (int i = 1; i < nam1; i++)
            for (int j = 1; j < nam2; j++) {
                S_(i, j) = PB * S_(i, j - 1) + oo2p * i * S_(i - 1, j - 1);
                if (j > 1) S_(i, j) += oo2p * (j - 1) * S_(i, j - 2);
            }
#undef S_
    }
}

/* osoei_terms.c:81-166 -- the overlap tables (first 3 blocks) and the 1D
 * kinetic tables T_ij (blocks 3..5) */
static void oei_ke_terms(double a1, const double *A, double a2, const double *B, int nam1, int nam2,
                         double *terms)
{
    const double oop = 1.0 / (a1 + a2);
    const double mu = a1 * a2 * oop;
    const double oo2p = 0.5 * oop;
    const double a1sq = 2.0 * a1 * a1;
    for (int d = 0; d < 3; d++) {
        const double AB = A[d] - B[d];
        const double P = (a1 * A[d] + a2 * B[d]) * oop;
        const double PA = P - A[d], PB = P - B[d];
        double *s = terms + d * nam1 * nam2;
        double *t = terms + (d + 3) * nam1 * nam2;
#define S_(i, j) s[(i) * nam2 + (j)]
#define T_(i, j) t[(i) * nam2 + (j)]
        S_(0, 0) = sqrt(3.14159265358979323846 * oop) * exp(-mu * AB * AB);
        T_(0, 0) = S_(0, 0) * (a1 - a1sq * (PA * PA + oo2p));
        for (int i = 1; i < nam1; i++) {
            S_(i, 0) = PA * S_(i - 1, 0);
            if (i > 1) S_(i, 0) += (i - 1) * oo2p * S_(i - 2, 0);
            T_(i, 0) = PA * T_(i - 1, 0) + 2 * mu * S_(i, 0);
            if (i > 1) T_(i, 0) += (i - 1) * oo2p * T_(i - 2, 0) - a2 * oop * (i - 1) * S_(i - 2, 0);
        }
        for (int j = 1; j < nam2; j++) {
            S_(0, j) = PB * S_(0, j - 1);
            if (j > 1) S_(0, j) += (j - 1) * oo2p * S_(0, j - 2);
            T_(0, j) = PB * T_(0, j - 1) + 2 * mu * S_(0, j);
            if (j > 1) T_(0, j) += (j - 1) * oo2p * T_(0, j - 2) - a1 * oop * (j - 1) * S_(0, j - 2);
        }
        for (int i = 1; i < nam1; i++)
            for (int j = 1; j < nam2; j++) {
                S_(i, j) = PB * S_(i, j - 1) + oo2p * i * S_(i - 1, j - 1);
                if (j > 1) S_(i, j) += oo2p * (j - 1) * S_(i, j - 2);
                T_(i, j) = PB * T_(i, j - 1) + oo2p * i * T_(i - 1, j - 1) + 2 * mu * S_(i, j);
                if (j > 1) T_(i, j) += oo2p * (j - 1) * T_(i, j - 2) - a1 * oop * (j - 1) * S_(i, j - 2);
            }
#undef S_
#undef T_
    }
}

/* osoei_overlap.c:10-66 (kind 0) and osoei_ke.c:10-67 (kind 1): one block
 * of ncart(a) x ncart(b), primitive sums of products of the 1D tables */
static void oei_sk(int kind, const or_shell *A, const or_shell *B, double *out)
{
    const int nam1 = A->am + 1, nam2 = B->am + 1, n12 = nam1 * nam2;
    const int na = ncart(A->am), nb = ncart(B->am);
    double w[6 * (OEI_MAXL + 2) * (OEI_MAXL + 2)];
    const double xa[3] = {A->x, A->y, A->z}, xb[3] = {B->x, B->y, B->z};
    for (int k = 0; k < na * nb; k++) out[k] = 0.0;
    for (int i = 0; i < A->nprim; i++)
        for (int j = 0; j < B->nprim; j++) {
            const double pf = A->coef[i] * B->coef[j];
            if (kind == 0) oei_overlap_terms(A->alpha[i], xa, B->alpha[j], xb, nam1, nam2, w);
            else oei_ke_terms(A->alpha[i], xa, B->alpha[j], xb, nam1, nam2, w);
            int o = 0;
            for (int n = 0; n < na; n++)
                for (int m = 0; m < nb; m++, o++) {
                    const int *e1 = g_oc[A->am][n].e, *e2 = g_oc[B->am][m].e;
                    const int ix = e1[0] * nam2 + e2[0], iy = e1[1] * nam2 + e2[1], iz = e1[2] * nam2 + e2[2];
                    double v;
                    if (kind == 0) v = w[ix] * w[n12 + iy] * w[2 * n12 + iz];
                    else
                        v = w[3 * n12 + ix] * w[n12 + iy] * w[2 * n12 + iz] +
                            w[ix] * w[4 * n12 + iy] * w[2 * n12 + iz] +
                            w[ix] * w[n12 + iy] * w[5 * n12 + iz];
                    out[o] += pf * v;
                }
        }
}

/* osoei_potential.c:11-218 -- nuclear attraction, -sum_C Z_C <a|1/r_C|b>:
 * per point charge and primitive pair, Theta^(m)_{i,0} up the first centre
 * from 2 pi / p e^{-mu AB^2} F_m(p PC^2), then Theta^(m)_{i,j} up the second
 * centre, each component along its recurrence direction. */
static void oei_potential(int ncenter, const double *Z, const double *cxyz, const or_shell *A,
                          const or_shell *B, double *out)
{
    const int am1 = A->am, am2 = B->am, L = am1 + am2;
    const int na = ncart(am1), nb = ncart(am2);
    /* th[i][j]: (M(i,j)+1) levels of ncart(i)*ncart(j) values, M(i,j) = L - i - j */
    double *th[OEI_MAXL + 1][OEI_MAXL + 1];
    size_t tot = 0;
    for (int i = 0; i <= am1; i++)
        for (int j = 0; j <= am2; j++) tot += (size_t)(L - i - j + 1) * ncart(i) * ncart(j);
    double *thbuf = (double *)malloc(sizeof(double) * tot);
    tot = 0;
    for (int i = 0; i <= am1; i++)
        for (int j = 0; j <= am2; j++) {
            th[i][j] = thbuf + tot;
            tot += (size_t)(L - i - j + 1) * ncart(i) * ncart(j);
        }
    const double xa[3] = {A->x, A->y, A->z}, xb[3] = {B->x, B->y, B->z};
    double AB2 = 0.0;
    for (int d = 0; d < 3; d++) AB2 += (xa[d] - xb[d]) * (xa[d] - xb[d]);
    for (int k = 0; k < na * nb; k++) out[k] = 0.0;
    for (int c = 0; c < ncenter; c++)
        for (int ia = 0; ia < A->nprim; ia++)
            for (int ib = 0; ib < B->nprim; ib++) {
                const double a1 = A->alpha[ia], a2 = B->alpha[ib];
                const double p = a1 + a2, oop = 1.0 / p, mu = a1 * a2 * oop, oo2p = 0.5 * oop;
                double PA[3], PB[3], PC[3], PC2 = 0.0;
                for (int d = 0; d < 3; d++) {
                    const double P = (a1 * xa[d] + a2 * xb[d]) * oop;
                    PA[d] = P - xa[d];
                    PB[d] = P - xb[d];
                    PC[d] = P - cxyz[3 * c + d];
                    PC2 += PC[d] * PC[d];
                }
                double F[2 * OEI_MAXL + 2];
                oracle_boys(PC2 * p, L, F);
                const double pre = 2.0 * 3.14159265358979323846 * oop * exp(-mu * AB2);
                for (int m = 0; m <= L; m++) th[0][0][m] = F[m] * pre;
                for (int i = 0; i <= am1; i++) {
                    const int ni = ncart(i), ni1 = i > 0 ? ncart(i - 1) : 0, ni2 = i > 1 ? ncart(i - 2) : 0;
                    if (i > 0) {
                        for (int m = 0; m <= L - i; m++)
                            for (int n = 0; n < ni; n++) {
                                const oei_cart *cc = &g_oc[i][n];
                                const int d = cc->dir;
                                double v = PA[d] * th[i - 1][0][m * ni1 + cc->dn1[d]] -
                                           PC[d] * th[i - 1][0][(m + 1) * ni1 + cc->dn1[d]];
                                if (cc->e[d] > 1)
                                    v += oo2p * (cc->e[d] - 1) *
                                         (th[i - 2][0][m * ni2 + cc->dn2[d]] - th[i - 2][0][(m + 1) * ni2 + cc->dn2[d]]);
                                th[i][0][m * ni + n] = v;
                            }
                    }
                    for (int j = 1; j <= am2; j++) {
                        const int nj = ncart(j), nj1 = ncart(j - 1), nj2 = j > 1 ? ncart(j - 2) : 0;
                        for (int m = 0; m <= L - i - j; m++)
                            for (int n = 0; n < ni; n++)
                                for (int o = 0; o < nj; o++) {
                                    const oei_cart *cb = &g_oc[j][o];
                                    const int d = cb->dir;
                                    const int ei = g_oc[i][n].e[d];
                                    const int i1 = cb->dn1[d];
                                    double v = PB[d] * th[i][j - 1][(m * ni + n) * nj1 + i1] -
                                               PC[d] * th[i][j - 1][((m + 1) * ni + n) * nj1 + i1];
                                    if (ei > 0) {
                                        const int a1i = g_oc[i][n].dn1[d];
                                        v += oo2p * ei * th[i - 1][j - 1][(m * ni1 + a1i) * nj1 + i1] -
                                             oo2p * ei * th[i - 1][j - 1][((m + 1) * ni1 + a1i) * nj1 + i1];
                                    }
                                    if (cb->e[d] > 1) {
                                        const int i2 = cb->dn2[d];
                                        v += oo2p * (cb->e[d] - 1) *
                                             (th[i][j - 2][(m * ni + n) * nj2 + i2] -
                                              th[i][j - 2][((m + 1) * ni + n) * nj2 + i2]);
                                    }
                                    th[i][j][(m * ni + n) * nj + o] = v;
                                }
                    }
                }
                const double f = Z[c] * A->coef[ia] * B->coef[ib];
                for (int k = 0; k < na * nb; k++) out[k] -= th[am1][am2][k] * f;
            }
    free(thbuf);
}

/* kind 0 overlap, 1 kinetic energy, 2 nuclear attraction of the point
 * charges (Z, cxyz); block n = ncart(a) x ncart(b) values at out + offsets[n] */
int oracle_osoei(int kind, int nsh, const int *am, const int *nprim, const int *prim_off,
                 const double *xyz, const double *alpha, const double *coef, int npair, const int *pairs,
                 const long long *offsets, int ncenter, const double *Z, const double *cxyz, double *out)
{
    ensure_init();
    oei_tables();
    or_shell *s = (or_shell *)malloc(sizeof(or_shell) * (size_t)nsh);
    load_shells(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    for (int n = 0; n < npair; n++) {
        const or_shell *A = &s[pairs[2 * n]], *B = &s[pairs[2 * n + 1]];
        if (A->am > OEI_MAXL || B->am > OEI_MAXL) { free(s); return -1; }
        if (kind == 2) oei_potential(ncenter, Z, cxyz, A, B, out + offsets[n]);
        else oei_sk(kind, A, B, out + offsets[n]);
    }
    free(s);
    return 0;
}

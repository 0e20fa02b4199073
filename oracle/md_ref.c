/*
 * md_ref.c -- independent brute-force ERI reference in long double.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.c).  This pins the Obara-Saika
 * restatement in oracle.c with a different algorithm: McMurchie-Davidson
 * Hermite expansion (Helgaker, Jorgensen, Olsen, "Molecular
 * Electronic-Structure Theory", eqs. 9.5.6-9.5.7 for E, 9.9.18-9.9.20 for R,
 * 9.9.33 for the ERI), its own Boys function (positive series below x = 50,
 * asymptotic above), no triangle/same-shell shortcuts, every primitive
 * combination summed, every value in long double.
 *
 * Integral definition (the one simint computes, skel/simint/shell/shell.c:493-498
 * and generator/ostei/OSTEI_Writer.cpp:695-702): unnormalised Cartesian
 * Gaussians x^i y^j z^k exp(-a r^2) contracted with the caller's coefficients
 * (after simint_normalize_shells), output in Psi4 Cartesian order,
 * out[((a nb + b) nc + c) nd + d].
 *
 * First derivatives are formed by Richardson-extrapolated central finite
 * differences of the long-double integral with respect to the centre
 * coordinates, so they share nothing with the analytic (a+1)/(a-1) formula.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MD_MAXL 8
#define MD_MAXN (4 * MD_MAXL + 2)

typedef long double ld;

/* oracle.c: simint's split Boys scheme (reference values, double) */
void oracle_boys(double x, int n, double *F);
static int g_refboys = 0;   /* 1: use the reference scheme's F_m (test mode) */

static ld md_boys(int n, ld x)
{
    if (x > 50.0L) {
        /* Gamma(n+1/2) / (2 x^(n+1/2)); e^-x < 2e-22 */
        ld g = sqrtl(3.14159265358979323846264338327950288L);
        for (int k = 0; k < n; k++) g *= (k + 0.5L);
        return g / (2.0L * powl(x, n + 0.5L));
    }
    ld term = 1.0L / (2 * n + 1), sum = term;
    for (int k = 1; k < 4000; k++) {
        term *= 2.0L * x / (2 * n + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-24L) break;
    }
    return expl(-x) * sum;
}

/* Hermite expansion coefficients E^{ij}_t (t = 0..i+j) for one direction */
static void hermite_E(int imax, int jmax, ld p, ld XPA, ld XPB, ld E[MD_MAXL + 2][MD_MAXL + 2][2 * MD_MAXL + 5])
{
    memset(E, 0, sizeof(ld) * (MD_MAXL + 2) * (MD_MAXL + 2) * (2 * MD_MAXL + 5));
    const ld o2p = 1.0L / (2.0L * p);
    E[0][0][0] = 1.0L;
    for (int i = 0; i <= imax; i++)
        for (int j = 0; j <= jmax; j++) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; t++) {
                ld v;
                if (i > 0) {
                    v = XPA * E[i - 1][j][t];
                    if (t > 0) v += o2p * E[i - 1][j][t - 1];
                    if (t + 1 <= i - 1 + j) v += (t + 1) * E[i - 1][j][t + 1];
                } else {
                    v = XPB * E[i][j - 1][t];
                    if (t > 0) v += o2p * E[i][j - 1][t - 1];
                    if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
                }
                E[i][j][t] = v;
            }
        }
}

static inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static void cart_list(int l, int (*e)[3])
{
    int n = 0;
    for (int ix = l; ix >= 0; ix--)
        for (int iy = l - ix; iy >= 0; iy--) { e[n][0] = ix; e[n][1] = iy; e[n][2] = l - ix - iy; n++; }
}

/* position of x^ix y^iy z^iz in the order cart_list writes */
static inline int cart_index(int ix, int iy, int iz)
{
    const int l = ix + iy + iz;
    return (l - ix) * (l - ix + 1) / 2 + (l - ix - iy);
}

typedef struct {
    int am, nprim;
    ld c[3];
    const double *alpha, *coef;
} md_shell;

/* contracted (ab|cd) for one quartet, accumulated into out (long double).
 * wcen >= 0: each primitive quartet is weighted by 2 * (exponent of the
 * primitive on centre wcen), the factor of the analytic derivative
 * d/dA_x G_a = 2 alpha G_{a+1_x} - a_x G_{a-1_x}. */
static int g_absmode = 0;   /* 1: accumulate |primitive contribution| */

static void md_quartet_w(const md_shell *S[4], int wcen, ld *out)
{
    const int la = S[0]->am, lb = S[1]->am, lc = S[2]->am, ld_ = S[3]->am;
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld_);
    int ea[MD_MAXL * 8][3], eb[MD_MAXL * 8][3], ec[MD_MAXL * 8][3], ed[MD_MAXL * 8][3];
    cart_list(la, ea); cart_list(lb, eb); cart_list(lc, ec); cart_list(ld_, ed);
    const int N = la + lb + lc + ld_;
    const int nall = na * nb * nc * nd;
    for (int k = 0; k < nall; k++) out[k] = 0.0L;

    static __thread ld Eab[3][MD_MAXL + 2][MD_MAXL + 2][2 * MD_MAXL + 5];
    static __thread ld Ecd[3][MD_MAXL + 2][MD_MAXL + 2][2 * MD_MAXL + 5];
    const int D = N + 1;
    ld *Rn = (ld *)malloc(sizeof(ld) * D * D * D);
    ld *Rn1 = (ld *)malloc(sizeof(ld) * D * D * D);
#define RI(t, u, v) (((t) * D + (u)) * D + (v))

    const ld pi = 3.14159265358979323846264338327950288L;
    for (int i = 0; i < S[0]->nprim; i++)
        for (int j = 0; j < S[1]->nprim; j++)
            for (int k = 0; k < S[2]->nprim; k++)
                for (int l = 0; l < S[3]->nprim; l++) {
                    const ld a = S[0]->alpha[i], b = S[1]->alpha[j], c = S[2]->alpha[k], d = S[3]->alpha[l];
                    const ld p = a + b, q = c + d;
                    ld P[3], Q[3], AB2 = 0, CD2 = 0, RPQ2 = 0;
                    for (int x = 0; x < 3; x++) {
                        P[x] = (a * S[0]->c[x] + b * S[1]->c[x]) / p;
                        Q[x] = (c * S[2]->c[x] + d * S[3]->c[x]) / q;
                        AB2 += (S[0]->c[x] - S[1]->c[x]) * (S[0]->c[x] - S[1]->c[x]);
                        CD2 += (S[2]->c[x] - S[3]->c[x]) * (S[2]->c[x] - S[3]->c[x]);
                        RPQ2 += (P[x] - Q[x]) * (P[x] - Q[x]);
                    }
                    const ld alpha = p * q / (p + q);
                    ld cc = (ld)S[0]->coef[i] * S[1]->coef[j] * S[2]->coef[k] * S[3]->coef[l];
                    if (wcen >= 0) {
                        const ld ex[4] = {a, b, c, d};
                        cc *= 2.0L * ex[wcen];
                    }
                    const ld pref = cc * 2.0L * powl(pi, 2.5L) / (p * q * sqrtl(p + q))
                                    * expl(-a * b / p * AB2) * expl(-c * d / q * CD2);
                    for (int x = 0; x < 3; x++) {
                        hermite_E(la, lb, p, P[x] - S[0]->c[x], P[x] - S[1]->c[x], Eab[x]);
                        hermite_E(lc, ld_, q, Q[x] - S[2]->c[x], Q[x] - S[3]->c[x], Ecd[x]);
                    }
                    /* R^n_{tuv}, n from N down to 0 */
                    const ld T = alpha * RPQ2;
                    double Fref[64];
                    if (g_refboys) oracle_boys((double)T, N, Fref);
                    for (int n = N; n >= 0; n--) {
                        const int lim = N - n;
                        ld f = g_refboys ? (ld)Fref[n] : md_boys(n, T);
                        ld m2a = 1.0L;
                        for (int z = 0; z < n; z++) m2a *= -2.0L * alpha;
                        for (int t = 0; t <= lim; t++)
                            for (int u = 0; u + t <= lim; u++)
                                for (int v = 0; v + u + t <= lim; v++) {
                                    ld r;
                                    if (t > 0) {
                                        r = (P[0] - Q[0]) * Rn1[RI(t - 1, u, v)];
                                        if (t > 1) r += (t - 1) * Rn1[RI(t - 2, u, v)];
                                    } else if (u > 0) {
                                        r = (P[1] - Q[1]) * Rn1[RI(t, u - 1, v)];
                                        if (u > 1) r += (u - 1) * Rn1[RI(t, u - 2, v)];
                                    } else if (v > 0) {
                                        r = (P[2] - Q[2]) * Rn1[RI(t, u, v - 1)];
                                        if (v > 1) r += (v - 1) * Rn1[RI(t, u, v - 2)];
                                    } else {
                                        r = m2a * f;
                                    }
                                    Rn[RI(t, u, v)] = r;
                                }
                        ld *tmp = Rn; Rn = Rn1; Rn1 = tmp;
                    }
                    /* now Rn1 holds level 0 */
                    const ld *R = Rn1;
                    int idx = 0;
                    for (int ia = 0; ia < na; ia++)
                        for (int ib = 0; ib < nb; ib++)
                            for (int ic = 0; ic < nc; ic++)
                                for (int id = 0; id < nd; id++, idx++) {
                                    ld s = 0.0L;
                                    const int tx = ea[ia][0] + eb[ib][0], ty = ea[ia][1] + eb[ib][1], tz = ea[ia][2] + eb[ib][2];
                                    const int ux = ec[ic][0] + ed[id][0], uy = ec[ic][1] + ed[id][1], uz = ec[ic][2] + ed[id][2];
                                    for (int t = 0; t <= tx; t++) {
                                        const ld e1 = Eab[0][ea[ia][0]][eb[ib][0]][t];
                                        if (e1 == 0.0L) continue;
                                        for (int u = 0; u <= ty; u++) {
                                            const ld e2 = e1 * Eab[1][ea[ia][1]][eb[ib][1]][u];
                                            if (e2 == 0.0L) continue;
                                            for (int v = 0; v <= tz; v++) {
                                                const ld e3 = e2 * Eab[2][ea[ia][2]][eb[ib][2]][v];
                                                if (e3 == 0.0L) continue;
                                                ld s2 = 0.0L;
                                                for (int tt = 0; tt <= ux; tt++)
                                                    for (int uu = 0; uu <= uy; uu++)
                                                        for (int vv = 0; vv <= uz; vv++) {
                                                            const ld f = Ecd[0][ec[ic][0]][ed[id][0]][tt]
                                                                       * Ecd[1][ec[ic][1]][ed[id][1]][uu]
                                                                       * Ecd[2][ec[ic][2]][ed[id][2]][vv];
                                                            const ld sg = ((tt + uu + vv) & 1) ? -1.0L : 1.0L;
                                                            s2 += sg * f * R[RI(t + tt, u + uu, v + vv)];
                                                        }
                                                s += e3 * s2;
                                            }
                                        }
                                    }
                                    out[idx] += g_absmode ? fabsl(pref * s) : pref * s;
                                }
                }
    free(Rn); free(Rn1);
#undef RI
}

static void md_quartet(const md_shell *S[4], ld *out) { md_quartet_w(S, -1, out); }

static void load(int nsh, const int *am, const int *nprim, const int *prim_off, const double *xyz,
                 const double *alpha, const double *coef, md_shell *s)
{
    for (int i = 0; i < nsh; i++) {
        s[i].am = am[i]; s[i].nprim = nprim[i];
        for (int x = 0; x < 3; x++) s[i].c[x] = xyz[3 * i + x];
        s[i].alpha = alpha + prim_off[i];
        s[i].coef = coef + prim_off[i];
    }
}

/* Select the Boys values md_eri uses: 0 = exact (series / asymptotic in long
 * double), 1 = simint's split scheme (oracle_boys), which isolates the
 * recurrence algebra from the Boys-function approximation. */
void md_set_reference_boys(int on) { g_refboys = on; }

/* contracted integrals for a list of shell quartets (same layout and offsets
 * as oracle_eri_quartets with deriv = 0) */
int md_eri_quartets(int nsh, const int *am, const int *nprim, const int *prim_off,
                    const double *xyz, const double *alpha, const double *coef,
                    int nq, const int *quartets, const long long *offsets, double *out)
{
    md_shell *s = (md_shell *)malloc(sizeof(md_shell) * (size_t)nsh);
    load(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    for (int n = 0; n < nq; n++) {
        const int *q = quartets + 4 * n;
        const md_shell *S[4] = {&s[q[0]], &s[q[1]], &s[q[2]], &s[q[3]]};
        if (S[0]->am > MD_MAXL || S[1]->am > MD_MAXL || S[2]->am > MD_MAXL || S[3]->am > MD_MAXL) {
            free(s);
            return -1;
        }
        const int nall = ncart(S[0]->am) * ncart(S[1]->am) * ncart(S[2]->am) * ncart(S[3]->am);
        ld *buf = (ld *)malloc(sizeof(ld) * (size_t)nall);
        md_quartet(S, buf);
        for (int k = 0; k < nall; k++) out[offsets[n] + k] = (double)buf[k];
        free(buf);
    }
    free(s);
    return 0;
}

/* first derivatives by Richardson-extrapolated central differences:
 * out[idx*12 + centre*3 + xyz], same layout as oracle_eri_quartets deriv=1 */
int md_eri_deriv1_quartets(int nsh, const int *am, const int *nprim, const int *prim_off,
                           const double *xyz, const double *alpha, const double *coef,
                           int nq, const int *quartets, const long long *offsets, double h,
                           double *out)
{
    md_shell *s = (md_shell *)malloc(sizeof(md_shell) * (size_t)nsh);
    load(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    for (int n = 0; n < nq; n++) {
        const int *q = quartets + 4 * n;
        md_shell sh[4];
        for (int c = 0; c < 4; c++) sh[c] = s[q[c]];
        const int nall = ncart(sh[0].am) * ncart(sh[1].am) * ncart(sh[2].am) * ncart(sh[3].am);
        ld *p1 = (ld *)malloc(sizeof(ld) * nall), *m1 = (ld *)malloc(sizeof(ld) * nall);
        ld *p2 = (ld *)malloc(sizeof(ld) * nall), *m2 = (ld *)malloc(sizeof(ld) * nall);
        for (int cen = 0; cen < 4; cen++)
            for (int x = 0; x < 3; x++) {
                /* the same shell may sit on several centres: move every
                 * occurrence of this centre's shell index only once per
                 * derivative, i.e. move the centre, not the shell object */
                const ld c0 = sh[cen].c[x];
                const md_shell *S[4] = {&sh[0], &sh[1], &sh[2], &sh[3]};
                sh[cen].c[x] = c0 + h;       md_quartet(S, p1);
                sh[cen].c[x] = c0 - h;       md_quartet(S, m1);
                sh[cen].c[x] = c0 + 0.5L * h; md_quartet(S, p2);
                sh[cen].c[x] = c0 - 0.5L * h; md_quartet(S, m2);
                sh[cen].c[x] = c0;
                for (int k = 0; k < nall; k++) {
                    const ld d1 = (p1[k] - m1[k]) / (2.0L * h);
                    const ld d2 = (p2[k] - m2[k]) / (1.0L * h);
                    out[offsets[n] + (long long)k * 12 + cen * 3 + x] = (double)((4.0L * d2 - d1) / 3.0L);
                }
            }
        free(p1); free(m1); free(p2); free(m2);
    }
    free(s);
    return 0;
}

/* first derivatives from the analytic formula, every term in long double:
 * d/dX (..a..) = sum over primitives [2 alpha (..a+1_x..) - a_x (..a-1_x..)]
 * for all four centres (no translational-invariance shortcut).  Same layout
 * as md_eri_deriv1_quartets.  This is the exact value of what simint's
 * first-derivative kernels compute (for a pair built from one shell, the
 * symmetrised weights alpha_i + alpha_j of shell.c:512-526 give the same sum,
 * the primitive pair being symmetric on one centre). */
int md_eri_deriv1_terms(int nsh, const int *am, const int *nprim, const int *prim_off,
                        const double *xyz, const double *alpha, const double *coef,
                        int nq, const int *quartets, const long long *offsets, int scale, double *out);

int md_eri_deriv1_analytic_quartets(int nsh, const int *am, const int *nprim, const int *prim_off,
                                    const double *xyz, const double *alpha, const double *coef,
                                    int nq, const int *quartets, const long long *offsets, double *out)
{
    return md_eri_deriv1_terms(nsh, am, nprim, prim_off, xyz, alpha, coef, nq, quartets, offsets, 0, out);
}

/* scale: 1 = instead of the derivative, the sum over primitives of the
 * magnitudes of both terms, |2 alpha (a+1)| + a_x |(a-1)|, i.e. the size of
 * what the derivative is formed from; FP64 evaluations of the formula can
 * differ by a few ulps of this, not of the (possibly much smaller) result. */
int md_eri_deriv1_terms(int nsh, const int *am, const int *nprim, const int *prim_off,
                        const double *xyz, const double *alpha, const double *coef,
                        int nq, const int *quartets, const long long *offsets, int scale, double *out)
{
    g_absmode = scale;
    md_shell *s = (md_shell *)malloc(sizeof(md_shell) * (size_t)nsh);
    load(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    for (int n = 0; n < nq; n++) {
        const int *q = quartets + 4 * n;
        md_shell sh[4];
        for (int c = 0; c < 4; c++) sh[c] = s[q[c]];
        if (sh[0].am + 1 > MD_MAXL || sh[1].am + 1 > MD_MAXL || sh[2].am + 1 > MD_MAXL || sh[3].am + 1 > MD_MAXL) {
            free(s);
            return -1;
        }
        const int n4[4] = {ncart(sh[0].am), ncart(sh[1].am), ncart(sh[2].am), ncart(sh[3].am)};
        const int nall = n4[0] * n4[1] * n4[2] * n4[3];
        int e4[4][MD_MAXL * 8][3];
        for (int c = 0; c < 4; c++) cart_list(sh[c].am, e4[c]);
        for (int cen = 0; cen < 4; cen++) {
            md_shell sp[4] = {sh[0], sh[1], sh[2], sh[3]}, sm[4] = {sh[0], sh[1], sh[2], sh[3]};
            sp[cen].am += 1;
            sm[cen].am -= 1;
            const md_shell *Sp[4] = {&sp[0], &sp[1], &sp[2], &sp[3]};
            const md_shell *Sm[4] = {&sm[0], &sm[1], &sm[2], &sm[3]};
            int np4[4] = {n4[0], n4[1], n4[2], n4[3]}, nm4[4] = {n4[0], n4[1], n4[2], n4[3]};
            np4[cen] = ncart(sp[cen].am);
            ld *plus = (ld *)malloc(sizeof(ld) * (size_t)np4[0] * np4[1] * np4[2] * np4[3]);
            md_quartet_w(Sp, cen, plus);
            ld *minus = NULL;
            if (sh[cen].am > 0) {
                nm4[cen] = ncart(sm[cen].am);
                minus = (ld *)malloc(sizeof(ld) * (size_t)nm4[0] * nm4[1] * nm4[2] * nm4[3]);
                md_quartet(Sm, minus);
            }
            for (int k = 0; k < nall; k++) {
                int comp[4], r = k;
                for (int c = 3; c >= 0; c--) { comp[c] = r % n4[c]; r /= n4[c]; }
                for (int x = 0; x < 3; x++) {
                    int e[3] = {e4[cen][comp[cen]][0], e4[cen][comp[cen]][1], e4[cen][comp[cen]][2]};
                    int cp[4] = {comp[0], comp[1], comp[2], comp[3]};
                    e[x] += 1;
                    cp[cen] = cart_index(e[0], e[1], e[2]);
                    ld v = plus[((cp[0] * np4[1] + cp[1]) * np4[2] + cp[2]) * np4[3] + cp[3]];
                    e[x] -= 2;
                    if (e[x] >= 0) {
                        int cm[4] = {comp[0], comp[1], comp[2], comp[3]};
                        cm[cen] = cart_index(e[0], e[1], e[2]);
                        const ld mv = (ld)(e[x] + 1) * minus[((cm[0] * nm4[1] + cm[1]) * nm4[2] + cm[2]) * nm4[3] + cm[3]];
                        v = scale ? v + mv : v - mv;
                    }
                    out[offsets[n] + (long long)k * 12 + cen * 3 + x] = (double)v;
                }
            }
            free(plus);
            free(minus);
        }
    }
    g_absmode = 0;
    free(s);
    return 0;
}

double md_boys_value(int n, double x) { return (double)md_boys(n, (ld)x); }

/* -------------------------------------------------------------------------
 * One-electron integrals in long double, from the Hermite expansion
 * (Helgaker, Jorgensen & Olsen, ch. 9): overlap S = E^{ij}_0 (pi/p)^{1/2} per
 * direction; kinetic energy from the 1D identity
 *   T_ij = -2 b^2 S_{i,j+2} + b (2j+1) S_ij - j(j-1)/2 S_{i,j-2};
 * nuclear attraction -Z 2 pi/p e^{-mu AB^2} sum_tuv E_t E_u E_v R_tuv(p, P-C).
 * Independent of the Obara-Saika tables the oracle and the kernels use.
 * kind 0 overlap, 1 kinetic, 2 potential; blocks as oracle_osoei.
 * ------------------------------------------------------------------------- */
int md_oei_pairs(int kind, int nsh, const int *am, const int *nprim, const int *prim_off,
                 const double *xyz, const double *alpha, const double *coef, int npair, const int *pairs,
                 const long long *offsets, int ncenter, const double *Z, const double *cxyz, double *out)
{
    md_shell *s = (md_shell *)malloc(sizeof(md_shell) * (size_t)nsh);
    load(nsh, am, nprim, prim_off, xyz, alpha, coef, s);
    const ld pi = 3.14159265358979323846264338327950288L;
    static __thread ld E[3][MD_MAXL + 2][MD_MAXL + 2][2 * MD_MAXL + 5];
    for (int n = 0; n < npair; n++) {
        const md_shell *A = &s[pairs[2 * n]], *B = &s[pairs[2 * n + 1]];
        if (A->am + 2 > MD_MAXL || B->am + 2 > MD_MAXL) { free(s); return -1; }
        const int na = ncart(A->am), nb = ncart(B->am);
        int ea[64][3], eb[64][3];
        cart_list(A->am, ea);
        cart_list(B->am, eb);
        ld acc[64 * 64];
        for (int k = 0; k < na * nb; k++) acc[k] = 0.0L;
        for (int i = 0; i < A->nprim; i++)
            for (int j = 0; j < B->nprim; j++) {
                const ld a = A->alpha[i], b = B->alpha[j], p = a + b, mu = a * b / p;
                const ld cc = (ld)A->coef[i] * B->coef[j];
                ld P[3], K[3];
                for (int x = 0; x < 3; x++) {
                    P[x] = (a * A->c[x] + b * B->c[x]) / p;
                    const ld X = A->c[x] - B->c[x];
                    K[x] = expl(-mu * X * X);
                    hermite_E(A->am, B->am + 2, p, P[x] - A->c[x], P[x] - B->c[x], E[x]);
                }
                if (kind < 2) {
                    const ld sq = sqrtl(pi / p);
                    for (int u = 0; u < na; u++)
                        for (int v = 0; v < nb; v++) {
                            ld S1[3], T1[3];
                            for (int x = 0; x < 3; x++) {
                                const int ii = ea[u][x], jj = eb[v][x];
                                S1[x] = E[x][ii][jj][0] * sq * K[x];
                                if (kind == 1) {
                                    ld t = -2.0L * b * b * E[x][ii][jj + 2][0] + b * (2 * jj + 1) * E[x][ii][jj][0];
                                    if (jj >= 2) t -= 0.5L * jj * (jj - 1) * E[x][ii][jj - 2][0];
                                    T1[x] = t * sq * K[x];
                                }
                            }
                            ld val = kind == 0 ? S1[0] * S1[1] * S1[2]
                                               : T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2];
                            acc[u * nb + v] += cc * val;
                        }
                    continue;
                }
                const int N = A->am + B->am, D = N + 1;
                ld *Rn = (ld *)malloc(sizeof(ld) * D * D * D), *Rn1 = (ld *)malloc(sizeof(ld) * D * D * D);
#define RI(t, u, v) (((t) * D + (u)) * D + (v))
                for (int c = 0; c < ncenter; c++) {
                    ld PC[3], R2 = 0.0L;
                    for (int x = 0; x < 3; x++) { PC[x] = P[x] - (ld)cxyz[3 * c + x]; R2 += PC[x] * PC[x]; }
                    const ld T = p * R2;
                    double Fref[64];
                    if (g_refboys) oracle_boys((double)T, N, Fref);
                    for (int nn = N; nn >= 0; nn--) {
                        const int lim = N - nn;
                        ld m2p = 1.0L;
                        for (int z = 0; z < nn; z++) m2p *= -2.0L * p;
                        for (int t = 0; t <= lim; t++)
                            for (int u = 0; u + t <= lim; u++)
                                for (int v = 0; v + u + t <= lim; v++) {
                                    ld r;
                                    if (t > 0) {
                                        r = PC[0] * Rn1[RI(t - 1, u, v)];
                                        if (t > 1) r += (t - 1) * Rn1[RI(t - 2, u, v)];
                                    } else if (u > 0) {
                                        r = PC[1] * Rn1[RI(t, u - 1, v)];
                                        if (u > 1) r += (u - 1) * Rn1[RI(t, u - 2, v)];
                                    } else if (v > 0) {
                                        r = PC[2] * Rn1[RI(t, u, v - 1)];
                                        if (v > 1) r += (v - 1) * Rn1[RI(t, u, v - 2)];
                                    } else {
                                        r = m2p * (g_refboys ? (ld)Fref[nn] : md_boys(nn, T));
                                    }
                                    Rn[RI(t, u, v)] = r;
                                }
                        ld *tmp = Rn; Rn = Rn1; Rn1 = tmp;
                    }
                    const ld pre = -(ld)Z[c] * cc * 2.0L * pi / p * K[0] * K[1] * K[2];
                    for (int u = 0; u < na; u++)
                        for (int v = 0; v < nb; v++) {
                            ld sum = 0.0L;
                            for (int t = 0; t <= ea[u][0] + eb[v][0]; t++)
                                for (int uu = 0; uu <= ea[u][1] + eb[v][1]; uu++)
                                    for (int vv = 0; vv <= ea[u][2] + eb[v][2]; vv++)
                                        sum += E[0][ea[u][0]][eb[v][0]][t] * E[1][ea[u][1]][eb[v][1]][uu] *
                                               E[2][ea[u][2]][eb[v][2]][vv] * Rn1[RI(t, uu, vv)];
                            acc[u * nb + v] += pre * sum;
                        }
                }
#undef RI
                free(Rn);
                free(Rn1);
            }
        for (int k = 0; k < na * nb; k++) out[offsets[n] + k] = (double)acc[k];
    }
    free(s);
    return 0;
}

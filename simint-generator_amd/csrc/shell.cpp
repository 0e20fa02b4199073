// Shells and shell pairs: the host side of the simint API.
//
// Semantics follow skel/simint/shell/shell.c (allocation 131-185 and
// 325-382, normalisation 248-278, Gaussian-product fill 425-582, primitive
// sort 37-128, concatenation 605-873), with the layout fixed at
// SIMINT_SIMD_LEN == 1 (no padding primitives) and SIMINT_OSTEI_MAXDER > 0
// (alpha2/beta2 always present).  Two deliberate differences, both on paths
// where the reference misbehaves:
//   * simint_copy_shell copies the 2*nprim doubles of the source, not
//     dest->memsize bytes (which over-reads a smaller source buffer);
//   * simint_cat_multi_shellpair copies each input pair's own primitives
//     (the reference re-reads the first pair's primitives for every pair of
//     a multi-pair input; results agree whenever inputs hold one pair each).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "simint/constants.h"
#include "simint/shell/shell.h"
#include "simint/shell/shell_screen.h"

namespace {

// pi^(3/2) (2l-1)!! / 2^l, the l-dependent part of the shell norm
double norm_factor(int l)
{
    long double v = 5.568327996831707845284817982118835702014L;  // pi^(3/2)
    for (int k = 1; k <= 2 * l - 1; k += 2) v *= k;
    for (int k = 0; k < l; k++) v *= 0.5L;
    return (double)v;
}

char *bytes(void *p) { return static_cast<char *>(p); }

void pair_swap(simint_multi_shellpair *P, int i, int j)
{
    double *f[14] = {P->x, P->y, P->z, P->PA_x, P->PA_y, P->PA_z, P->PB_x, P->PB_y, P->PB_z,
                     P->alpha, P->prefac, P->screen, P->alpha2, P->beta2};
    for (double *a : f) {
        const double t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
}

// descending quicksort on the screen value (shell.c:37-71)
void pair_quicksort(simint_multi_shellpair *P, int left, int right)
{
    int i = left, j = right;
    const double mid = P->screen[(i + j) / 2];
    while (i <= j) {
        while (P->screen[i] > mid) i++;
        while (P->screen[j] < mid) j--;
        if (i <= j) {
            pair_swap(P, i, j);
            i++;
            j--;
        }
    }
    if (i < right) pair_quicksort(P, i, right);
    if (j > left) pair_quicksort(P, left, j);
}

void sort_pairs(simint_multi_shellpair *P)
{
    if (P->nprim == 1) return;
    int idx = 0;
    for (int ab = 0; ab < P->nshell12; ab++) {
        if (P->nprim12[ab] > 1) pair_quicksort(P, idx, idx + P->nprim12[ab] - 1);
        idx += P->nprim12[ab];
    }
}

// one allocation, arrays in the reference's order (shell.c:131-185)
void allocate_base(int npair, int nprim, simint_multi_shellpair *P, int screen_method)
{
    const size_t dprim = (size_t)nprim * sizeof(double);
    const size_t ipair = (size_t)npair * sizeof(int);
    const size_t dpair = (size_t)npair * sizeof(double);
    int narr = 11 + (screen_method ? 1 : 0) + 2;
    const size_t memsize = dprim * narr + dpair * 3 + ipair;
    if (P->memsize < memsize) {
        simint_free_multi_shellpair(P);
        P->ptr = SIMINT_ALLOC(memsize > 0 ? memsize : 1);
        if (!P->ptr) {
            fprintf(stderr, "simint: cannot allocate %zu bytes for a multi_shellpair\n", memsize);
            abort();
        }
        P->memsize = memsize;
    }
    int k = 0;
    P->x = (double *)(bytes(P->ptr) + dprim * (k++));
    P->y = (double *)(bytes(P->ptr) + dprim * (k++));
    P->z = (double *)(bytes(P->ptr) + dprim * (k++));
    P->PA_x = (double *)(bytes(P->ptr) + dprim * (k++));
    P->PA_y = (double *)(bytes(P->ptr) + dprim * (k++));
    P->PA_z = (double *)(bytes(P->ptr) + dprim * (k++));
    P->PB_x = (double *)(bytes(P->ptr) + dprim * (k++));
    P->PB_y = (double *)(bytes(P->ptr) + dprim * (k++));
    P->PB_z = (double *)(bytes(P->ptr) + dprim * (k++));
    P->alpha = (double *)(bytes(P->ptr) + dprim * (k++));
    P->prefac = (double *)(bytes(P->ptr) + dprim * (k++));
    P->screen = screen_method ? (double *)(bytes(P->ptr) + dprim * (k++)) : nullptr;
    P->alpha2 = (double *)(bytes(P->ptr) + dprim * (k++));
    P->beta2 = (double *)(bytes(P->ptr) + dprim * (k++));
    P->AB_x = (double *)(bytes(P->ptr) + dprim * k);
    P->AB_y = (double *)(bytes(P->ptr) + dprim * k + dpair);
    P->AB_z = (double *)(bytes(P->ptr) + dprim * k + 2 * dpair);
    P->nprim12 = (int *)(bytes(P->ptr) + dprim * k + 3 * dpair);
}

int pair_nprim(const simint_shell *A, const simint_shell *B)
{
    return compare_shell(A, B) ? A->nprim * (A->nprim + 1) / 2 : A->nprim * B->nprim;
}

// interleave A x B into the pair list the *2 entry points take
simint_shell *make_pair_list(int na, const simint_shell *A, int nb, const simint_shell *B)
{
    simint_shell *AB = (simint_shell *)malloc(sizeof(simint_shell) * 2 * (size_t)(na * nb > 0 ? na * nb : 1));
    if (!AB) {
        fprintf(stderr, "simint: cannot allocate the %d x %d shell pair list\n", na, nb);
        abort();
    }
    int k = 0;
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            AB[k] = A[i];
            AB[k + 1] = B[j];
            k += 2;
        }
    return AB;
}

}  // namespace

extern "C" {

void simint_initialize_shell(simint_shell *G)
{
    G->ptr = nullptr;
    G->memsize = 0;
}

void simint_initialize_shells(int n, simint_shell *G)
{
    for (int i = 0; i < n; i++) simint_initialize_shell(G + i);
}

void simint_allocate_shell(int nprim, simint_shell *G)
{
    const size_t memsize = 2 * (size_t)nprim * sizeof(double);
    if (G->memsize < memsize) {
        simint_free_shell(G);
        G->ptr = malloc(memsize > 0 ? memsize : 1);
        G->memsize = memsize;
    }
    G->alpha = (double *)G->ptr;
    G->coef = (double *)(bytes(G->ptr) + (size_t)nprim * sizeof(double));
}

void simint_free_shell(simint_shell *G)
{
    if (G->ptr != nullptr) free(G->ptr);
    simint_initialize_shell(G);
}

void simint_free_shells(int n, simint_shell *G)
{
    for (int i = 0; i < n; i++) simint_free_shell(G + i);
}

void simint_copy_shell(const simint_shell *src, simint_shell *dest)
{
    simint_allocate_shell(src->nprim, dest);
    dest->nprim = src->nprim;
    dest->am = src->am;
    dest->x = src->x;
    dest->y = src->y;
    dest->z = src->z;
    memcpy(dest->alpha, src->alpha, sizeof(double) * (size_t)src->nprim);
    memcpy(dest->coef, src->coef, sizeof(double) * (size_t)src->nprim);
}

void simint_normalize_shells(int n, simint_shell *G)
{
    for (int i = 0; i < n; ++i) {
        const double m = (double)G[i].am + 1.5;
        const double m2 = 0.5 * m;
        double sum = 0.0;
        for (int j = 0; j < G[i].nprim; j++) {
            const double a1 = G[i].alpha[j], c1 = G[i].coef[j];
            for (int k = 0; k < G[i].nprim; k++) {
                const double a2 = G[i].alpha[k], c2 = G[i].coef[k];
                sum += (c1 * c2 * pow(a1 * a2, m2)) / (pow(a1 + a2, m));
            }
        }
        const double norm = 1.0 / sqrt(sum * norm_factor(G[i].am));
        for (int j = 0; j < G[i].nprim; ++j) G[i].coef[j] *= norm * pow(G[i].alpha[j], m2);
    }
}

void simint_create_shell(int nprim, int am, double x, double y, double z, const double *alpha,
                         const double *coef, simint_shell *G)
{
    simint_allocate_shell(nprim, G);
    G->am = am;
    G->nprim = nprim;
    G->x = x;
    G->y = y;
    G->z = z;
    memcpy(G->alpha, alpha, (size_t)nprim * sizeof(double));
    memcpy(G->coef, coef, (size_t)nprim * sizeof(double));
}

void simint_create_zero_shell(simint_shell *G)
{
    const double alpha[1] = {0.0};
    const double coef[1] = {1.0};
    simint_create_shell(1, 0, 0.0, 0.0, 0.0, alpha, coef, G);
}

void simint_initialize_multi_shellpair(simint_multi_shellpair *P)
{
    P->ptr = nullptr;
    P->memsize = 0;
}

void simint_initialize_multi_shellpairs(int n, simint_multi_shellpair *P)
{
    for (int i = 0; i < n; i++) simint_initialize_multi_shellpair(P + i);
}

void simint_free_multi_shellpair(simint_multi_shellpair *P)
{
    if (P->ptr != nullptr) SIMINT_FREE(P->ptr);
    simint_initialize_multi_shellpair(P);
}

void simint_free_multi_shellpairs(int n, simint_multi_shellpair *P)
{
    for (int i = 0; i < n; i++) simint_free_multi_shellpair(P + i);
}

void simint_allocate_multi_shellpair(int na, const simint_shell *A, int nb, const simint_shell *B,
                                     simint_multi_shellpair *P, int screen_method)
{
    simint_shell *AB = make_pair_list(na, A, nb, B);
    simint_allocate_multi_shellpair2(na * nb, AB, P, screen_method);
    free(AB);
}

void simint_allocate_multi_shellpair2(int npair, const simint_shell *AB, simint_multi_shellpair *P,
                                      int screen_method)
{
    int nprim = 0;
    for (int i = 0; i < npair; i++) nprim += pair_nprim(&AB[2 * i], &AB[2 * i + 1]);
    allocate_base(npair, nprim, P, screen_method);
}

void simint_fill_multi_shellpair(int na, const simint_shell *A, int nb, const simint_shell *B,
                                 simint_multi_shellpair *P, int screen_method)
{
    simint_shell *AB = make_pair_list(na, A, nb, B);
    simint_fill_multi_shellpair2(na * nb, AB, P, screen_method);
    free(AB);
}

void simint_fill_multi_shellpair2(int npair, const simint_shell *AB, simint_multi_shellpair *P,
                                  int screen_method)
{
    P->nshell12 = npair;
    P->nshell12_clip = npair;
    P->nprim = 0;
    // The reference starts the maximum at 1e99 (shell.c:434), so pairs made
    // here never trigger the shell-level early exit; kept for identical
    // behaviour.  simint_cat_* compute a real maximum.
    P->screen_max = 1.e99;

    int idx = 0;
    for (int ab = 0; ab < npair; ab++) {
        const simint_shell *A = &AB[2 * ab];
        const simint_shell *B = &AB[2 * ab + 1];
        if (screen_method && P->screen) {
            const double m = simint_primscreen(A, B, P->screen + idx, screen_method);
            P->screen_max = (m > P->screen_max ? m : P->screen_max);
        }
        const int same = compare_shell(A, B);
        P->am1 = A->am;
        P->am2 = B->am;
        const double Xab_x = A->x - B->x;
        const double Xab_y = A->y - B->y;
        const double Xab_z = A->z - B->z;
        const double Xab = Xab_x * Xab_x + Xab_y * Xab_y + Xab_z * Xab_z;
        for (int i = 0; i < A->nprim; ++i) {
            const double ai = A->alpha[i];
            const double AxAa = A->x * ai, AyAa = A->y * ai, AzAa = A->z * ai;
            const int jend = same ? i + 1 : B->nprim;
            for (int j = 0; j < jend; ++j) {
                const double aj = B->alpha[j];
                const double ab_sum = ai + aj;
                const double ab_mul = ai * aj;
                double pf = A->coef[i] * B->coef[j] * exp(-Xab * ab_mul / ab_sum) * SQRT_TWO_PI_52 / ab_sum;
                if (same && i != j) pf *= 2.0;
                P->prefac[idx] = pf;
                P->x[idx] = (AxAa + aj * B->x) / ab_sum;
                P->y[idx] = (AyAa + aj * B->y) / ab_sum;
                P->z[idx] = (AzAa + aj * B->z) / ab_sum;
                P->PA_x[idx] = P->x[idx] - A->x;
                P->PA_y[idx] = P->y[idx] - A->y;
                P->PA_z[idx] = P->z[idx] - A->z;
                P->PB_x[idx] = P->x[idx] - B->x;
                P->PB_y[idx] = P->y[idx] - B->y;
                P->PB_z[idx] = P->z[idx] - B->z;
                P->alpha[idx] = ab_sum;
                if (same && i != j) {
                    // the factor 2 already sits in prefac (shell.c:513-520)
                    P->alpha2[idx] = ai + aj;
                    P->beta2[idx] = ai + aj;
                } else {
                    P->alpha2[idx] = 2.0 * ai;
                    P->beta2[idx] = 2.0 * aj;
                }
                ++idx;
            }
        }
        P->AB_x[ab] = Xab_x;
        P->AB_y[ab] = Xab_y;
        P->AB_z[ab] = Xab_z;
        const int np = pair_nprim(A, B);
        P->nprim12[ab] = np;
        P->nprim += np;
    }
    if (screen_method && P->screen) sort_pairs(P);
}

void simint_create_multi_shellpair(int na, const simint_shell *A, int nb, const simint_shell *B,
                                   simint_multi_shellpair *P, int screen_method)
{
    simint_allocate_multi_shellpair(na, A, nb, B, P, screen_method);
    simint_fill_multi_shellpair(na, A, nb, B, P, screen_method);
}

void simint_create_multi_shellpair2(int npair, const simint_shell *AB, simint_multi_shellpair *P,
                                    int screen_method)
{
    simint_allocate_multi_shellpair2(npair, AB, P, screen_method);
    simint_fill_multi_shellpair2(npair, AB, P, screen_method);
}

void simint_cat_multi_shellpair(int nmpair, const simint_multi_shellpair **Pin, simint_multi_shellpair *Pout,
                                int screen_method)
{
    if (nmpair <= 0) return;
    int nshell12 = 0, nprim = 0;
    for (int i = 0; i < nmpair; i++) {
        nshell12 += Pin[i]->nshell12;
        for (int j = 0; j < Pin[i]->nshell12; j++) nprim += Pin[i]->nprim12[j];
    }
    allocate_base(nshell12, nprim, Pout, screen_method);
    Pout->am1 = Pin[0]->am1;
    Pout->am2 = Pin[0]->am2;
    Pout->screen_max = 0.0;
    Pout->nprim = 0;
    int idx = 0, sasb = 0;
    for (int i = 0; i < nmpair; i++) {
        const simint_multi_shellpair *S = Pin[i];
        int src = 0;
        for (int j = 0; j < S->nshell12; j++) {
            for (int p = 0; p < S->nprim12[j]; p++, src++, idx++) {
                Pout->x[idx] = S->x[src];
                Pout->y[idx] = S->y[src];
                Pout->z[idx] = S->z[src];
                Pout->PA_x[idx] = S->PA_x[src];
                Pout->PA_y[idx] = S->PA_y[src];
                Pout->PA_z[idx] = S->PA_z[src];
                Pout->PB_x[idx] = S->PB_x[src];
                Pout->PB_y[idx] = S->PB_y[src];
                Pout->PB_z[idx] = S->PB_z[src];
                Pout->alpha[idx] = S->alpha[src];
                Pout->prefac[idx] = S->prefac[src];
                Pout->alpha2[idx] = S->alpha2[src];
                Pout->beta2[idx] = S->beta2[src];
                if (screen_method) Pout->screen[idx] = S->screen ? S->screen[src] : 0.0;
            }
            Pout->AB_x[sasb] = S->AB_x[j];
            Pout->AB_y[sasb] = S->AB_y[j];
            Pout->AB_z[sasb] = S->AB_z[j];
            Pout->nprim12[sasb] = S->nprim12[j];
            sasb++;
        }
        if (S->screen_max > Pout->screen_max) Pout->screen_max = S->screen_max;
        Pout->nprim += S->nprim;
    }
    Pout->nshell12 = nshell12;
    Pout->nshell12_clip = nshell12;
}

void simint_cat_shellpairs(int nmpair, const simint_multi_shellpair **Pin, simint_multi_shellpair *Pout,
                           int screen_method)
{
    if (nmpair <= 0) return;
    // every input holds one shell pair (shell.c:733-873)
    int nprim = 0;
    for (int i = 0; i < nmpair; i++) nprim += Pin[i]->nprim12[0];
    allocate_base(nmpair, nprim, Pout, screen_method);
    Pout->am1 = Pin[0]->am1;
    Pout->am2 = Pin[0]->am2;
    Pout->screen_max = 0.0;
    Pout->nprim = 0;
    int idx = 0;
    for (int i = 0; i < nmpair; i++) {
        const simint_multi_shellpair *S = Pin[i];
        const size_t n = (size_t)S->nprim12[0];
        const size_t bytes_n = sizeof(double) * n;
        memcpy(&Pout->x[idx], S->x, bytes_n);
        memcpy(&Pout->y[idx], S->y, bytes_n);
        memcpy(&Pout->z[idx], S->z, bytes_n);
        memcpy(&Pout->PA_x[idx], S->PA_x, bytes_n);
        memcpy(&Pout->PA_y[idx], S->PA_y, bytes_n);
        memcpy(&Pout->PA_z[idx], S->PA_z, bytes_n);
        memcpy(&Pout->PB_x[idx], S->PB_x, bytes_n);
        memcpy(&Pout->PB_y[idx], S->PB_y, bytes_n);
        memcpy(&Pout->PB_z[idx], S->PB_z, bytes_n);
        memcpy(&Pout->alpha[idx], S->alpha, bytes_n);
        memcpy(&Pout->prefac[idx], S->prefac, bytes_n);
        memcpy(&Pout->alpha2[idx], S->alpha2, bytes_n);
        memcpy(&Pout->beta2[idx], S->beta2, bytes_n);
        if (screen_method) {
            for (size_t p = 0; p < n; p++) Pout->screen[idx + p] = S->screen ? S->screen[p] : 0.0;
        }
        idx += (int)n;
        if (S->screen_max > Pout->screen_max) Pout->screen_max = S->screen_max;
        Pout->AB_x[i] = S->AB_x[0];
        Pout->AB_y[i] = S->AB_y[0];
        Pout->AB_z[i] = S->AB_z[0];
        Pout->nprim12[i] = S->nprim12[0];
        Pout->nprim += S->nprim;
    }
    Pout->nshell12 = nmpair;
    Pout->nshell12_clip = nmpair;
}

}  // extern "C"

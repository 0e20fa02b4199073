/* Gaussian shells and shell pairs (skel/simint/shell/shell.h:18-375).
 *
 * The two structures are laid out exactly as the reference's with
 * SIMINT_OSTEI_MAXDER > 0 (so alpha2/beta2 are always present), and every
 * function keeps its signature and meaning.  Pairs are stored unpadded
 * (SIMINT_SIMD_LEN == 1). */
#pragma once

#include <stddef.h>
#include "simint/ostei/ostei_config.h"

#define SIMINT_SCREEN_NONE         0
#define SIMINT_SCREEN_SCHWARZ      1
#define SIMINT_SCREEN_FASTSCHWARZ  2

#ifdef __cplusplus
extern "C" {
#endif

/* One contracted Cartesian Gaussian shell. */
struct simint_shell
{
    int am;          /* angular momentum */
    int nprim;       /* number of primitives */

    double x;        /* centre (bohr) */
    double y;
    double z;

    double * alpha;  /* exponents, nprim */
    double * coef;   /* contraction coefficients, nprim */

    size_t memsize;  /* bytes owned through ptr */
    void * ptr;      /* the one allocation behind alpha and coef */
};

/* Gaussian-product data for a set of shell pairs sharing (am1, am2), one
 * entry per primitive pair, structure of arrays. */
struct simint_multi_shellpair
{
    int am1;            /* AM of the first shell of every pair */
    int am2;            /* AM of the second shell */
    int nprim;          /* total primitive pairs stored */

    int nshell12;       /* number of shell pairs */
    int nshell12_clip;  /* number of pairs to compute (<= nshell12) */
    int * nprim12;      /* primitive pairs per shell pair */

    double * AB_x;      /* A - B per shell pair */
    double * AB_y;
    double * AB_z;

    double * x;         /* P = (a A + b B)/(a + b) per primitive pair */
    double * y;
    double * z;
    double * PA_x;      /* P - A */
    double * PA_y;
    double * PA_z;
    double * PB_x;      /* P - B */
    double * PB_y;
    double * PB_z;

    double * alpha;     /* a + b */

    double * alpha2;    /* 2a (a + b for an off-diagonal same-shell pair) */
    double * beta2;     /* 2b (a + b for an off-diagonal same-shell pair) */

    double * prefac;    /* cA cB exp(-ab/(a+b) AB^2) sqrt(2 pi^(5/2)) / (a+b) */
    double * screen;    /* primitive (ab|ab) bound, NULL without screening */
    double screen_max;  /* largest screen value (see simint_create_multi_shellpair) */

    size_t memsize;     /* bytes owned through ptr */
    void * ptr;         /* the one allocation behind all arrays */
};

/* Two shells are "the same shell" when they share storage and size; a
 * same-shell pair stores only the primitive triangle (shell.h:89-94). */
static inline
int compare_shell(struct simint_shell const * A,
                  struct simint_shell const * B)
{
    return A->nprim == B->nprim && A->ptr == B->ptr;
}

void simint_initialize_shell(struct simint_shell * G);
void simint_initialize_shells(int n, struct simint_shell * G);
void simint_allocate_shell(int nprim, struct simint_shell * G);
void simint_free_shell(struct simint_shell * G);
void simint_free_shells(int n, struct simint_shell * G);
void simint_copy_shell(struct simint_shell const * src,
                       struct simint_shell * dest);
void simint_normalize_shells(int n, struct simint_shell * G);
void simint_create_shell(int nprim, int am, double x, double y, double z,
                         double const * alpha,
                         double const * coef,
                         struct simint_shell * G);
void simint_create_zero_shell(struct simint_shell * G);

void simint_initialize_multi_shellpair(struct simint_multi_shellpair * P);
void simint_initialize_multi_shellpairs(int n, struct simint_multi_shellpair * P);
void simint_allocate_multi_shellpair(int na, struct simint_shell const * A,
                                     int nb, struct simint_shell const * B,
                                     struct simint_multi_shellpair * P,
                                     int screen_method);
void simint_allocate_multi_shellpair2(int npair, struct simint_shell const * AB,
                                      struct simint_multi_shellpair * P,
                                      int screen_method);
void simint_free_multi_shellpair(struct simint_multi_shellpair * P);
void simint_free_multi_shellpairs(int n, struct simint_multi_shellpair * P);
void simint_fill_multi_shellpair(int na, struct simint_shell const * A,
                                 int nb, struct simint_shell const * B,
                                 struct simint_multi_shellpair * P,
                                 int screen_method);
void simint_fill_multi_shellpair2(int npair, struct simint_shell const * AB,
                                  struct simint_multi_shellpair * P,
                                  int screen_method);
void simint_create_multi_shellpair(int na, struct simint_shell const * A,
                                   int nb, struct simint_shell const * B,
                                   struct simint_multi_shellpair * P,
                                   int screen_method);
void simint_create_multi_shellpair2(int npair,
                                    struct simint_shell const * AB,
                                    struct simint_multi_shellpair * P,
                                    int screen_method);
void simint_cat_multi_shellpair(int nmpair,
                                struct simint_multi_shellpair const ** Pin,
                                struct simint_multi_shellpair * Pout,
                                int screen_method);
void simint_cat_shellpairs(int nmpair,
                           struct simint_multi_shellpair const ** Pin,
                           struct simint_multi_shellpair * Pout,
                           int screen_method);

#ifdef __cplusplus
}
#endif

/* A plain C client of the drop-in API, in the style of the reference's
 * skel/examples/example1.c:16-173: it includes only <simint/simint.h> (and
 * the class-dispatch header for the simint_osteifunc type), links against
 * libsimint.so, and uses the calls exactly as a simint program does.
 *
 *   client <jobs.txt> <out.bin>
 *
 * jobs.txt:  nshell, then one line per shell "am nprim x y z alpha... coef...";
 *            njob, then one line per job "kind i j k l":
 *              kind 0  simint_compute_eri       on P = (i j), Q = (k l)
 *              kind 1  simint_compute_eri_deriv on P = (i j), Q = (k l)
 *              kind 2  simint_compute_overlap   of shells i, j
 *              kind 3  simint_compute_ke        of shells i, j
 *              kind 4  simint_compute_potential of shells i, j, unit charges at
 *                      every shell centre
 * out.bin:   per job: int32 return code, int64 value count, the values. */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>

#include <simint/simint.h>
#include <simint/ostei/ostei.h>

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* the dispatch-table function type must stay declarable by client code */
static simint_osteifunc unused_table[1];

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: client jobs.txt out.bin\n");
        return 2;
    }
    (void)unused_table;
    FILE *in = fopen(argv[1], "r");
    FILE *out = fopen(argv[2], "wb");
    if (!in || !out) return 2;
    simint_init();

    int nshell = 0;
    if (fscanf(in, "%d", &nshell) != 1 || nshell <= 0) return 2;
    struct simint_shell *sh = malloc(sizeof(struct simint_shell) * nshell);
    double *cx = malloc(sizeof(double) * nshell), *cy = malloc(sizeof(double) * nshell);
    double *cz = malloc(sizeof(double) * nshell), *cq = malloc(sizeof(double) * nshell);
    int maxam = 0;
    for (int s = 0; s < nshell; s++) {
        int am, nprim;
        double x, y, z;
        if (fscanf(in, "%d %d %lf %lf %lf", &am, &nprim, &x, &y, &z) != 5) return 2;
        double *al = malloc(sizeof(double) * nprim), *co = malloc(sizeof(double) * nprim);
        for (int p = 0; p < nprim; p++)
            if (fscanf(in, "%lf", &al[p]) != 1) return 2;
        for (int p = 0; p < nprim; p++)
            if (fscanf(in, "%lf", &co[p]) != 1) return 2;
        simint_initialize_shell(&sh[s]);
        simint_create_shell(nprim, am, x, y, z, al, co, &sh[s]);
        free(al);
        free(co);
        cx[s] = x;
        cy[s] = y;
        cz[s] = z;
        cq[s] = 1.0;
        if (am > maxam) maxam = am;
    }
    simint_normalize_shells(nshell, sh);

    /* workspace sized as the reference's callers size it (test_eri.cpp:71) */
    double *work = SIMINT_ALLOC(simint_eri_workmem(1, maxam));
    int njob = 0;
    if (fscanf(in, "%d", &njob) != 1) return 2;
    for (int j = 0; j < njob; j++) {
        int kind, a, b, c, d;
        if (fscanf(in, "%d %d %d %d %d", &kind, &a, &b, &c, &d) != 5) return 2;
        int64_t n = 0;
        int rc = 0;
        double *res = NULL;
        if (kind <= 1) {
            struct simint_multi_shellpair P, Q;
            simint_initialize_multi_shellpair(&P);
            simint_initialize_multi_shellpair(&Q);
            simint_create_multi_shellpair(1, &sh[a], 1, &sh[b], &P, SIMINT_SCREEN_NONE);
            simint_create_multi_shellpair(1, &sh[c], 1, &sh[d], &Q, SIMINT_SCREEN_NONE);
            n = (int64_t)ncart(sh[a].am) * ncart(sh[b].am) * ncart(sh[c].am) * ncart(sh[d].am) * (kind ? 12 : 1);
            res = SIMINT_ALLOC(sizeof(double) * n);
            rc = kind ? simint_compute_eri_deriv(1, &P, &Q, 0.0, work, res)
                      : simint_compute_eri(&P, &Q, 0.0, work, res);
            simint_free_multi_shellpair(&P);
            simint_free_multi_shellpair(&Q);
        } else {
            n = (int64_t)ncart(sh[a].am) * ncart(sh[b].am);
            res = SIMINT_ALLOC(sizeof(double) * n);
            if (kind == 2) rc = simint_compute_overlap(&sh[a], &sh[b], res);
            else if (kind == 3) rc = simint_compute_ke(&sh[a], &sh[b], res);
            else rc = simint_compute_potential(nshell, cq, cx, cy, cz, &sh[a], &sh[b], res);
        }
        int32_t rc32 = rc;
        fwrite(&rc32, sizeof(rc32), 1, out);
        fwrite(&n, sizeof(n), 1, out);
        fwrite(res, sizeof(double), (size_t)n, out);
        SIMINT_FREE(res);
    }
    SIMINT_FREE(work);
    for (int s = 0; s < nshell; s++) simint_free_shell(&sh[s]);
    free(sh);
    free(cx);
    free(cy);
    free(cz);
    free(cq);
    fclose(in);
    fclose(out);
    simint_finalize();
    return 0;
}

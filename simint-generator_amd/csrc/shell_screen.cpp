// Shell-pair screening values (skel/simint/shell/shell_screen.c:9-270).
//
// Fast Schwarz is the closed-form (00|00) bound of each primitive pair.
// Schwarz computes the (ij|ij) integrals of every uncontracted primitive
// pair on the GPU -- the diagonal quartets of all primitive pairs of one
// shell pair in a single launch, not every pair against every pair -- and
// keeps the largest diagonal element (not square-rooted; the compute entry
// points compare against screen_tol^2).
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "simint/constants.h"
#include "simint/simint_eri.h"
#include "simint/shell/shell_screen.h"

namespace {

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// (ij|ij) of the s-type primitive pair (shell_screen.c:226-270)
inline double fast_value(const simint_shell *A, const simint_shell *B, int i, int j)
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

}  // namespace

namespace simint_rt {
// frontend.hip: the largest |(bb|bb)| diagonal element of every pair of P,
// from the (bb|bb) quartets alone (npair quartets, not npair^2)
int host_pair_bounds(const simint_multi_shellpair *P, double *out);
}

extern "C" {

double simint_shellscreen(const simint_shell *A, const simint_shell *B, int screen_method)
{
    if (screen_method == SIMINT_SCREEN_FASTSCHWARZ) return simint_shellscreen_fastschwarz(A, B);
    return simint_shellscreen_schwarz(A, B);
}

double simint_primscreen(const simint_shell *A, const simint_shell *B, double *out, int screen_method)
{
    if (screen_method == SIMINT_SCREEN_FASTSCHWARZ) return simint_primscreen_fastschwarz(A, B, out);
    return simint_primscreen_schwarz(A, B, out);
}

double simint_shellscreen_schwarz(const simint_shell *A, const simint_shell *B)
{
    simint_multi_shellpair P;
    simint_initialize_multi_shellpair(&P);
    simint_create_multi_shellpair(1, A, 1, B, &P, SIMINT_SCREEN_NONE);
    const int n12 = ncart(A->am) * ncart(B->am);
    std::vector<double> ints((size_t)n12 * n12);
    const int rc = simint_compute_eri(&P, &P, 0.0, nullptr, ints.data());
    simint_free_multi_shellpair(&P);
    if (rc < 0) return NAN;
    double mx = 0.0;
    for (double v : ints) mx = fabs(v) > mx ? fabs(v) : mx;
    return mx;
}

double simint_shellscreen_fastschwarz(const simint_shell *A, const simint_shell *B)
{
    const int same = compare_shell(A, B);
    double val = 0.0;
    for (int i = 0; i < A->nprim; i++) {
        const int jend = same ? i + 1 : B->nprim;
        for (int j = 0; j < jend; j++) val += fast_value(A, B, i, j);
    }
    return val;
}

double simint_primscreen_schwarz(const simint_shell *A, const simint_shell *B, double *out)
{
    // uncontract: one single-primitive shell pair per primitive pair
    const int same = compare_shell(A, B);
    std::vector<simint_shell> pairs;
    for (int i = 0; i < A->nprim; i++) {
        const int jend = same ? i + 1 : B->nprim;
        for (int j = 0; j < jend; j++) {
            simint_shell a, b;
            simint_initialize_shell(&a);
            simint_initialize_shell(&b);
            simint_create_shell(1, A->am, A->x, A->y, A->z, A->alpha + i, A->coef + i, &a);
            simint_create_shell(1, B->am, B->x, B->y, B->z, B->alpha + j, B->coef + j, &b);
            pairs.push_back(a);
            pairs.push_back(b);
        }
    }
    const int npair = (int)(pairs.size() / 2);
    simint_multi_shellpair P;
    simint_initialize_multi_shellpair(&P);
    simint_create_multi_shellpair2(npair, pairs.data(), &P, SIMINT_SCREEN_NONE);
    std::vector<double> mx(npair > 0 ? npair : 1);
    const bool ok = simint_rt::host_pair_bounds(&P, mx.data()) == 0;
    simint_free_multi_shellpair(&P);
    for (auto &s : pairs) simint_free_shell(&s);
    if (!ok) return NAN;
    double total = 0.0;
    for (int k = 0; k < npair; k++) {
        if (out) out[k] = mx[k];
        total = mx[k] > total ? mx[k] : total;
    }
    return total;
}

double simint_primscreen_fastschwarz(const simint_shell *A, const simint_shell *B, double *out)
{
    const int same = compare_shell(A, B);
    double total = 0.0;
    int idx = 0;
    for (int i = 0; i < A->nprim; i++) {
        const int jend = same ? i + 1 : B->nprim;
        for (int j = 0; j < jend; j++) {
            const double G = fast_value(A, B, i, j);
            if (out) out[idx] = G;
            total = G > total ? G : total;
            idx++;
        }
    }
    return total;
}

}  // extern "C"

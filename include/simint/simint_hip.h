/* Device-resident and batched entry points of the MI355X ERI engine.
 *
 * These have no counterpart in the reference's API; they sit next to the
 * drop-in simint_compute_eri for callers that keep shell-pair data and
 * integrals in GPU memory (HBM) and submit many quartets per call.  Device
 * pointers are plain HIP device addresses.  All work is queued on the
 * calling thread's library stream (simint_hip_stream); call
 * simint_hip_synchronize before reading results. */
#pragma once

#include "simint/shell/shell.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Shell pairs of one (am1, am2) resident in device memory, in the order of
 * the source structure (its first nshell12_clip pairs). */
typedef struct simint_hip_pairs simint_hip_pairs;
/* A shell set resident in device memory (simint_hip_shells_create below). */
typedef struct simint_hip_shells simint_hip_shells;

int  simint_hip_pairs_create(struct simint_multi_shellpair const * P, simint_hip_pairs ** out);
void simint_hip_pairs_free(simint_hip_pairs * p);
int  simint_hip_pairs_count(simint_hip_pairs const * p);
/* device memory (bytes) the pair set holds */
size_t simint_hip_pairs_bytes(simint_hip_pairs const * p);

/* (ab|cd) for bra pairs b = b0 .. b0+nbra-1 of P and, for each, ket pairs
 * k = 0 .. nket[b-b0]-1 of Q: the per-bra prefix a Schwarz screen over
 * bound-sorted pairs produces.  Quartet (b, k) is written to
 *   d_out + out_off[b-b0] + k * ncart(abcd)       (doubles, Psi4 order).
 * nket and out_off are host arrays.  screen_tol as in simint_compute_eri.
 * Returns the number of quartets queued, or a negative SIMINT_ERROR_*. */
long long simint_hip_compute_eri_prefix(simint_hip_pairs const * P, simint_hip_pairs const * Q,
                                        int b0, int nbra, int const * nket,
                                        long long const * out_off,
                                        double screen_tol, double * d_out);

/* Quartet lists: quartet i is bra pair d_qbra[i] of P against ket pair
 * d_qket[i] of Q, its ncart(abcd) integrals written at d_out + d_qout[i]
 * (or at d_out + i * ncart(abcd) when d_qout is NULL).  All arrays are
 * device arrays, and the call only queues kernels (no host copies or
 * synchronisation), so it may be captured into a graph.  Quartets of any
 * mix of primitive counts; ordering the list by primitive count keeps the
 * lanes of a wavefront in step.  Returns nquartet or a negative SIMINT_ERROR_*. */
long long simint_hip_compute_eri_list(simint_hip_pairs const * P, simint_hip_pairs const * Q,
                                      long long nquartet, int const * d_qbra, int const * d_qket,
                                      long long const * d_qout, double screen_tol, double * d_out);

/* First derivatives of a quartet list (the layout of simint_compute_eri_deriv:
 * 12 doubles per Cartesian, [Ax Ay Az Bx ... Dz]) at d_out + d_qout[i] (or
 * i * 12 ncart(abcd)).  d_work: device scratch of simint_hip_deriv1_worksize
 * doubles.  The class needs its (a+1) classes built (AM <= SIMINT_OSTEI_DERIV1_MAXAM). */
long long simint_hip_deriv1_worksize(simint_hip_pairs const * P, simint_hip_pairs const * Q, long long nquartet);
long long simint_hip_compute_eri_deriv1_list(simint_hip_pairs const * P, simint_hip_pairs const * Q,
                                             long long nquartet, int const * d_qbra, int const * d_qket,
                                             long long const * d_qout, double screen_tol, double * d_work,
                                             double * d_out);

/* No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same as
 * simint_hip_compute_eri_deriv1_list, except that the derivative scratch is a
 * fixed array inside the library (SIMINT_HIP_NOALLOC_SCRATCH_DOUBLES doubles,
 * 64 MiB, held by the device until the process ends) instead of d_work, so
 * the call allocates nothing.  simint_init must have run first (it looks up
 * the array's device address); a list whose scratch,
 * simint_hip_deriv1_worksize(P, Q, nquartet) doubles, exceeds the array is
 * refused with SIMINT_ERROR_ARGUMENT before anything is queued.  Every call
 * on the device shares the one array and leaves the last call's values in
 * it (nothing needs to start at zero): calls that overlap on the device,
 * from other threads or on other streams, race and must be ordered by the
 * caller. */
#define SIMINT_HIP_NOALLOC_SCRATCH_DOUBLES (8ull << 20)
long long simint_hip_compute_eri_deriv1_list_noalloc(simint_hip_pairs const * P, simint_hip_pairs const * Q,
                                                     long long nquartet, int const * d_qbra, int const * d_qket,
                                                     long long const * d_qout, double screen_tol, double * d_out);

/* First derivatives over ket ranges (as _range, FP64): quartet (b0+i, k),
 * k in [ket_lo[i], ket_hi[i]), written densely in (bra, ket) order -- quartet
 * number sum_{j<i}(ket_hi[j]-ket_lo[j]) + (k - ket_lo[i]) at d_out + number *
 * 12 ncart(abcd).  d_work: simint_hip_deriv1_worksize(P, Q, nquartet) doubles. */
long long simint_hip_compute_eri_deriv1_range(simint_hip_pairs const * P, simint_hip_pairs const * Q,
                                              int b0, int nbra, int const * ket_lo, int const * ket_hi,
                                              double screen_tol, double * d_work, double * d_out);

/* Graph capture of the calling thread's library stream: every list call
 * between begin and end is recorded (not run); simint_hip_graph_launch
 * replays the whole batch as one graph launch. */
typedef struct simint_hip_graph simint_hip_graph;
int  simint_hip_capture_begin(void);
int  simint_hip_capture_end(simint_hip_graph ** out);
int  simint_hip_graph_launch(simint_hip_graph * g);
void simint_hip_graph_free(simint_hip_graph * g);

/* As _prefix with a ket range per bra: kets k = ket_lo[i] .. ket_hi[i]-1
 * (ket_lo may be NULL: from 0), quartet (b, k) written at
 * d_out + out_off[i] + (k - ket_lo[i]) * ncart(abcd).  precision 0: FP64;
 * 1: mixed -- the VRR in FP32, with the Boys function, the prefactors, the
 * primitive contraction and the HRR in FP64 (classes whose selected kernel is
 * not the lane-per-quartet family, or above d, run FP64 regardless).  The
 * caller decides which quartets may take FP32 (their Schwarz bound). */
long long simint_hip_compute_eri_range(simint_hip_pairs const * P, simint_hip_pairs const * Q,
                                       int b0, int nbra, int const * ket_lo, int const * ket_hi,
                                       long long const * out_off, double screen_tol, int precision,
                                       double * d_out);

/* (bb|bb) for every pair b of P, at d_out + b * ncart(abab). */
long long simint_hip_compute_eri_diag(simint_hip_pairs const * P, double * d_out);

/* One-electron integrals of many shell pairs (no counterpart in the
 * reference, whose osoei functions take one pair per call).  A shell set is
 * copied to the device once; pair i is shells d_pairs[2i] and d_pairs[2i+1]
 * of the set, its ncart(a) x ncart(b) values (Psi4 order) written at
 * d_out + d_off[i].  kind 0: overlap, 1: kinetic energy, 2: nuclear
 * attraction -sum_C Z_C <a|1/r_C|b> of the ncenter point charges in
 * d_charges (device, one row of Z, x, y, z per charge).  Angular momentum up
 * to 5.  Queues the kernel on the calling thread's stream; returns npair or
 * a negative SIMINT_ERROR_*. */
int  simint_hip_shells_create(int nshell, struct simint_shell const * shells, simint_hip_shells ** out);
void simint_hip_shells_free(simint_hip_shells * s);
long long simint_hip_compute_osoei(int kind, simint_hip_shells const * S, long long npair, int const * d_pairs,
                                   long long const * d_off, int ncenter, double const * d_charges, double * d_out);

/* Device-side front end (no counterpart in the reference, which fills shell
 * pairs and screening values on the host, skel/simint/shell/shell.c:425-582,
 * shell_screen.c:226-270):
 *   simint_hip_pairs_from_shells -- shell pairs (host list of shell index
 *     pairs into a device shell set, all of one (am1, am2)) filled by a
 *     kernel: the values simint_create_multi_shellpair2 computes, the
 *     fast-Schwarz screening values and their descending sort when
 *     screen_method is SIMINT_SCREEN_FASTSCHWARZ (SIMINT_SCREEN_NONE also
 *     accepted);
 *   simint_hip_pair_bounds -- the Schwarz bound of every pair, the largest
 *     |(bb|bb)| diagonal element, into d_bound (device, one per pair);
 *   simint_hip_order_pairs -- the permutation that sorts the pairs by
 *     primitive count, then bound, both descending (stable), into d_order;
 *   simint_hip_screen_counts -- for bras with bounds d_bra_bound and a
 *     bound-sorted (descending) ket group d_ket_bound, the number of kets with
 *     bound_b bound_k >= screen_tol^2 (with `same`, a bra's own group: at most
 *     b + 1), into d_counts (device).
 * The last three queue work on the calling thread's stream. */
int simint_hip_pairs_from_shells(simint_hip_shells const * S, int npair, int const * pairs, int screen_method,
                                 simint_hip_pairs ** out);
long long simint_hip_pair_bounds(simint_hip_pairs const * P, double * d_bound);
int simint_hip_order_pairs(simint_hip_pairs const * P, double const * d_bound, int * d_order);
int simint_hip_screen_counts(int nbra, double const * d_bra_bound, int nket, double const * d_ket_bound, int same,
                             double screen_tol, int * d_counts);
/* A device pair set's primitive-pair data copied back to host arrays (for
 * inspection and tests): fields holds 14 x nprim doubles, the arrays x y z
 * PA_x PA_y PA_z PB_x PB_y PB_z alpha prefac alpha2 beta2 screen one after
 * another; nprim12 the count per pair; ab the 3 x npair A-B components.  Any
 * of them may be NULL.  Returns nprim (the primitive pairs held) or a
 * negative SIMINT_ERROR_*; waits for the copies. */
long long simint_hip_pairs_export(simint_hip_pairs const * P, double * fields, int * nprim12, double * ab);

/* Coulomb and exchange digestion (a Fock build's consumer of the integrals;
 * no counterpart in the reference, which returns integrals only).
 * simint_hip_compute_jk_range computes the quartets of one range launch as
 * simint_hip_compute_eri_range does (integrals into d_out) and then, on the
 * same stream, contracts them with the density:
 *     J_mn = sum_ls (mn|ls) D_ls,     K_mn = sum_ls (ml|ns) D_ls,
 * counting each quartet with the weight of its permutational images, so
 * that digesting every unique shell quartet (ab|cd) -- a >= b, c >= d,
 * (ab) >= (cd) in any fixed pair order -- exactly once and then calling
 * simint_hip_jk_symmetrize gives J and K over all basis functions.
 * jk->bra_fun / ket_fun (device, 2 ints per pair of P / Q): the first basis
 * function of each shell of the pair; two shells are the same shell when
 * their first functions are equal.  jk->same != 0: P and Q are one pair set
 * and quartet (b, b) is its own bra-ket image.  density, J, K: nbf x nbf,
 * row-major, density symmetric; J and K accumulate (atomically, so the
 * summation order and the last bits vary between runs); either may be
 * NULL.  Returns the quartets digested or a negative SIMINT_ERROR_*. */
typedef struct simint_hip_jk {
    int nbf;
    int same;
    const int *bra_fun;
    const int *ket_fun;
    const double *density;
    double *J;
    double *K;
} simint_hip_jk;
long long simint_hip_compute_jk_range(simint_hip_pairs const * P, simint_hip_pairs const * Q, int b0, int nbra,
                                      int const * ket_lo, int const * ket_hi, long long const * out_off,
                                      double screen_tol, int precision, double * d_out,
                                      simint_hip_jk const * jk);
/* J <- (J + J^T) / 2 and K <- (K + K^T) / 2 (device, nbf x nbf, either may
 * be NULL), on the calling thread's stream. */
int simint_hip_jk_symmetrize(int nbf, double * d_J, double * d_K);
/* Exchange updates of a bra pair's rows are summed in the local data share
 * over all the bra's kets before they reach K (1, the default, where the
 * rows fit) or added to K per quartet (0).  Process-wide; returns the
 * previous setting. */
int simint_hip_set_jk_rows(int enable);

/* Cartesian order of the drop-in calls' output (simint_compute_eri,
 * simint_compute_eri_deriv and the one-electron calls), process-wide.
 * The reference fixes it when its library is generated (Psi4 or GAMESS,
 * generator/CMakeLists.txt:9-10); here it is switched at run time:
 * GAMESS orders d as xx yy zz xy xz yz and f as xxx yyy zzz xxy xxz xyy yyz
 * xzz yzz xyz, the other AMs as Psi4.  The device-resident calls always
 * write Psi4 order.  Returns the previous ordering or SIMINT_ERROR_ARGUMENT. */
#define SIMINT_ORDERING_PSI4   0
#define SIMINT_ORDERING_GAMESS 1
int simint_hip_set_ordering(int ordering);
int simint_hip_ordering(void);

int   simint_hip_synchronize(void);
void *simint_hip_stream(void);                /* hipStream_t of the calling thread */
/* Queue the calling thread's later calls on `stream` (a hipStream_t; NULL:
 * back to the thread's own stream), e.g. to spread the classes of one batch
 * over several streams that a graph capture joins.  The calls that stage
 * host data through the library's per-thread buffers (_prefix, _range,
 * _diag, the drop-in calls) synchronise only the current stream before
 * reusing those buffers: synchronise the previous stream yourself before
 * switching between such calls; list calls take device arrays and need no care. */
int   simint_hip_set_stream(void * stream);
/* Lanes per quartet (1, 2, 4, 8 or 16) of the calling thread's later
 * lane-per-quartet simint_hip_compute_eri_list launches: each lane sums a
 * strided share of the quartet's primitive quartets and the shares are added
 * across the lanes before the HRR.  0 (the default): chosen by list length,
 * so that a launch has about 65536 lanes.  Returns the previous setting or
 * a negative SIMINT_ERROR_*.  No counterpart in the reference (its kernels
 * vectorise over kets on the CPU). */
int   simint_hip_set_list_split(int lanes);
const char *simint_hip_last_error(void);

/* Kernel family used for a class: 1 = lane-per-quartet (registers),
 * 2 = wave-per-quartet (LDS-staged), 3 / 4 = split lane-per-quartet (the
 * contracted HRR inputs processed in groups of about 64 / 128, through a
 * device staging area), 0 = class not built.  The setter lets a caller (the
 * autotuner, tests) pick a family; it returns the previous one or a
 * negative value if that family is not built for the class. */
int simint_hip_class_kernel(int la, int lb, int lc, int ld);
int simint_hip_set_class_kernel(int la, int lb, int lc, int ld, int family);
/* The same for the quartet-list calls: 1 = lane-per-quartet, 2 = wave-per-quartet. */
int simint_hip_class_list_kernel(int la, int lb, int lc, int ld);
int simint_hip_set_class_list_kernel(int la, int lb, int lc, int ld, int family);

#ifdef __cplusplus
}
#endif

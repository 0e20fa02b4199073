// Device-side data layout shared by the host runtime and the generated
// class kernels.
//
// Shell-pair primitive data lives in HBM in two views:
//   * PairsLinear  -- one structure-of-arrays over all primitive pairs, pair p
//                     owning [poff[p], poff[p] + np[p]).  Used for the bra,
//                     which is uniform across a wavefront, so its loads are
//                     scalar (SGPR) loads.
//   * PairsTiled   -- pairs grouped into tiles of 64 (one per lane); tile t
//                     stores primitive j of lane l at toff[t] + 64*j + l, so a
//                     wavefront reading "its" primitive j touches one
//                     contiguous 512-byte run per field.  Padded lanes and
//                     padded primitives carry prefac = 0, alpha = 1 and zero
//                     coordinates, so they contribute exactly zero.
// The per-primitive fields are those of simint_multi_shellpair
// (skel/simint/shell/shell.h:47-85) plus 1/alpha, precomputed on the host.
#pragma once
#include <stdint.h>

#define SIMINT_TILE 64

struct PairsLinear {
    const double *alpha, *inv_alpha, *prefac;
    const double *x, *y, *z;
    const double *pax, *pay, *paz;
    const double *pbx, *pby, *pbz;
    const double *screen;
    const double *alpha2, *beta2;
    const int *poff;               // per pair
    const int *np;                 // per pair
    const double *abx, *aby, *abz; // per pair
};

struct PairsTiled {
    const double *alpha, *inv_alpha, *prefac;
    const double *x, *y, *z;
    const double *pax, *pay, *paz;
    const double *pbx, *pby, *pbz;
    const double *screen;
    const double *alpha2, *beta2;
    const long long *toff;         // per tile (element offset, multiple of 64)
    const int *tnp;                // per tile: primitive rows
    const double *abx, *aby, *abz; // per tile lane: t*64 + l
    const long long *col;          // per tile lane: output column (-1 = none)
};

// One wavefront's work: bra pair `bra` against lanes [lane_lo, lane_hi) of
// ket tile `tile`, writing lane l's integrals at
// out + obase + col[tile*64 + l] * stride.
struct WorkItem {
    int bra;
    int tile;
    short lane_lo;
    short lane_hi;
    int pad;
    long long obase;
};

// Quartet-list work (any bra against any ket): quartet q of a launch is
// bra pair qbra[q] against ket pair qket[q], both in the linear view, its
// ncart(abcd) integrals written contiguously at out + qout[q] (or, with
// qout null, at out + q * ostride).
struct EriListArgs {
    PairsLinear bra;
    PairsLinear ket;
    const int *qbra;
    const int *qket;
    const long long *qout;
    long long ostride;
    long long nquartet;
    double *out;
    const double *boys_grid;
    double tol2;
    double ket_screen_max;
    int split;                 // lanes per quartet (1, 2, 4, 8) of the lane-per-quartet list kernels
};

struct EriArgs {
    PairsLinear bra;
    PairsTiled ket;
    const WorkItem *items;
    long long nitems;
    double *out;
    const double *boys_grid;       // 366 x 32, row-major
    double tol2;                   // primitive screening threshold (squared); <= 0 disables
    double ket_screen_max;         // simint_multi_shellpair.screen_max of the ket set
    long long stride;              // doubles per quartet in the output
    double *scratch;               // per-work-item staging (split kernels only)
};

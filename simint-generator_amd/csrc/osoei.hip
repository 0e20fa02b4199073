// One-electron integrals on the GPU: overlap, kinetic energy and the nuclear
// attraction of point charges (skel/simint/osoei/osoei_terms.c:7-166,
// osoei_overlap.c:10-66, osoei_ke.c:10-67, osoei_potential.c:11-218).
//
// One 64-lane workgroup per shell pair, primitive pairs in the reference's
// order (first shell outer), point charges outermost for the potential:
//   * overlap / kinetic energy: three lanes build the 1D Obara-Saika tables
//     S_ij (and T_ij) of one primitive pair, one direction each, in the local
//     data share; every lane then forms its own Cartesian products
//     S_x S_y S_z (T_x S_y S_z + S_x T_y S_z + S_x S_y T_z) and accumulates
//     them, times c_a c_b, in registers;
//   * potential: the lanes fill the Obara-Saika nuclear-attraction blocks
//     Theta^(m)_{i,j} block by block (first centre up from the Boys values,
//     then the second centre), each block's (m, component) entries spread
//     over the lanes, and accumulate -Z c_a c_b Theta^(0)_{la,lb}.
// Each block keeps only the auxiliary orders its consumers read:
// (i,0) for m <= la+lb-i and (i,j>0) for m <= lb-j (the reference keeps one
// more order in the second blocks, whose values never reach the result).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string.h>
#include <vector>

#include "simint/simint.h"
#include "simint/simint_hip.h"
#include "boys_device.h"
#include "runtime_internal.h"

#define OEI_MAXL 5
#define OEI_NC ((OEI_MAXL + 1) * (OEI_MAXL + 2) / 2)

namespace {

// Cartesian component n of AM l in Psi4 order (x from l down, then y from
// l-x down): exponents, the recurrence direction (smallest non-zero
// exponent, rightmost on ties: the reference's recur_lookup tables) and the
// indices of the components one and two steps down in each direction.
struct OeiCart {
    signed char e[3];
    signed char dir;
    signed char dn1[3];
    signed char dn2[3];
};
struct OeiCartTable {
    OeiCart c[OEI_MAXL + 1][OEI_NC];
};

constexpr int oei_index(int l, int x, int y) { return (l - x) * (l - x + 1) / 2 + (l - x - y); }

constexpr OeiCartTable oei_make_table()
{
    OeiCartTable t{};
    for (int l = 0; l <= OEI_MAXL; l++) {
        int n = 0;
        for (int x = l; x >= 0; x--)
            for (int y = l - x; y >= 0; y--, n++) {
                OeiCart &c = t.c[l][n];
                const int e[3] = {x, y, l - x - y};
                c.e[0] = (signed char)e[0];
                c.e[1] = (signed char)e[1];
                c.e[2] = (signed char)e[2];
                c.dir = -1;
                if (l > 0) {
                    int m = 99;
                    for (int d = 0; d < 3; d++)
                        if (e[d] > 0 && e[d] < m) m = e[d];
                    for (int d = 2; d >= 0; d--)
                        if (e[d] == m) {
                            c.dir = (signed char)d;
                            break;
                        }
                }
                for (int d = 0; d < 3; d++) {
                    int e1[3] = {e[0], e[1], e[2]}, e2[3] = {e[0], e[1], e[2]};
                    e1[d] -= 1;
                    e2[d] -= 2;
                    c.dn1[d] = (signed char)(e1[d] >= 0 ? oei_index(l - 1, e1[0], e1[1]) : -1);
                    c.dn2[d] = (signed char)(e2[d] >= 0 ? oei_index(l - 2, e2[0], e2[1]) : -1);
                }
            }
    }
    return t;
}

__constant__ OeiCartTable kOei = oei_make_table();

constexpr int oei_nc(int l) { return (l + 1) * (l + 2) / 2; }

// doubles of the potential's blocks for la = lb = MAXL
constexpr int oei_th_size(int maxl)
{
    int s = 0;
    for (int i = 0; i <= maxl; i++) s += (2 * maxl - i + 1) * oei_nc(i);
    for (int i = 0; i <= maxl; i++)
        for (int j = 1; j <= maxl; j++) s += (maxl - j + 1) * oei_nc(i) * oei_nc(j);
    return s;
}

struct OeiArgs {
    const int *am, *nprim, *poff;   // per shell
    const double *xyz;              // 3 per shell
    const double *alpha, *coef;     // per primitive
    const int *pairs;               // 2 per pair
    const long long *off;           // output offset per pair
    long long npair;
    int ncenter;
    const double *charges;          // Z, x, y, z per charge
    double *out;
    const double *boys;
};

constexpr double kPi = 3.14159265358979323846;

template <int L>
__device__ __forceinline__ void boys_upto(const double *grid, double x, double *F)
{
    boys_eval<L>(grid, x, F);
}

// F_m(x), m = 0..L, for a run-time L <= 2 MAXL
template <int MAXL>
__device__ __forceinline__ void boys_any(const double *grid, int L, double x, double *F)
{
    switch (L) {
    case 0: boys_upto<0>(grid, x, F); break;
    case 1: boys_upto<1>(grid, x, F); break;
    case 2: boys_upto<2>(grid, x, F); break;
    case 3: boys_upto<3>(grid, x, F); break;
    case 4: boys_upto<4>(grid, x, F); break;
    case 5: boys_upto<5>(grid, x, F); break;
    case 6: boys_upto<6>(grid, x, F); break;
    default:
        if constexpr (MAXL > 3) {
            switch (L) {
            case 7: boys_upto<7>(grid, x, F); break;
            case 8: boys_upto<8>(grid, x, F); break;
            case 9: boys_upto<9>(grid, x, F); break;
            case 10: boys_upto<10>(grid, x, F); break;
            default: break;
            }
        }
        break;
    }
}

// overlap (KIND 0) and kinetic energy (KIND 1)
template <int KIND, int MAXL>
__global__ void __launch_bounds__(64) osoei_sk(const OeiArgs a)
{
    constexpr int NT = (MAXL + 1) * (MAXL + 1);
    constexpr int NPL = (oei_nc(MAXL) * oei_nc(MAXL) + 63) / 64;
    __shared__ double tab[6 * NT];
    const int lane = threadIdx.x;
    for (long long pr = blockIdx.x; pr < a.npair; pr += gridDim.x) {
        const int A = a.pairs[2 * pr], B = a.pairs[2 * pr + 1];
        const int am1 = a.am[A], am2 = a.am[B];
        const int nam1 = am1 + 1, nam2 = am2 + 1, n12 = nam1 * nam2;
        const int nc1 = oei_nc(am1), nc2 = oei_nc(am2), nout = nc1 * nc2;
        const double *xa = a.xyz + 3 * A, *xb = a.xyz + 3 * B;
        double acc[NPL];
#pragma unroll
        for (int k = 0; k < NPL; k++) acc[k] = 0.0;
        for (int i = 0; i < a.nprim[A]; i++)
            for (int j = 0; j < a.nprim[B]; j++) {
                const double a1 = a.alpha[a.poff[A] + i], a2 = a.alpha[a.poff[B] + j];
                if (lane < 3) {
                    // osoei_terms.c: the 1D tables of direction d = lane
                    const int d = lane;
                    const double A_ = xa[d], B_ = xb[d];
                    const double AB = A_ - B_;
                    const double oop = 1.0 / (a1 + a2);
                    const double mu = a1 * a2 * oop;
                    const double oo2p = 0.5 * oop;
                    const double P = (a1 * A_ + a2 * B_) * oop;
                    const double PA = P - A_, PB = P - B_;
                    double *s = tab + d * n12;
                    double *t = tab + (d + 3) * n12;
                    s[0] = sqrt(kPi * oop) * exp(-mu * (AB * AB));
                    if (KIND == 1) t[0] = s[0] * (a1 - (2.0 * a1 * a1) * (PA * PA + oo2p));
                    for (int ii = 1; ii < nam1; ii++) {
                        double v = PA * s[(ii - 1) * nam2];
                        if (ii > 1) v += (ii - 1) * oo2p * s[(ii - 2) * nam2];
                        s[ii * nam2] = v;
                        if (KIND == 1) {
                            double w = PA * t[(ii - 1) * nam2] + 2 * mu * v;
                            if (ii > 1) w += (ii - 1) * oo2p * t[(ii - 2) * nam2] - a2 * oop * (ii - 1) * s[(ii - 2) * nam2];
                            t[ii * nam2] = w;
                        }
                    }
                    for (int jj = 1; jj < nam2; jj++) {
                        double v = PB * s[jj - 1];
                        if (jj > 1) v += (jj - 1) * oo2p * s[jj - 2];
                        s[jj] = v;
                        if (KIND == 1) {
                            double w = PB * t[jj - 1] + 2 * mu * v;
                            if (jj > 1) w += (jj - 1) * oo2p * t[jj - 2] - a1 * oop * (jj - 1) * s[jj - 2];
                            t[jj] = w;
                        }
                    }
                    for (int ii = 1; ii < nam1; ii++)
                        for (int jj = 1; jj < nam2; jj++) {
                            double v = PB * s[ii * nam2 + jj - 1] + oo2p * ii * s[(ii - 1) * nam2 + jj - 1];
                            if (jj > 1) v += oo2p * (jj - 1) * s[ii * nam2 + jj - 2];
                            s[ii * nam2 + jj] = v;
                            if (KIND == 1) {
                                double w = PB * t[ii * nam2 + jj - 1] + oo2p * ii * t[(ii - 1) * nam2 + jj - 1] + 2 * mu * v;
                                if (jj > 1) w += oo2p * (jj - 1) * t[ii * nam2 + jj - 2] - a1 * oop * (jj - 1) * s[ii * nam2 + jj - 2];
                                t[ii * nam2 + jj] = w;
                            }
                        }
                }
                __syncthreads();
                const double pf = a.coef[a.poff[A] + i] * a.coef[a.poff[B] + j];
#pragma unroll
                for (int k = 0; k < NPL; k++) {
                    const int o = lane + 64 * k;
                    if (o < nout) {
                        const int n = o / nc2, m = o - n * nc2;
                        const signed char *e1 = kOei.c[am1][n].e, *e2 = kOei.c[am2][m].e;
                        const int ix = e1[0] * nam2 + e2[0], iy = e1[1] * nam2 + e2[1], iz = e1[2] * nam2 + e2[2];
                        double v;
                        if (KIND == 0) {
                            v = tab[ix] * tab[n12 + iy] * tab[2 * n12 + iz];
                        } else {
                            v = tab[3 * n12 + ix] * tab[n12 + iy] * tab[2 * n12 + iz] +
                                tab[ix] * tab[4 * n12 + iy] * tab[2 * n12 + iz] +
                                tab[ix] * tab[n12 + iy] * tab[5 * n12 + iz];
                        }
                        acc[k] += pf * v;
                    }
                }
                __syncthreads();
            }
        double *o = a.out + a.off[pr];
#pragma unroll
        for (int k = 0; k < NPL; k++)
            if (lane + 64 * k < nout) o[lane + 64 * k] = acc[k];
    }
}

// nuclear attraction of the point charges
template <int MAXL>
__global__ void __launch_bounds__(64) osoei_potential(const OeiArgs a)
{
    constexpr int NPL = (oei_nc(MAXL) * oei_nc(MAXL) + 63) / 64;
    __shared__ double th[oei_th_size(MAXL)];
    __shared__ int boff[MAXL + 1][MAXL + 1];
    const int lane = threadIdx.x;
    for (long long pr = blockIdx.x; pr < a.npair; pr += gridDim.x) {
        const int A = a.pairs[2 * pr], B = a.pairs[2 * pr + 1];
        const int am1 = a.am[A], am2 = a.am[B], L = am1 + am2;
        const int nc1 = oei_nc(am1), nc2 = oei_nc(am2), nout = nc1 * nc2;
        __syncthreads();   // the previous pair's last reads of boff are done
        if (lane == 0) {
            int s = 0;
            for (int i = 0; i <= am1; i++) {
                boff[i][0] = s;
                s += (L - i + 1) * oei_nc(i);
                for (int j = 1; j <= am2; j++) {
                    boff[i][j] = s;
                    s += (am2 - j + 1) * oei_nc(i) * oei_nc(j);
                }
            }
        }
        __syncthreads();
        const double xa[3] = {a.xyz[3 * A], a.xyz[3 * A + 1], a.xyz[3 * A + 2]};
        const double xb[3] = {a.xyz[3 * B], a.xyz[3 * B + 1], a.xyz[3 * B + 2]};
        const double AB[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
        const double AB2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        double acc[NPL];
#pragma unroll
        for (int k = 0; k < NPL; k++) acc[k] = 0.0;
        for (int c = 0; c < a.ncenter; c++) {
            const double Z = a.charges[4 * c];
            const double C[3] = {a.charges[4 * c + 1], a.charges[4 * c + 2], a.charges[4 * c + 3]};
            for (int ia = 0; ia < a.nprim[A]; ia++) {
                const double a1 = a.alpha[a.poff[A] + ia];
                const double a1xyz[3] = {a1 * xa[0], a1 * xa[1], a1 * xa[2]};
                for (int ib = 0; ib < a.nprim[B]; ib++) {
                    const double a2 = a.alpha[a.poff[B] + ib];
                    const double p = a1 + a2;
                    const double oop = 1.0 / (a1 + a2);
                    const double mu = a1 * a2 * oop;
                    const double oo2p = 0.5 * oop;
                    double PA[3], PB[3], PC[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double P = (a1xyz[d] + a2 * xb[d]) * oop;
                        PA[d] = P - xa[d];
                        PB[d] = P - xb[d];
                        PC[d] = P - C[d];
                    }
                    const double PC2 = PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2];
                    double F[2 * MAXL + 1];
                    boys_any<MAXL>(a.boys, L, PC2 * p, F);
                    const double e = exp(-mu * AB2);
                    // th[0][0][m] = F_m 2 pi / p e^{-mu AB^2} (osoei_potential.c:79-85)
#pragma unroll
                    for (int m = 0; m <= 2 * MAXL; m++)
                        if (m <= L && lane == m) th[m] = F[m] * 2 * kPi * oop * e;
                    __syncthreads();
                    for (int i = 0; i <= am1; i++) {
                        const int ni = oei_nc(i);
                        const int ni1 = i > 0 ? oei_nc(i - 1) : 0, ni2 = i > 1 ? oei_nc(i - 2) : 0;
                        if (i > 0) {
                            // (i,0) up the first centre (osoei_potential.c:104-134)
                            double *w = th + boff[i][0];
                            const double *w14 = th + boff[i - 1][0];
                            const double *w25 = i > 1 ? th + boff[i - 2][0] : th;
                            const int ne = (L - i + 1) * ni;
                            for (int q = lane; q < ne; q += 64) {
                                const int m = q / ni, n = q - m * ni;
                                const OeiCart &cc = kOei.c[i][n];
                                const int d = cc.dir;
                                double v = PA[d] * w14[m * ni1 + cc.dn1[d]] - PC[d] * w14[(m + 1) * ni1 + cc.dn1[d]];
                                if (cc.e[d] > 1)
                                    v += oo2p * (cc.e[d] - 1) * (w25[m * ni2 + cc.dn2[d]] - w25[(m + 1) * ni2 + cc.dn2[d]]);
                                w[q] = v;
                            }
                            __syncthreads();
                        }
                        for (int j = 1; j <= am2; j++) {
                            // (i,j) up the second centre (osoei_potential.c:136-207)
                            const int nj = oei_nc(j), nj1 = oei_nc(j - 1), nj2 = j > 1 ? oei_nc(j - 2) : 0;
                            double *w = th + boff[i][j];
                            const double *w14 = th + boff[i][j - 1];
                            const double *w36 = j > 1 ? th + boff[i][j - 2] : th;
                            const double *w25 = i > 0 ? th + boff[i - 1][j - 1] : th;
                            const int nij = ni * nj;
                            const int ne = (am2 - j + 1) * nij;
                            for (int q = lane; q < ne; q += 64) {
                                const int m = q / nij, r = q - m * nij;
                                const int n = r / nj, o = r - n * nj;
                                const OeiCart &cb = kOei.c[j][o];
                                const int d = cb.dir;
                                const int ei = kOei.c[i][n].e[d];
                                const int i1 = cb.dn1[d];
                                double v = PB[d] * w14[(m * ni + n) * nj1 + i1] - PC[d] * w14[((m + 1) * ni + n) * nj1 + i1];
                                if (ei > 0) {
                                    const int a1i = kOei.c[i][n].dn1[d];
                                    v += oo2p * ei * w25[(m * ni1 + a1i) * nj1 + i1] -
                                         oo2p * ei * w25[((m + 1) * ni1 + a1i) * nj1 + i1];
                                }
                                if (cb.e[d] > 1) {
                                    const int i2 = cb.dn2[d];
                                    v += oo2p * (cb.e[d] - 1) *
                                         (w36[(m * ni + n) * nj2 + i2] - w36[((m + 1) * ni + n) * nj2 + i2]);
                                }
                                w[q] = v;
                            }
                            __syncthreads();
                        }
                    }
                    const double ca = a.coef[a.poff[A] + ia], cb = a.coef[a.poff[B] + ib];
                    const double *res = th + boff[am1][am2];
#pragma unroll
                    for (int k = 0; k < NPL; k++) {
                        const int o = lane + 64 * k;
                        if (o < nout) acc[k] -= res[o] * Z * ca * cb;
                    }
                    __syncthreads();   // th is overwritten by the next primitive pair
                }
            }
        }
        double *o = a.out + a.off[pr];
#pragma unroll
        for (int k = 0; k < NPL; k++)
            if (lane + 64 * k < nout) o[lane + 64 * k] = acc[k];
    }
}

template <int MAXL>
hipError_t launch_osoei(int kind, const OeiArgs &a, hipStream_t s)
{
    const unsigned grid = (unsigned)std::min<long long>(a.npair, 1 << 20);
    if (kind == 0) hipLaunchKernelGGL((osoei_sk<0, MAXL>), dim3(grid), dim3(64), 0, s, a);
    else if (kind == 1) hipLaunchKernelGGL((osoei_sk<1, MAXL>), dim3(grid), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((osoei_potential<MAXL>), dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

// host image of a shell set: am, nprim, poff, xyz, alpha, coef (8-byte
// aligned sub-arrays in one buffer)
struct ShellImage {
    std::vector<char> buf;
    size_t o_am = 0, o_np = 0, o_po = 0, o_xyz = 0, o_al = 0, o_co = 0;
    int maxam = 0;
    size_t add(size_t bytes)
    {
        const size_t off = (buf.size() + 255) & ~(size_t)255;
        buf.resize(off + std::max(bytes, (size_t)8));
        return off;
    }
    template <class T> T *at(size_t off) { return reinterpret_cast<T *>(buf.data() + off); }
};

int check_shell(const simint_shell *s)
{
    if (!s || s->am < 0 || s->nprim < 0) return SIMINT_ERROR_ARGUMENT;
    if (s->nprim > 0 && (!s->alpha || !s->coef)) return SIMINT_ERROR_ARGUMENT;
    if (s->am > OEI_MAXL) return SIMINT_ERROR_UNSUPPORTED;
    return 0;
}

void pack_shells(ShellImage &im, int n, const simint_shell *const *sh)
{
    long long np = 0;
    for (int i = 0; i < n; i++) np += sh[i]->nprim;
    im.o_am = im.add(sizeof(int) * n);
    im.o_np = im.add(sizeof(int) * n);
    im.o_po = im.add(sizeof(int) * n);
    im.o_xyz = im.add(sizeof(double) * 3 * n);
    im.o_al = im.add(sizeof(double) * np);
    im.o_co = im.add(sizeof(double) * np);
    int acc = 0;
    for (int i = 0; i < n; i++) {
        im.at<int>(im.o_am)[i] = sh[i]->am;
        im.at<int>(im.o_np)[i] = sh[i]->nprim;
        im.at<int>(im.o_po)[i] = acc;
        im.at<double>(im.o_xyz)[3 * i] = sh[i]->x;
        im.at<double>(im.o_xyz)[3 * i + 1] = sh[i]->y;
        im.at<double>(im.o_xyz)[3 * i + 2] = sh[i]->z;
        if (sh[i]->nprim) {
            memcpy(im.at<double>(im.o_al) + acc, sh[i]->alpha, sizeof(double) * sh[i]->nprim);
            memcpy(im.at<double>(im.o_co) + acc, sh[i]->coef, sizeof(double) * sh[i]->nprim);
        }
        acc += sh[i]->nprim;
        im.maxam = std::max(im.maxam, sh[i]->am);
    }
}

OeiArgs shell_args(char *base, const ShellImage &im)
{
    OeiArgs a{};
    a.am = reinterpret_cast<const int *>(base + im.o_am);
    a.nprim = reinterpret_cast<const int *>(base + im.o_np);
    a.poff = reinterpret_cast<const int *>(base + im.o_po);
    a.xyz = reinterpret_cast<const double *>(base + im.o_xyz);
    a.alpha = reinterpret_cast<const double *>(base + im.o_al);
    a.coef = reinterpret_cast<const double *>(base + im.o_co);
    return a;
}

hipError_t launch_any(int kind, int maxam, const OeiArgs &a, hipStream_t s)
{
    return maxam <= 3 ? launch_osoei<3>(kind, a, s) : launch_osoei<OEI_MAXL>(kind, a, s);
}

// one pair through the calling thread's staging buffers (the drop-in calls)
int oei_one(int kind, const simint_shell *s1, const simint_shell *s2, int ncenter, const double *Z,
            const double *x, const double *y, const double *z, double *integrals)
{
    int v = check_shell(s1);
    if (!v) v = check_shell(s2);
    if (v) {
        simint_rt::set_error(v == SIMINT_ERROR_UNSUPPORTED ? "one-electron integrals: AM above 5"
                                                           : "one-electron integrals: bad shell");
        return v;
    }
    if (!integrals || ncenter < 0 || (ncenter > 0 && (!Z || !x || !y || !z))) return SIMINT_ERROR_ARGUMENT;
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    const simint_shell *sh[2] = {s1, s2};
    ShellImage im;
    pack_shells(im, 2, sh);
    const size_t o_pair = im.add(sizeof(int) * 2);
    const size_t o_off = im.add(sizeof(long long));
    const size_t o_ch = im.add(sizeof(double) * 4 * std::max(ncenter, 1));
    im.at<int>(o_pair)[0] = 0;
    im.at<int>(o_pair)[1] = 1;
    im.at<long long>(o_off)[0] = 0;
    for (int c = 0; c < ncenter; c++) {
        double *r = im.at<double>(o_ch) + 4 * c;
        r[0] = Z[c];
        r[1] = x[c];
        r[2] = y[c];
        r[3] = z[c];
    }
    const int nout = oei_nc(s1->am) * oei_nc(s2->am);
    void *host = nullptr, *din = nullptr, *dout = nullptr;
    // room for the block and its copy in the drop-in calls' Cartesian order
    hipError_t e = simint_rt::thread_buffers(im.buf.size(), im.buf.size(), 2 * sizeof(double) * nout, &host, &din,
                                             &dout);
    if (e != hipSuccess) { simint_rt::set_hip_error("one-electron buffers", e); return SIMINT_ERROR_HIP; }
    const hipStream_t s = simint_rt::stream();
    memcpy(host, im.buf.data(), im.buf.size());
    if ((e = hipMemcpyAsync(din, host, im.buf.size(), hipMemcpyHostToDevice, s)) != hipSuccess) {
        simint_rt::set_hip_error("hipMemcpyAsync(H2D)", e);
        return SIMINT_ERROR_HIP;
    }
    char *base = static_cast<char *>(din);
    OeiArgs a = shell_args(base, im);
    a.pairs = reinterpret_cast<const int *>(base + o_pair);
    a.off = reinterpret_cast<const long long *>(base + o_off);
    a.npair = 1;
    a.ncenter = kind == 2 ? ncenter : 0;
    a.charges = reinterpret_cast<const double *>(base + o_ch);
    a.out = static_cast<double *>(dout);
    a.boys = simint_rt::boys_grid();
    if ((e = launch_any(kind, im.maxam, a, s)) != hipSuccess) {
        simint_rt::set_hip_error("one-electron kernel launch", e);
        hipStreamSynchronize(s);
        return SIMINT_ERROR_HIP;
    }
    const double *result = static_cast<double *>(dout);
    bool moved = false;
    if ((e = simint_rt::to_drop_in_order(s1->am, s2->am, result, static_cast<double *>(dout) + nout, s, &moved)) !=
        hipSuccess) {
        simint_rt::set_hip_error("ordering permutation", e);
        hipStreamSynchronize(s);
        return SIMINT_ERROR_HIP;
    }
    if (moved) result += nout;
    if ((e = hipMemcpyAsync(integrals, result, sizeof(double) * nout, hipMemcpyDeviceToHost, s)) != hipSuccess) {
        simint_rt::set_hip_error("hipMemcpyAsync(D2H)", e);
        hipStreamSynchronize(s);
        return SIMINT_ERROR_HIP;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) {
        simint_rt::set_hip_error("hipStreamSynchronize", e);
        return SIMINT_ERROR_HIP;
    }
    return 1;
}

}  // namespace

extern "C" {

int simint_compute_osoei_overlap(const simint_shell *sh1, const simint_shell *sh2, double *restrict integrals)
{
    return oei_one(0, sh1, sh2, 0, nullptr, nullptr, nullptr, nullptr, integrals);
}

int simint_compute_osoei_ke(const simint_shell *sh1, const simint_shell *sh2, double *restrict integrals)
{
    return oei_one(1, sh1, sh2, 0, nullptr, nullptr, nullptr, nullptr, integrals);
}

int simint_compute_osoei_potential(int ncenter, double *Z, double *x, double *y, double *z, const simint_shell *sh1,
                                   const simint_shell *sh2, double *restrict integrals)
{
    return oei_one(2, sh1, sh2, ncenter, Z, x, y, z, integrals);
}

int simint_compute_overlap(simint_shell *a, simint_shell *b, double *restrict integrals)
{
    return simint_compute_osoei_overlap(a, b, integrals);
}

int simint_compute_ke(simint_shell *a, simint_shell *b, double *restrict integrals)
{
    return simint_compute_osoei_ke(a, b, integrals);
}

int simint_compute_potential(int ncenter, double *Z, double *x, double *y, double *z, const simint_shell *sh1,
                             const simint_shell *sh2, double *restrict integrals)
{
    return simint_compute_osoei_potential(ncenter, Z, x, y, z, sh1, sh2, integrals);
}

int simint_hip_shells_create(int nshell, const simint_shell *shells, simint_hip_shells **out)
{
    if (!out || nshell < 0 || (nshell > 0 && !shells)) return SIMINT_ERROR_ARGUMENT;
    std::vector<const simint_shell *> sh(nshell);
    for (int i = 0; i < nshell; i++) {
        sh[i] = shells + i;
        const int v = check_shell(sh[i]);
        if (v) {
            simint_rt::set_error(v == SIMINT_ERROR_UNSUPPORTED ? "one-electron integrals: AM above 5"
                                                               : "one-electron integrals: bad shell");
            return v;
        }
    }
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    ShellImage im;
    pack_shells(im, nshell, sh.data());
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, im.buf.size());
    if (e != hipSuccess) { simint_rt::set_hip_error("hipMalloc(shells)", e); return SIMINT_ERROR_HIP; }
    e = hipMemcpy(d, im.buf.data(), im.buf.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(d);
        simint_rt::set_hip_error("hipMemcpy(shells)", e);
        return SIMINT_ERROR_HIP;
    }
    simint_hip_shells *s = new simint_hip_shells;
    s->n = nshell;
    s->maxam = im.maxam;
    s->dev = d;
    const OeiArgs sa = shell_args(static_cast<char *>(d), im);
    s->host_am = new int[std::max(nshell, 1)];
    s->host_nprim = new int[std::max(nshell, 1)];
    for (int i = 0; i < nshell; i++) {
        s->host_am[i] = shells[i].am;
        s->host_nprim[i] = shells[i].nprim;
    }
    s->am = sa.am;
    s->nprim = sa.nprim;
    s->poff = sa.poff;
    s->xyz = sa.xyz;
    s->alpha = sa.alpha;
    s->coef = sa.coef;
    *out = s;
    return 0;
}

void simint_hip_shells_free(simint_hip_shells *s)
{
    if (!s) return;
    if (s->dev) hipFree(s->dev);
    delete[] s->host_am;
    delete[] s->host_nprim;
    delete s;
}

long long simint_hip_compute_osoei(int kind, const simint_hip_shells *S, long long npair, const int *d_pairs,
                                   const long long *d_off, int ncenter, const double *d_charges, double *d_out)
{
    if (!S || kind < 0 || kind > 2 || npair < 0 || ncenter < 0) return SIMINT_ERROR_ARGUMENT;
    if (npair > 0 && (!d_pairs || !d_off || !d_out)) return SIMINT_ERROR_ARGUMENT;
    if (kind == 2 && ncenter > 0 && !d_charges) return SIMINT_ERROR_ARGUMENT;
    if (!simint_rt::ensure_device()) return SIMINT_ERROR_NO_DEVICE;
    if (npair == 0) return 0;
    OeiArgs a{};
    a.am = S->am;
    a.nprim = S->nprim;
    a.poff = S->poff;
    a.xyz = S->xyz;
    a.alpha = S->alpha;
    a.coef = S->coef;
    a.pairs = d_pairs;
    a.off = d_off;
    a.npair = npair;
    a.ncenter = kind == 2 ? ncenter : 0;
    a.charges = d_charges;
    a.out = d_out;
    a.boys = simint_rt::boys_grid();
    const hipError_t e = launch_any(kind, S->maxam, a, simint_rt::stream());
    if (e != hipSuccess) { simint_rt::set_hip_error("one-electron kernel launch", e); return SIMINT_ERROR_HIP; }
    return npair;
}

}  // extern "C"

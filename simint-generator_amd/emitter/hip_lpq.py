"""HIP writer: lane-per-quartet (LPQ) kernels for one (ab|cd) class.

Mapping.  One wavefront (64 lanes) takes one work item: a bra shell pair,
uniform across the wave, against the 64 ket shell pairs of a ket tile, one
per lane.  Bra primitive data are wave-uniform scalar loads; ket primitive j
of every lane is one contiguous 512-byte run per field (tiled layout, see
csrc/device_types.h).  Each lane loops over bra x ket primitive pairs,
evaluates the Boys function, the Obara-Saika VRR as straight-line FP64 code
held in registers, and accumulates the contracted HRR inputs in registers;
then it runs the bra and ket HRR in registers and writes its ncart(abcd)
integrals in Psi4 order.

The arithmetic of every step is the reference's (setup:
generator/ostei/OSTEI_Writer.cpp:595-702; VRR: OSTEI_VRR_Writer.cpp:95-125;
HRR: OSTEI_HRR_Writer.cpp:22-99); only the (component, m) values that feed
the result are emitted.
"""
from .rrgraph import graph, carts, cart_index, ncart, I, J, K, L

AMCHAR = "spdfghik"


def class_name(am):
    return "_".join(AMCHAR[x] for x in am)


def _fmt_int(k):
    return "%d.0" % k


class _Writer:
    def __init__(self):
        self.lines = []

    def __call__(self, s=""):
        self.lines.append(s)

    def text(self):
        return "\n".join(self.lines) + "\n"


def emit_base(am, w, indent):
    """The primitive quartet's base values, into variables declared by the
    caller: Fp[m] = F_m(x) times the prefactor (m = 0..L), one_over_PQalpha_sum,
    alpha, PQx/PQy/PQz -- the same arithmetic emit_body uses."""
    Ltot = graph(*am).Ltot
    w(indent + "{")
    w(indent + "    const double PQalpha_mul = P_alpha * Q_alpha;")
    w(indent + "    const double PQalpha_sum = P_alpha + Q_alpha;")
    w(indent + "    one_over_PQalpha_sum = 1.0 / PQalpha_sum;")
    w(indent + "    PQx = P_x - Q_x;")
    w(indent + "    PQy = P_y - Q_y;")
    w(indent + "    PQz = P_z - Q_z;")
    w(indent + "    double R2 = PQx * PQx;")
    w(indent + "    R2 = fma(PQy, PQy, R2);")
    w(indent + "    R2 = fma(PQz, PQz, R2);")
    w(indent + "    alpha = PQalpha_mul * one_over_PQalpha_sum;")
    w(indent + "    const double F_x = R2 * alpha;")
    w(indent + "    double F[%d];" % (Ltot + 1))
    w(indent + "    boys_eval<%d>(grid, F_x, F);" % Ltot)
    w(indent + "    const double prefac = (P_prefac * Q_prefac) * sqrt(one_over_PQalpha_sum);")
    for m in range(Ltot + 1):
        w(indent + "    Fp[%d] = F[%d] * prefac;" % (m, m))
    w(indent + "}")


def emit_body(am, w, indent="        ", fp="double", keep=None, acc_tops=None, base="compute"):
    """Emit the per-primitive-quartet code (setup, Boys, VRR, accumulation).

    Assumes in scope: P_alpha, P_inv, P_prefac, P_x/y/z, P_PA or P_PB (x,y,z),
    Q_* likewise, grid, and accumulators acc<k> (double).  fp = "float":
    the Gaussian-product setup, the Boys function and the prefactor stay
    FP64; the VRR runs in FP32 on rounded copies of its inputs, and each
    result is accumulated into the FP64 accumulators.  base = "given": the
    values emit_base computes are already in scope (Fp[m], alpha,
    one_over_PQalpha_sum, PQx/PQy/PQz), so only the factors derived from
    them and the VRR are emitted."""
    given = base == "given"
    mixed = fp == "float"
    fm = "fmaf" if mixed else "fma"
    g = graph(*am)
    Ltot = g.Ltot
    # keep: the VRR nodes to emit (a closure of acc_tops); acc_tops: list of
    # (accumulator name, top key) -- by default every node and every top
    nodes = g.node_list if keep is None else [n for n in g.node_list if (n.q, n.m) in keep]
    if acc_tops is None:
        acc_tops = [("acc%d" % t, key) for t, key in enumerate(g.top_keys)]
    # The Boys values come from one evaluation at the class's total AM even
    # when a group of tops needs fewer orders, as in the reference (one
    # boys_F_split per primitive quartet): for x >= 36.5 the lower orders
    # then carry the same error pattern in every group, which the HRR
    # transfers between tops of different groups cancel -- with per-group
    # orders they did not, and (df|dd) lost 1e-7 normwise.
    used_steps = set(n.step for n in nodes if n.step is not None)
    # classes without HRR take 1/(p+q) and its square root from one refined
    # reciprocal square root (a few ulp): no HRR transfer amplifies it.  Only
    # in the lane-per-quartet kernels (their file carries the helpers); the
    # setup is FP64 in the mixed kernels too.
    fast = keep is None and not g.bra_hrr_needed() and not g.ket_hrr_needed() and not given
    if not given:
        _emit_setup(w, indent, Ltot, fast)
    _emit_vrr(g, nodes, acc_tops, w, indent, fp, mixed, fm, given, used_steps)


def _emit_setup(w, indent, Ltot, fast):
    w(indent + "const double PQalpha_mul = P_alpha * Q_alpha;")
    w(indent + "const double PQalpha_sum = P_alpha + Q_alpha;")
    if fast:
        w(indent + "const double rs_PQalpha_sum = rsqrt_nr(PQalpha_sum);")
        w(indent + "const double one_over_PQalpha_sum = rs_PQalpha_sum * rs_PQalpha_sum;")
    else:
        w(indent + "const double one_over_PQalpha_sum = 1.0 / PQalpha_sum;")
    w(indent + "const double PQx = P_x - Q_x;")
    w(indent + "const double PQy = P_y - Q_y;")
    w(indent + "const double PQz = P_z - Q_z;")
    w(indent + "double R2 = PQx * PQx;")
    w(indent + "R2 = fma(PQy, PQy, R2);")
    w(indent + "R2 = fma(PQz, PQz, R2);")
    w(indent + "const double alpha = PQalpha_mul * one_over_PQalpha_sum;")
    w(indent + "const double F_x = R2 * alpha;")
    w(indent + "double F[%d];" % (Ltot + 1))
    if fast:
        w(indent + "boys_eval_rs<%d>(grid, F_x, F);" % Ltot)
        w(indent + "const double prefac = (P_prefac * Q_prefac) * rs_PQalpha_sum;")
    else:
        w(indent + "boys_eval<%d>(grid, F_x, F);" % Ltot)
        w(indent + "const double prefac = (P_prefac * Q_prefac) * sqrt(one_over_PQalpha_sum);")


def _emit_vrr(g, nodes, acc_tops, w, indent, fp, mixed, fm, given, used_steps):
    has_bra = bool(used_steps & {I, J})
    has_ket = bool(used_steps & {K, L})
    # integer constants k/(2p), k/(2q), k/(2pq)
    need = {"2p": set(), "2q": set(), "2pq": set()}
    for n in nodes:
        if n.step is None:
            continue
        bra = n.step in (I, J)
        if n.src[2] is not None:
            need["2p" if bra else "2q"].add(n.fac[0])
        if n.src[4] is not None:
            need["2p" if bra else "2q"].add(n.fac[1])
        if n.src[6] is not None:
            need["2pq"].add(n.fac[2])
        if n.src[7] is not None:
            need["2pq"].add(n.fac[3])
    cv = (lambda e: "(float)(%s)" % e) if mixed else (lambda e: e)
    if need["2p"]:
        w(indent + "const %s one_over_2p = %s;" % (fp, cv("0.5 * P_inv")))
    if need["2q"]:
        w(indent + "const %s one_over_2q = %s;" % (fp, cv("0.5 * Q_inv")))
    if need["2pq"]:
        w(indent + "const %s one_over_2pq = %s;" % (fp, cv("0.5 * one_over_PQalpha_sum")))
    for kind in ("2p", "2q", "2pq"):
        for k in sorted(need[kind]):
            if k > 1:
                w(indent + "const %s c%d_%s = %s%s * one_over_%s;" % (fp, k, kind, _fmt_int(k), "f" if mixed else "",
                                                                     kind))
    if has_bra:
        w(indent + "const double a_over_p_d = (-alpha) * P_inv;")
        w(indent + "const %s a_over_p = %s;" % (fp, cv("a_over_p_d")))
        w(indent + "const %s aop_PQ[3] = {%s, %s, %s};" % (fp, cv("a_over_p_d * PQx"), cv("a_over_p_d * PQy"),
                                                           cv("a_over_p_d * PQz")))
    if has_ket:
        w(indent + "const double a_over_q_pos = alpha * Q_inv;")
        w(indent + "const %s aoq_PQ[3] = {%s, %s, %s};" % (fp, cv("a_over_q_pos * PQx"), cv("a_over_q_pos * PQy"),
                                                           cv("a_over_q_pos * PQz")))
        w(indent + "const %s a_over_q = %s;" % (fp, cv("-a_over_q_pos")))
    if mixed:
        # FP32 copies of the pair geometry the VRR steps use
        for side, dirs in (("P", (I, J)), ("Q", (K, L))):
            for arr in ("PA", "PB"):
                if any(n.step in dirs and ((arr == "PA") == (n.step in (I, K))) for n in nodes if n.step is not None):
                    w(indent + "const float f%s_%s[3] = {(float)%s_%s[0], (float)%s_%s[1], (float)%s_%s[2]};"
                      % (side, arr, side, arr, side, arr, side, arr))

    def cname(k, kind):
        return ("one_over_%s" % kind) if k == 1 else ("c%d_%s" % (k, kind))

    # each top is added to its accumulator right after it is computed, so
    # the value is dead unless a later node reads it (the node list is a
    # depth-first post-order over the tops)
    acc_of = {}
    for accname, key in acc_tops:
        acc_of.setdefault(key, []).append(accname)

    def accumulate(key, v):
        for accname in acc_of.get(key, ()):
            w(indent + "%s += %s;" % (accname, ("(double)%s" % v) if mixed else v))

    idx = {}
    for n in nodes:
        key = (n.q, n.m)
        v = "v%d" % len(idx)
        idx[key] = v
        if n.step is None:
            w(indent + "const %s %s = %s;" % (fp, v, cv(("Fp[%d]" if given else "F[%d] * prefac") % n.m)))
            accumulate(key, v)
            continue
        d = n.dir
        bra = n.step in (I, J)
        pfx = "f" if mixed else ""
        if n.step == I:
            pa = "%sP_PA[%d]" % (pfx, d)
        elif n.step == J:
            pa = "%sP_PB[%d]" % (pfx, d)
        elif n.step == K:
            pa = "%sQ_PA[%d]" % (pfx, d)
        else:
            pa = "%sQ_PB[%d]" % (pfx, d)
        aop = ("aop_PQ[%d]" if bra else "aoq_PQ[%d]") % d
        aover = "a_over_p" if bra else "a_over_q"
        k2 = "2p" if bra else "2q"
        s = [idx[x] if x is not None else None for x in n.src]
        expr = "%s * %s" % (pa, s[0])
        expr = "%s(%s, %s, %s)" % (fm, aop, s[1], expr)
        if s[2] is not None:
            expr = "%s(%s, %s(%s, %s, %s), %s)" % (fm, cname(n.fac[0], k2), fm, aover, s[3], s[2], expr)
        if s[4] is not None:
            expr = "%s(%s, %s(%s, %s, %s), %s)" % (fm, cname(n.fac[1], k2), fm, aover, s[5], s[4], expr)
        if s[6] is not None:
            expr = "%s(%s, %s, %s)" % (fm, cname(n.fac[2], "2pq"), s[6], expr)
        if s[7] is not None:
            expr = "%s(%s, %s, %s)" % (fm, cname(n.fac[3], "2pq"), s[7], expr)
        w(indent + "const %s %s = %s;" % (fp, v, expr))
        accumulate(key, v)


STAGE = 16   # output components per LDS slice (coalesced stores)


def _emit_flush(w, indent, k0, sl):
    """Write components [k0, k0 + sl) of the wave's quartets from the LDS
    slice: consecutive threads store consecutive doubles of one quartet's
    block, so every store instruction writes sl-double runs instead of 64
    scattered doubles."""
    w(indent + "__syncthreads();")
    w(indent + "for (int e = lane; e < %d; e += 64) {" % (64 * sl))
    w(indent + "    const int q = e / %d, j = e - q * %d;" % (sl, sl))
    w(indent + "    if (q >= wi.lane_lo && q < wi.lane_hi) A.out[ob[q] + %d + j] = xb[q * %d + j];" % (k0, STAGE + 1))
    w(indent + "}")
    w(indent + "__syncthreads();")


def emit_hrr(am, w, indent="    ", out="o", acc="acc", fp="double", staged=False):
    """Emit the bra then ket HRR on acc<k>, storing into out[k] (Psi4 order).
    fp = "float": the HRR runs in FP32 on rounded copies of the FP64
    contracted values and AB/CD; results are widened to FP64 on store.
    staged: results go through the LDS slices xb / ob (see _emit_flush)
    instead of straight to out[k]; every lane of the wave must run it."""
    mixed = fp == "float"
    fm = "fmaf" if mixed else "fma"
    g = graph(*am)
    la, lb, lc, ld = am
    bra_ops, bra_map, ket_ops, ket_map = g.hrr_program()
    ket_tops = g.ket_tops()
    bra_tops = g.bra_tops()
    nf = len(ket_tops)
    etop = {e: i for i, e in enumerate(bra_tops)}
    ftop = {f: i for i, f in enumerate(ket_tops)}

    if mixed:
        for t in range(len(g.tops)):
            w(indent + "const float f%s%d = (float)%s%d;" % (acc, t, acc, t))
        if g.bra_hrr_needed():
            w(indent + "const float fhAB[3] = {(float)hAB[0], (float)hAB[1], (float)hAB[2]};")
        if g.ket_hrr_needed():
            w(indent + "const float fhCD[3] = {(float)hCD[0], (float)hCD[1], (float)hCD[2]};")
    pre = "f" if mixed else ""

    def topvar(e, f):
        return "%s%s%d" % (pre, acc, etop[e] * nf + ftop[f])

    # bra HRR for every top ket component f -> names hb_<f>_<a>_<b>
    names = {}
    cnt = [0]

    def newname(prefix):
        cnt[0] += 1
        return "%s%d" % (prefix, cnt[0])

    def bra_val(src, f):
        if src[0] == "top":
            return topvar(src[1], f)
        return names[("b", f, src[1], src[2])]

    for f in ket_tops:
        for (dst, s0, s1, sign, d) in bra_ops:
            nm = newname("hb")
            names[("b", f, dst[1], dst[2])] = nm
            ab = ("%shAB[%d]" % (pre, d)) if sign > 0 else ("(-%shAB[%d])" % (pre, d))
            w(indent + "const %s %s = %s(%s, %s, %s);" % (fp, nm, fm, ab, bra_val(s1, f), bra_val(s0, f)))

    # ket HRR for every final bra component
    nb, nc, nd = ncart(lb), ncart(lc), ncart(ld)
    ntot = ncart(la) * nb * nc * nd
    for ia, a in enumerate(carts(la)):
        for ib, b in enumerate(carts(lb)):
            bm = bra_map[(a, b)]

            def abf(fc):
                return bra_val(bm, fc)

            def ket_val(src):
                if src[0] == "top":
                    return abf(src[1])
                return names[("k", a, b, src[1], src[2])]

            for (dst, s0, s1, sign, d) in ket_ops:
                nm = newname("hk")
                names[("k", a, b, dst[1], dst[2])] = nm
                cd = ("%shCD[%d]" % (pre, d)) if sign > 0 else ("(-%shCD[%d])" % (pre, d))
                w(indent + "const %s %s = %s(%s, %s, %s);" % (fp, nm, fm, cd, ket_val(s1), ket_val(s0)))
            for ic, c in enumerate(carts(lc)):
                for idd, dd in enumerate(carts(ld)):
                    k = ((ia * nb + ib) * nc + ic) * nd + idd
                    val = ket_val(ket_map[(c, dd)])
                    if mixed:
                        val = "(double)" + val
                    if staged:
                        w(indent + "xb[lane * %d + %d] = %s;" % (STAGE + 1, k % STAGE, val))
                        if (k + 1) % STAGE == 0 or k + 1 == ntot:
                            _emit_flush(w, indent, k - k % STAGE, k % STAGE + 1)
                    else:
                        w(indent + "%s[%d] = %s;" % (out, k, val))


def _emit_fast_setup(w, Ltot):
    """Helpers of the classes without HRR (emit_body's fast setup): x^-1/2
    from the hardware reciprocal square root refined by two Newton steps
    y <- y + y (1/2 - (x/2) y^2) (within a few ulp of the correctly rounded
    value; IEEE 1/x and sqrt are each about three times the instructions),
    and the Boys function with its long-range branch for L < 4 (all orders
    from longfac[m] x^-(m+1/2), boys.h:18-25) on that square root (L >= 4
    keeps boys_eval's branch)."""
    w("namespace {")
    w("__device__ __forceinline__ double rsqrt_nr(double x)")
    w("{")
    w("    double y = __builtin_amdgcn_rsq(x);")
    w("    const double h = 0.5 * x;")
    w("    double t = fma(-h * y, y, 0.5);")
    w("    y = fma(y, t, y);")
    w("    t = fma(-h * y, y, 0.5);")
    w("    return fma(y, t, y);")
    w("}")
    w("")
    w("template <int L>")
    w("__device__ __forceinline__ void boys_eval_rs(const double *__restrict__ grid, double x, double *F)")
    w("{")
    w("    if (L >= 4 || x < SIMINT_BOYS_MAXX) {")
    w("        boys_eval<L>(grid, x, F);")
    w("        return;")
    w("    }")
    w("    constexpr double longfac[] = SIMINT_BOYS_LONGFAC;")
    w("    double x2 = rsqrt_nr(x);")
    w("    const double x1 = x2 * x2;")
    w("#pragma unroll")
    w("    for (int n = 0; n <= L; n++) {")
    w("        F[n] = longfac[n] * x2;")
    w("        x2 *= x1;")
    w("    }")
    w("}")
    w("}  // namespace")
    w("")


def emit_lpq(am, mixed=False):
    """Source of one class: the prefix kernel, the quartet-list kernel and,
    with mixed=True, the FP32-VRR prefix kernel (eri_lpqm_<class>)."""
    w = _Writer()
    w("// Generated by simint-generator_amd/emitter/hip_lpq.py -- class (%s)" % class_name(am).replace("_", ""))
    w('#include "../device_types.h"')
    w('#include "../boys_device.h"')
    w("")
    g = graph(*am)
    if not g.bra_hrr_needed() and not g.ket_hrr_needed():
        _emit_fast_setup(w, g.Ltot)
    _emit_lpq_prefix(am, w, "double")
    emit_lpq_list(am, w)
    if mixed:
        w("")
        _emit_lpq_prefix(am, w, "float")
    return w.text()


def _emit_lpq_prefix(am, w, fp, hrr_fp="double"):
    g = graph(*am)
    name = class_name(am)
    ntop = len(g.top_keys)
    kname = ("eri_lpqm_%s" if fp == "float" else "eri_lpq_%s") % name
    w("__global__ void __launch_bounds__(64) %s(const EriArgs A)" % kname)
    w("{")
    w("    const WorkItem wi = A.items[blockIdx.x];")
    w("    const int lane = threadIdx.x;")
    w("    const double *__restrict__ grid = A.boys_grid;")
    w("    const int ib0 = A.bra.poff[wi.bra];")
    w("    const int ib1 = ib0 + A.bra.np[wi.bra];")
    w("    const int nq = A.ket.tnp[wi.tile];")
    w("    const long long tl = (long long)wi.tile * SIMINT_TILE + lane;")
    w("    const long long kb = A.ket.toff[wi.tile] + lane;")
    w("    const bool screen = A.tol2 > 0.0;")
    for t in range(ntop):
        w("    double acc%d = 0.0;" % t)
    bra_pa = "pa" if g.bra_j else "pb"
    ket_pa = "pa" if g.ket_l else "pb"
    w("    for (int i = ib0; i < ib1; ++i) {")
    w("        const double P_alpha = A.bra.alpha[i];")
    w("        const double P_inv = A.bra.inv_alpha[i];")
    w("        const double P_prefac = A.bra.prefac[i];")
    w("        const double P_x = A.bra.x[i], P_y = A.bra.y[i], P_z = A.bra.z[i];")
    w("        const double P_%s[3] = {A.bra.%sx[i], A.bra.%sy[i], A.bra.%sz[i]};"
      % ("PA" if g.bra_j else "PB", bra_pa, bra_pa, bra_pa))
    w("        const double P_screen = screen ? A.bra.screen[i] : 0.0;")
    w("        if (screen && P_screen * A.ket_screen_max < A.tol2) continue;")
    w("        for (int j = 0; j < nq; ++j) {")
    w("            const long long kj = kb + (long long)SIMINT_TILE * j;")
    w("            if (screen && A.ket.screen[kj] * P_screen < A.tol2) continue;")
    w("            const double Q_alpha = A.ket.alpha[kj];")
    w("            const double Q_inv = A.ket.inv_alpha[kj];")
    w("            const double Q_prefac = A.ket.prefac[kj];")
    w("            const double Q_x = A.ket.x[kj], Q_y = A.ket.y[kj], Q_z = A.ket.z[kj];")
    w("            const double Q_%s[3] = {A.ket.%sx[kj], A.ket.%sy[kj], A.ket.%sz[kj]};"
      % ("PA" if g.ket_l else "PB", ket_pa, ket_pa, ket_pa))
    emit_body(am, w, indent="            ", fp=fp)
    w("        }")
    w("    }")
    # direct stores (through LDS slices measured 3% slower on these
    # compute-bound kernels)
    w("    if (lane < wi.lane_lo || lane >= wi.lane_hi) return;")
    w("    const long long colv = A.ket.col[tl];")
    w("    double *__restrict__ o = A.out + wi.obase + colv * A.stride;")
    if g.bra_hrr_needed():
        w("    const double hAB[3] = {A.bra.abx[wi.bra], A.bra.aby[wi.bra], A.bra.abz[wi.bra]};")
    if g.ket_hrr_needed():
        w("    const double hCD[3] = {A.ket.abx[tl], A.ket.aby[tl], A.ket.abz[tl]};")
    # the HRR stays FP64: its AB / CD transfers cancel strongly for the
    # far-apart pairs that have small Schwarz bounds, and FP32 there loses
    # far more than the ERI error budget
    emit_hrr(am, w, fp=hrr_fp)
    w("}")
    w("")
    w('extern "C" hipError_t simint_launch_%s(const EriArgs *a, hipStream_t s)' % kname[4:])
    w("{")
    w("    if (a->nitems <= 0) return hipSuccess;")
    w("    hipLaunchKernelGGL(%s, dim3((unsigned)a->nitems), dim3(64), 0, s, *a);" % kname)
    w("    return hipGetLastError();")
    w("}")
    w("")


def emit_lpq_list(am, w):
    """Quartet-list variant: S = A.split lanes per quartet (a power of two
    up to 16); lanes 64k + l take quartet (64k + l) / S, with its own bra and
    ket pair (both read from the linear view, per-lane gathers).  Lane
    sub = l % S sums primitive quartets sub, sub + S, ... (bra-major order,
    as the reference's loops); the S partial sums are added across the
    lanes (xor butterfly), then lane sub = 0 runs the HRR and writes the
    contiguous output block.  With S = 1 the summation order is the
    lane-per-quartet kernels'."""
    g = graph(*am)
    name = class_name(am)
    ntop = len(g.top_keys)
    w("__global__ void __launch_bounds__(64) eri_lpql_%s(const EriListArgs A)" % name)
    w("{")
    w("    const int S = A.split;")
    w("    const long long gl = (long long)blockIdx.x * 64 + threadIdx.x;")
    w("    const long long q = gl / S;")
    w("    const int sub = (int)(gl - q * S);")
    w("    const bool valid = q < A.nquartet;")
    w("    // lanes past the list stay until the cross-lane sums")
    w("    const long long qc = valid ? q : A.nquartet - 1;")
    w("    const int bp = A.qbra[qc], kp = A.qket[qc];")
    w("    const double *__restrict__ grid = A.boys_grid;")
    w("    const int ib0 = A.bra.poff[bp];")
    w("    const int jb0 = A.ket.poff[kp];")
    w("    const int nqk = A.ket.np[kp];")
    w("    const int K = valid ? A.bra.np[bp] * nqk : 0;")
    w("    const bool screen = A.tol2 > 0.0;")
    for t in range(ntop):
        w("    double acc%d = 0.0;" % t)
    bra_pa = "pa" if g.bra_j else "pb"
    ket_pa = "pa" if g.ket_l else "pb"
    w("    int ii = 0, jj = sub;")
    w("    for (int p = sub; p < K; p += S) {")
    w("        while (jj >= nqk) { jj -= nqk; ++ii; }")
    w("        const int i = ib0 + ii, kj = jb0 + jj;")
    w("        jj += S;")
    w("        const double P_screen = screen ? A.bra.screen[i] : 0.0;")
    w("        if (screen && (P_screen * A.ket_screen_max < A.tol2 || A.ket.screen[kj] * P_screen < A.tol2)) continue;")
    w("        const double P_alpha = A.bra.alpha[i];")
    w("        const double P_inv = A.bra.inv_alpha[i];")
    w("        const double P_prefac = A.bra.prefac[i];")
    w("        const double P_x = A.bra.x[i], P_y = A.bra.y[i], P_z = A.bra.z[i];")
    w("        const double P_%s[3] = {A.bra.%sx[i], A.bra.%sy[i], A.bra.%sz[i]};"
      % ("PA" if g.bra_j else "PB", bra_pa, bra_pa, bra_pa))
    w("        {")
    w("            const double Q_alpha = A.ket.alpha[kj];")
    w("            const double Q_inv = A.ket.inv_alpha[kj];")
    w("            const double Q_prefac = A.ket.prefac[kj];")
    w("            const double Q_x = A.ket.x[kj], Q_y = A.ket.y[kj], Q_z = A.ket.z[kj];")
    w("            const double Q_%s[3] = {A.ket.%sx[kj], A.ket.%sy[kj], A.ket.%sz[kj]};"
      % ("PA" if g.ket_l else "PB", ket_pa, ket_pa, ket_pa))
    emit_body(am, w, indent="            ")
    w("        }")
    w("    }")
    w("    for (int o = 1; o < S; o <<= 1) {")
    for t in range(ntop):
        w("        acc%d += __shfl_xor(acc%d, o);" % (t, t))
    w("    }")
    w("    if (!valid || sub != 0) return;")
    w("    double *__restrict__ o = A.out + (A.qout ? A.qout[q] : q * A.ostride);")
    if g.bra_hrr_needed():
        w("    const double hAB[3] = {A.bra.abx[bp], A.bra.aby[bp], A.bra.abz[bp]};")
    if g.ket_hrr_needed():
        w("    const double hCD[3] = {A.ket.abx[kp], A.ket.aby[kp], A.ket.abz[kp]};")
    emit_hrr(am, w)
    w("}")
    w("")
    w('extern "C" hipError_t simint_launch_lpql_%s(const EriListArgs *a, hipStream_t s)' % name)
    w("{")
    w("    if (a->nquartet <= 0) return hipSuccess;")
    w("    if (a->split < 1 || a->split > 16 || (a->split & (a->split - 1))) return hipErrorInvalidValue;")
    w("    hipLaunchKernelGGL(eri_lpql_%s, dim3((unsigned)((a->nquartet * a->split + 63) / 64)), dim3(64), 0, s, *a);"
      % name)
    w("    return hipGetLastError();")
    w("}")

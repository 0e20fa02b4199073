"""HIP writer: wave-per-quartet (WPQ) kernels for one (ab|cd) class.

For classes whose recurrence does not fit in one lane's registers.  One
wavefront takes one work item (a bra shell pair against a ket tile, exactly
as the lane-per-quartet kernels) and walks the tile's quartets one at a
time, cooperating on each:

1. primitive setup in parallel: lane l prepares primitive quartet l of the
   current batch of 64 (Gaussian-product quantities, Boys F_m scaled by the
   prefactor) into LDS;
2. VRR in LDS: the class's recurrence nodes (one per needed (component,
   m)) are grouped into dependency levels; the lanes split each level, read
   the sources from LDS and write the result back (node descriptors are a
   per-class constant table).  A thread loads the descriptors and then the
   sources of all its nodes of a level before it stores any result, with
   absent sources read from zero slots, so the level costs about one
   load latency rather than one per node;
3. contraction: lane l accumulates HRR inputs t = l, l+64, ... in registers;
4. HRR with uniform straight-line code: bra HRR with one lane per ket-top
   component, then ket HRR with one lane per final bra component, staged
   through LDS; results go straight to the output.

Arithmetic per step is the same as the lane-per-quartet kernels (and the
reference: generator/ostei/OSTEI_VRR_Writer.cpp:95-125,
OSTEI_HRR_Writer.cpp:22-99).
"""
from .rrgraph import graph, carts, ncart, I, J, K, L
from .hip_lpq import class_name, _Writer

# constants per primitive quartet kept in LDS (before F_0..F_L), in the
# blocks a recurrence step reads together: for direction s (0..2 bra x, y,
# z; 3..5 ket x, y, z) doubles 4s..4s+3 hold PA (or QC), aop_PQ (aoq_PQ),
# a_over_p (-a_over_q) and 1/2p (1/2q) of that side; 24 holds 1/2pq
NPC_FIXED = 25


def byte_desc(desc, starts, B, LPB):
    """Node descriptors as eight 32-bit words for a batch of B quartets:
    LDS byte offsets (from slot 0 of quartet slot 0) of the s0, s2, s4 pairs,
    of s6 and s7, of the node's own slot, and (from the primitive-constant
    block) of the step direction's four constants; the four integer factors
    as bytes.  Rows go level by level, each level padded to a multiple of
    LPB rows (zero sources, written to the sink slot) so that every thread
    of a level reads a row of that level.  Returns the rows and the first
    row of each level (level 0, the Boys bases, has none)."""
    zero, sink = len(desc), len(desc) + 2
    z = zero * B * 8
    pad = (z, z, z, z, z, sink * B * 8, 0, 0)
    rows, first = [], [0]
    for lv in range(1, len(starts) - 1):
        first.append(len(rows))
        S, E = starts[lv], starts[lv + 1]
        for k in range(S, E):
            w0, w1 = desc[k]
            s0, s2, s4, s6 = w0 & 0xffff, (w0 >> 16) & 0xffff, (w0 >> 32) & 0xffff, w0 >> 48
            s7 = w1 & 0xffff
            fl = (w1 >> 16) & 0xff
            sel = (fl & 1) * 3 + ((fl >> 1) & 3)
            fac = [(w1 >> sh) & 0xff for sh in (24, 32, 40, 48)]
            rows.append((s0 * B * 8, s2 * B * 8, s4 * B * 8, s6 * B * 8, s7 * B * 8, k * B * 8,
                         4 * sel * B * 8, fac[0] | fac[1] << 8 | fac[2] << 16 | fac[3] << 24))
        rows += [pad] * ((-(E - S)) % LPB)
    return rows, first


def levels(g):
    """Longest-path level of every VRR node (bases at level 0)."""
    lev = {}
    for n in g.node_list:
        key = (n.q, n.m)
        if n.step is None:
            lev[key] = 0
        else:
            lev[key] = 1 + max(lev[s] for s in n.src if s is not None)
    return lev


def plan(am):
    """Slot layout and node descriptors of the LDS-staged VRR.

    Slots are ordered by dependency level; within a level each component
    quartet owns a run of consecutive slots, one per m in its (contiguous)
    m range, so the (m, m+1) source pairs of a step sit in adjacent slots and
    one index addresses both.  Bases (ss|ss)^m occupy slots 0..L.
    Descriptor (two 64-bit words):
      w0 = s0 | s2 << 16 | s4 << 32 | s6 << 48
      w1 = s7 | flags << 16 | f0 << 24 | f1 << 32 | f2 << 40 | f3 << 48
      flags: bit0 ket step, bits1-2 direction, bit3..6 s2/s4/s6/s7 present.
    An absent source points at the zero slots (len(order), len(order) + 1,
    both 0.0 during the recurrence) with factor 0, so every node evaluates
    the same expression without branches."""
    g = graph(*am)
    lev = levels(g)
    comps = {}
    for n in g.node_list:
        comps.setdefault(n.q, []).append(n)
    bases = sorted(comps.pop(((0, 0, 0),) * 4), key=lambda n: n.m)
    order = list(bases)
    by_level = {}
    for q, ns in comps.items():
        by_level.setdefault(lev[(q, ns[0].m)], []).append(q)
    for lv in sorted(by_level):
        for q in sorted(by_level[lv]):
            ms = sorted(comps[q], key=lambda n: n.m)
            assert [n.m for n in ms] == list(range(ms[0].m, ms[-1].m + 1))
            order.extend(ms)
    slot = {(n.q, n.m): i for i, n in enumerate(order)}
    assert len(order) < 65536
    starts = [0]
    cur = 0
    for i, n in enumerate(order):
        lv = lev[(n.q, n.m)]
        while cur < lv:
            starts.append(i)
            cur += 1
    starts.append(len(order))
    zero = len(order)
    assert zero + 2 < 65536
    desc = []
    for n in order:
        if n.step is None:
            desc.append((0, 0))
            continue
        sl = [slot[x] if x is not None else zero for x in n.src]
        for a, b in ((0, 1), (2, 3), (4, 5)):
            if n.src[a] is not None:
                assert sl[b] == sl[a] + 1
        has = [n.src[2] is not None, n.src[4] is not None, n.src[6] is not None, n.src[7] is not None]
        flags = (1 if n.step in (K, L) else 0) | (n.dir << 1) | (has[0] << 3) | (has[1] << 4) \
            | (has[2] << 5) | (has[3] << 6)
        f = [x if h else 0 for x, h in zip(n.fac, has)]
        assert 0 <= min(f) and max(f) < 256
        w0 = sl[0] | (sl[2] << 16) | (sl[4] << 32) | (sl[6] << 48)
        w1 = sl[7] | (flags << 16) | (f[0] << 24) | (f[1] << 32) | (f[2] << 40) | (f[3] << 48)
        desc.append((w0, w1))
    tops = [slot[k] for k in g.top_keys]
    return g, order, starts, desc, tops


LDS_BUDGET = 56 * 1024
LDS_MAX = 160 * 1024 - 1024   # gfx950: 160 KB of LDS per CU, usable by one workgroup
THREADS = 256


def batch_size(am):
    """Quartets per batch: the largest power of two <= 16 whose LDS image
    (VRR/HRR work area per quartet plus the primitive-constant batch) fits the
    per-workgroup budget (so 2-4 workgroups share a CU)."""
    g, order, starts, desc, tops = plan(am)
    la, lb, lc, ld = am
    ws = max(len(order) + 3, len(tops) + ncart(la) * ncart(lb) * len(g.ket_tops()))
    npc = NPC_FIXED + g.Ltot + 1
    B = 16
    while B > 1:
        if 8 * (ws * B + PB_SETUP * npc * B) <= LDS_BUDGET:
            break
        B //= 2
    if 8 * (ws * B + PB_SETUP * npc * B) > LDS_MAX:
        raise ValueError("class %s needs %d bytes of LDS" % (am, 8 * (ws + PB_SETUP * npc)))
    # a quartet taking more than 64 KB leaves room for at most one workgroup
    # per CU alongside its register use anyway: take two quartets per batch
    # where they fit (twice the independent nodes per level barrier)
    if B == 1 and 8 * ws > 64 * 1024 and 8 * (2 * ws + 2 * PB_SETUP * npc) <= LDS_MAX:
        B = 2
    return B, ws, npc


PB_SETUP = 4   # primitive quartets per setup batch (sizing B)


LDS_CU = 160 * 1024


def setup_batch(B, ws, npc):
    """Primitive quartets prepared per setup batch: up to 32 (one per thread
    row, so the Boys evaluations of a batch run side by side), as long as
    the larger LDS image keeps as many workgroups per CU as PB_SETUP does."""
    base = 8 * (ws * B + PB_SETUP * npc * B)
    for PB in (32, 16, 8):
        lds = 8 * (ws * B + PB * npc * B)
        if PB <= THREADS // B and lds <= LDS_MAX and LDS_CU // lds == LDS_CU // base:
            return PB
    return PB_SETUP


LEVEL_CHUNK = 4   # nodes per thread whose loads are issued together
LEVEL_CHUNK_ONE = 8   # the same for classes with one quartet per batch:
                      # they run at one wave per SIMD, with the register
                      # room for twice the loads in flight


def emit_wpq(am):
    g, order, starts, desc, tops = plan(am)
    la, lb, lc, ld = am
    name = class_name(am)
    Ltot = g.Ltot
    nnode = len(order)
    ntop = len(tops)
    B, ws, npc = batch_size(am)
    LPB = THREADS // B
    nacc = (ntop + LPB - 1) // LPB
    bra_tops = g.bra_tops()
    ket_tops = g.ket_tops()
    ne, nf = len(bra_tops), len(ket_tops)
    nab = ncart(la) * ncart(lb)
    ncd = ncart(lc) * ncart(ld)
    w = _Writer()
    w("// Generated by simint-generator_amd/emitter/hip_wpq.py -- class (%s)" % name.replace("_", ""))
    w('#include "../device_types.h"')
    w('#include "../boys_device.h"')
    w("")
    w("namespace {")
    w("constexpr int NNODE = %d, NTOP = %d, NPC = %d, NE = %d, NF = %d, NAB = %d, NCD = %d;"
      % (nnode, ntop, npc, ne, nf, nab, ncd))
    w("constexpr int B = %d, LPB = %d, NACC = %d, PB = %d, WS = %d, ZS = %d;"
      % (B, LPB, nacc, setup_batch(B, ws, npc), ws, nnode))
    w("constexpr int NLEV = %d;" % (len(starts) - 1))
    w("__constant__ int lev_start[%d] = {%s};" % (len(starts), ", ".join(map(str, starts))))
    rows, _ = byte_desc(desc, starts, B, LPB)
    w("__constant__ __attribute__((aligned(16))) unsigned int vrr_desc[%d][8] = {" % max(1, len(rows)))
    for d in rows:
        w("    {%s}," % ", ".join("%du" % x for x in d))
    w("};")
    w("__constant__ unsigned short top_slot[%d] = {%s};" % (ntop, ", ".join(map(str, tops))))
    w("}  // namespace")
    w("")
    _emit_kernel(g, w, am, name, starts, Ltot, B, LPB, listmode=False)
    _emit_kernel(g, w, am, name, starts, Ltot, B, LPB, listmode=True)
    return w.text()


# quartets per workgroup in the quartet-list kernels: one batch of B, so
# that a class with few quartets still spreads over many workgroups


def _emit_kernel(g, w, am, name, starts, Ltot, B, LPB, listmode):
    """The cooperative kernel: over the ket lanes of a work item (one bra), or
    (listmode) over a chunk of a quartet list, each quartet with its own bra
    and ket pair read from the linear view."""
    w("// %d threads cooperate on B = %d quartets at a time; thread t serves quartet" % (THREADS, B))
    w("// slot t %% B and is rank t / B of that quartet's LPB = %d threads." % LPB)
    if listmode:
        w("__global__ void __launch_bounds__(%d) eri_wpql_%s(const EriListArgs A)" % (THREADS, name))
    else:
        w("__global__ void __launch_bounds__(%d) eri_wpq_%s(const EriArgs A)" % (THREADS, name))
    w("{")
    # one LDS array (at LDS address 0, so descriptor offsets address it
    # directly): VRR / HRR work area, primitive constants, list depths
    w("    __shared__ double lds[WS * B + PB * NPC * B + (B + 1) / 2];")
    w("    double *const work = lds;")
    w("    double *const pc = lds + WS * B;")
    bra_pa = "pa" if g.bra_j else "pb"
    ket_pa = "pa" if g.ket_l else "pb"
    if not listmode:
        w("    const WorkItem wi = A.items[blockIdx.x];")
    w("    const int tid = threadIdx.x;")
    w("    const int b = tid % B;")
    w("    const int r0 = tid / B;")
    w("    const double *__restrict__ grid = A.boys_grid;")
    w("    const bool screen = A.tol2 > 0.0;")
    if not listmode:
        w("    const int ib0 = A.bra.poff[wi.bra];")
        w("    const int npb = A.bra.np[wi.bra];")
        w("    const int nq = A.ket.tnp[wi.tile];")
        w("    const long long kt = A.ket.toff[wi.tile];")
        w("    const int K = npb * nq;")
        if g.bra_hrr_needed():
            w("    const double hAB[3] = {A.bra.abx[wi.bra], A.bra.aby[wi.bra], A.bra.abz[wi.bra]};")
        w("    for (int q0 = wi.lane_lo; q0 < wi.lane_hi; q0 += B) {")
        w("        const int nb = (wi.lane_hi - q0) < B ? (wi.lane_hi - q0) : B;")
        w("        const int ql = (q0 + b) < SIMINT_TILE ? (q0 + b) : (SIMINT_TILE - 1);")
        w("        const long long tl = (long long)wi.tile * SIMINT_TILE + ql;")
    else:
        w("    int *const kk = (int *)(lds + WS * B + PB * NPC * B);")
        w("    const long long qbeg = (long long)blockIdx.x * B;")
        w("    const long long qend = (qbeg + B) < A.nquartet ? (qbeg + B) : A.nquartet;")
        w("    for (long long q0 = qbeg; q0 < qend; q0 += B) {")
        w("        const int nb = (qend - q0) < B ? (int)(qend - q0) : B;")
        w("        const long long qi = (q0 + b) < qend ? (q0 + b) : (qend - 1);")
        w("        const int bp = A.qbra[qi], kp = A.qket[qi];")
        w("        const int ib0 = A.bra.poff[bp];")
        w("        const int nq = A.ket.np[kp];")
        w("        const int jb0 = A.ket.poff[kp];")
        w("        const int Kb = (b < nb) ? A.bra.np[bp] * nq : 0;")
        w("        __syncthreads();")
        w("        if (r0 == 0) kk[b] = Kb;")
        w("        __syncthreads();")
        w("        int K = 0;")
        w("#pragma unroll")
        w("        for (int t = 0; t < B; ++t) K = kk[t] > K ? kk[t] : K;")
        if g.bra_hrr_needed():
            w("        const double hAB[3] = {A.bra.abx[bp], A.bra.aby[bp], A.bra.abz[bp]};")
    w("        double acc[NACC];")
    w("#pragma unroll")
    w("        for (int s = 0; s < NACC; ++s) acc[s] = 0.0;")
    w("        for (int p0 = 0; p0 < K; p0 += PB) {")
    w("            const int np_batch = (K - p0) < PB ? (K - p0) : PB;")
    w("            __syncthreads();")
    w("            if (r0 < np_batch) {")
    if not listmode:
        w("                const int p = p0 + r0;")
        w("                const int i = ib0 + p / nq;")
        w("                const long long kj = kt + ql + (long long)SIMINT_TILE * (p % nq);")
    else:
        # quartets of the batch with fewer primitive quartets than the
        # largest contribute zero for the surplus rows
        w("                const int p = p0 + r0;")
        w("                const bool live = p < Kb;")
        w("                const int pe = live ? p : 0;")
        w("                const int i = ib0 + pe / nq;")
        w("                const long long kj = jb0 + pe % nq;")
    w("                const double P_alpha = A.bra.alpha[i], P_inv = A.bra.inv_alpha[i];")
    w("                const double Q_alpha = A.ket.alpha[kj], Q_inv = A.ket.inv_alpha[kj];")
    if not listmode:
        w("                double P_prefac = A.bra.prefac[i];")
    else:
        w("                double P_prefac = live ? A.bra.prefac[i] : 0.0;")
    w("                if (screen && (A.bra.screen[i] * A.ket_screen_max < A.tol2 ||")
    w("                               A.bra.screen[i] * A.ket.screen[kj] < A.tol2)) P_prefac = 0.0;")
    w("                const double Q_prefac = A.ket.prefac[kj];")
    w("                const double PQalpha_mul = P_alpha * Q_alpha;")
    w("                const double PQalpha_sum = P_alpha + Q_alpha;")
    w("                const double one_over_PQalpha_sum = 1.0 / PQalpha_sum;")
    w("                const double PQx = A.bra.x[i] - A.ket.x[kj];")
    w("                const double PQy = A.bra.y[i] - A.ket.y[kj];")
    w("                const double PQz = A.bra.z[i] - A.ket.z[kj];")
    w("                double R2 = PQx * PQx;")
    w("                R2 = fma(PQy, PQy, R2);")
    w("                R2 = fma(PQz, PQz, R2);")
    w("                const double alpha = PQalpha_mul * one_over_PQalpha_sum;")
    w("                double F[%d];" % (Ltot + 1))
    w("                boys_eval<%d>(grid, R2 * alpha, F);" % Ltot)
    w("                const double prefac = (P_prefac * Q_prefac) * sqrt(one_over_PQalpha_sum);")
    w("                double *c = pc + r0 * NPC * B + b;")
    w("                c[0 * B] = A.bra.%sx[i]; c[4 * B] = A.bra.%sy[i]; c[8 * B] = A.bra.%sz[i];" % (bra_pa, bra_pa, bra_pa))
    w("                c[12 * B] = A.ket.%sx[kj]; c[16 * B] = A.ket.%sy[kj]; c[20 * B] = A.ket.%sz[kj];" % (ket_pa, ket_pa, ket_pa))
    w("                const double a_over_p = (-alpha) * P_inv;")
    w("                const double a_over_q = alpha * Q_inv;")
    w("                c[1 * B] = a_over_p * PQx; c[5 * B] = a_over_p * PQy; c[9 * B] = a_over_p * PQz;")
    w("                c[13 * B] = a_over_q * PQx; c[17 * B] = a_over_q * PQy; c[21 * B] = a_over_q * PQz;")
    w("#pragma unroll")
    w("                for (int s = 0; s < 3; ++s) {")
    w("                    c[(4 * s + 2) * B] = a_over_p; c[(4 * s + 3) * B] = 0.5 * P_inv;")
    w("                    c[(4 * s + 14) * B] = -a_over_q; c[(4 * s + 15) * B] = 0.5 * Q_inv;")
    w("                }")
    w("                c[24 * B] = 0.5 * one_over_PQalpha_sum;")
    w("#pragma unroll")
    w("                for (int m = 0; m <= %d; ++m) c[(%d + m) * B] = F[m] * prefac;" % (Ltot, NPC_FIXED))
    w("            }")
    w("            __syncthreads();")
    w("            for (int pp = 0; pp < np_batch; ++pp) {")
    w("                const double *c = pc + pp * NPC * B + b;")
    w("                const double o2pq = c[24 * B];")
    w("                char *const wb = (char *)(work + b);")
    w("                const char *cb = (const char *)c;")
    w("                for (int m = r0; m <= %d; m += LPB) work[m * B + b] = c[(%d + m) * B];" % (Ltot, NPC_FIXED))
    w("                if (r0 == 0) { work[ZS * B + b] = 0.0; work[(ZS + 1) * B + b] = 0.0; }")
    w("                __syncthreads();")
    # Each thread's nodes of a level are independent: the sources of up to
    # LEVEL_CHUNK of them are loaded before any result is stored, so their
    # LDS latencies overlap.  The descriptors of a level (they do not depend
    # on LDS) are loaded while the level before it runs.  The thread row goes
    # through an empty asm so the descriptor loads stay in the loop (hoisted,
    # every level's descriptors would be live at once).
    w("                int rr = r0;")
    w('                asm volatile("" : "+v"(rr));')
    nlev = len(starts) - 1

    w("                const unsigned rofs = (unsigned)rr * 32u;")

    _, first = byte_desc(plan(am)[3], starts, B, LPB)

    def load_desc(lv):
        S, E = starts[lv], starts[lv + 1]
        nj = (E - S + LPB - 1) // LPB
        for j in range(nj):
            w("                const uint4 *dp%d_%d = (const uint4 *)((const char *)vrr_desc + (rofs + %du));"
              % (lv, j, (first[lv] + j * LPB) * 32))
            w("                const uint4 da%d_%d = dp%d_%d[0], db%d_%d = dp%d_%d[1];"
              % (lv, j, lv, j, lv, j, lv, j))

    def ld(base, off):
        return "*(const double *)(%s + %s)" % (base, off)

    if nlev > 1:
        load_desc(1)
    for lv in range(1, nlev):
        S, E = starts[lv], starts[lv + 1]
        nj = (E - S + LPB - 1) // LPB
        w("                // level %d: nodes [%d, %d), up to %d per thread" % (lv, S, E, nj))
        if lv + 1 < nlev:
            load_desc(lv + 1)
        chunk = LEVEL_CHUNK_ONE if B == 1 else LEVEL_CHUNK
        for j0 in range(0, nj, chunk):
            js = range(j0, min(nj, j0 + chunk))
            w("                {")
            for j in js:
                a, d = "da%d_%d" % (lv, j), "db%d_%d" % (lv, j)
                w("                    const double x0_%d = %s, x1_%d = %s;"
                  % (j, ld("wb", a + ".x"), j, ld("wb", a + ".x + 8 * B")))
                w("                    const double x2_%d = %s, x3_%d = %s;"
                  % (j, ld("wb", a + ".y"), j, ld("wb", a + ".y + 8 * B")))
                w("                    const double x4_%d = %s, x5_%d = %s;"
                  % (j, ld("wb", a + ".z"), j, ld("wb", a + ".z + 8 * B")))
                w("                    const double x6_%d = %s, x7_%d = %s;"
                  % (j, ld("wb", a + ".w"), j, ld("wb", d + ".x")))
                w("                    const char *cs%d = cb + %s.z;" % (j, d))
                w("                    const double kk%d = %s, aa%d = %s;"
                  % (j, ld("cs%d" % j, "0"), j, ld("cs%d" % j, "8 * B")))
                w("                    const double ao%d = %s, oo%d = %s;"
                  % (j, ld("cs%d" % j, "16 * B"), j, ld("cs%d" % j, "24 * B")))
            for j in js:
                f = "db%d_%d.w" % (lv, j)
                w("                    {")
                w("                        double t = kk%d * x0_%d;" % (j, j))
                w("                        t = fma(aa%d, x1_%d, t);" % (j, j))
                w("                        t = fma((double)(%s & 0xff) * oo%d, fma(ao%d, x3_%d, x2_%d), t);"
                  % (f, j, j, j, j))
                w("                        t = fma((double)((%s >> 8) & 0xff) * oo%d, fma(ao%d, x5_%d, x4_%d), t);"
                  % (f, j, j, j, j))
                w("                        t = fma((double)((%s >> 16) & 0xff) * o2pq, x6_%d, t);" % (f, j))
                w("                        t = fma((double)(%s >> 24) * o2pq, x7_%d, t);" % (f, j))
                w("                        *(double *)(wb + db%d_%d.y) = t;" % (lv, j))
                w("                    }")
            w("                }")
        w("                __syncthreads();")
    w("#pragma unroll")
    w("                for (int s = 0; s < NACC; ++s) {")
    w("                    const int t = r0 + LPB * s;")
    w("                    if (t < NTOP) acc[s] += work[top_slot[t] * B + b];")
    w("                }")
    w("                __syncthreads();")
    w("            }")
    w("        }")
    w("        double *tops = work;")
    w("        double *abf = work + NTOP * B;")
    w("        __syncthreads();")
    w("#pragma unroll")
    w("        for (int s = 0; s < NACC; ++s) {")
    w("            const int t = r0 + LPB * s;")
    w("            if (t < NTOP) tops[t * B + b] = acc[s];")
    w("        }")
    w("        __syncthreads();")
    _emit_hrr_coop(g, w, am, listmode)
    w("        __syncthreads();")
    w("    }")
    w("}")
    w("")
    if not listmode:
        w('extern "C" hipError_t simint_launch_wpq_%s(const EriArgs *a, hipStream_t s)' % name)
        w("{")
        w("    if (a->nitems <= 0) return hipSuccess;")
        w("    hipLaunchKernelGGL(eri_wpq_%s, dim3((unsigned)a->nitems), dim3(%d), 0, s, *a);" % (name, THREADS))
        w("    return hipGetLastError();")
        w("}")
        w("")
    else:
        w('extern "C" hipError_t simint_launch_wpql_%s(const EriListArgs *a, hipStream_t s)' % name)
        w("{")
        w("    if (a->nquartet <= 0) return hipSuccess;")
        w("    hipLaunchKernelGGL(eri_wpql_%s, dim3((unsigned)((a->nquartet + B - 1) / B)), dim3(%d), 0, s, *a);"
          % (name, THREADS))
        w("    return hipGetLastError();")
        w("}")


def _emit_hrr_coop(g, w, am, listmode=False):
    """HRR for the B quartets of a batch: bra HRR with lanes over (ket-top
    component f, quartet), then ket HRR with lanes over (final bra
    component, quartet).  LDS layout [value][quartet slot]."""
    la, lb, lc, ld = am
    bra_ops, bra_map, ket_ops, ket_map = g.hrr_program()
    bra_tops = g.bra_tops()
    ket_tops = g.ket_tops()
    nf = len(ket_tops)
    etop = {e: i for i, e in enumerate(bra_tops)}
    ftop = {f: i for i, f in enumerate(ket_tops)}
    bra_final = [(a, b) for a in carts(la) for b in carts(lb)]
    abidx = {ab: i for i, ab in enumerate(bra_final)}
    ind = "        "
    w(ind + "for (int f = r0; f < NF; f += LPB) {")
    w(ind + "    const double *T = tops + f * B + b;")
    names = {}
    cnt = [0]

    def bval(src):
        if src[0] == "top":
            return "T[%d * B]" % (etop[src[1]] * nf)
        return names[(src[1], src[2])]

    for (dst, s0, s1, sign, d) in bra_ops:
        cnt[0] += 1
        nm = "hb%d" % cnt[0]
        names[(dst[1], dst[2])] = nm
        ab = ("hAB[%d]" % d) if sign > 0 else ("(-hAB[%d])" % d)
        w(ind + "    const double %s = fma(%s, %s, %s);" % (nm, ab, bval(s1), bval(s0)))
    for ab in bra_final:
        w(ind + "    abf[(%d * NF + f) * B + b] = %s;" % (abidx[ab], bval(bra_map[ab])))
    w(ind + "}")
    w(ind + "__syncthreads();")
    w(ind + "if (b < nb) {")
    if listmode:
        if g.ket_hrr_needed():
            w(ind + "    const double hCD[3] = {A.ket.abx[kp], A.ket.aby[kp], A.ket.abz[kp]};")
        w(ind + "    double *__restrict__ o = A.out + (A.qout ? A.qout[qi] : qi * A.ostride);")
    else:
        if g.ket_hrr_needed():
            w(ind + "    const double hCD[3] = {A.ket.abx[tl], A.ket.aby[tl], A.ket.abz[tl]};")
        w(ind + "    double *__restrict__ o = A.out + wi.obase + A.ket.col[tl] * A.stride;")
    w(ind + "    for (int ab = r0; ab < NAB; ab += LPB) {")
    w(ind + "        const double *R = abf + ab * NF * B + b;")
    w(ind + "        double *oo = o + ab * NCD;")
    knames = {}
    cnt = [0]

    def kval(src):
        if src[0] == "top":
            return "R[%d * B]" % ftop[src[1]]
        return knames[(src[1], src[2])]

    for (dst, s0, s1, sign, d) in ket_ops:
        cnt[0] += 1
        nm = "hk%d" % cnt[0]
        knames[(dst[1], dst[2])] = nm
        cd = ("hCD[%d]" % d) if sign > 0 else ("(-hCD[%d])" % d)
        w(ind + "        const double %s = fma(%s, %s, %s);" % (nm, cd, kval(s1), kval(s0)))
    k = 0
    for c in carts(lc):
        for dd in carts(ld):
            w(ind + "        oo[%d] = %s;" % (k, kval(ket_map[(c, dd)])))
            k += 1
    w(ind + "    }")
    w(ind + "}")

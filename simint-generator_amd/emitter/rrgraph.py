"""Recurrence graphs for one (ab|cd) class: Obara-Saika VRR + HRR.

This is the emitter's model of the integral algorithm, written from the
mathematics and the reference generator's step-selection rules:

* HRR (generator/ostei/OSTEI_HRR_Algorithm_Base.cpp:160-283,
  generator/ostei/Makowski.hpp:26-82).  Bra moves I->J unless la < lb (then
  J->I); ket moves K->L unless lc < ld (then L->K).  A doublet is reduced in
  the direction of the lowest non-zero exponent of the stepped Gaussian,
  rightmost on ties:
      J: (a,b) = (a+1_i, b-1_i) + AB_i (a, b-1_i)
      I: (a,b) = (a-1_i, b+1_i) - AB_i (a-1_i, b)
  and the same for the ket with CD.  The HRR inputs ("tops") are the classes
  (e0|f0) (or (0e|, |0f)) for e in [la, la+lb], f in [lc, lc+ld].

* VRR (generator/ostei/OSTEI_VRR_Algorithm_Base.cpp:289-469,
  generator/ostei/Makowski.hpp:96-241).  Centres are reduced in the order of
  a stable sort of (3,2,1,0) by the largest AM each centre carries among the
  tops; a component quartet is reduced on the first centre in that order with
  non-zero AM, in the direction given by the same lowest-exponent rule.  The
  step has up to eight sources (arithmetic in OSTEI_VRR_Writer.cpp:95-125).

Counts follow the reference's conventions so they can be compared with its
worksize table (skel/simint/ostei/ostei_config.h.in:14-17) and op counts.
"""
from functools import lru_cache

I, J, K, L = 0, 1, 2, 3
STEPNAME = "IJKL"


def ncart(l):
    return (l + 1) * (l + 2) // 2


@lru_cache(maxsize=None)
def carts(l):
    """Cartesian exponents of AM l in Psi4 order (generator/Ordering_Psi4.cpp)."""
    out = []
    for ix in range(l, -1, -1):
        for iy in range(l - ix, -1, -1):
            out.append((ix, iy, l - ix - iy))
    return tuple(out)


def cart_index(e):
    l = sum(e)
    return (l - e[0]) * (l - e[0] + 1) // 2 + (l - e[0] - e[1])


def mk_dir(e):
    """Lowest non-zero exponent, rightmost on ties (Makowski.hpp:111-117)."""
    best = -1
    for k in (2, 1, 0):
        if e[k] > 0 and (best < 0 or e[k] < e[best]):
            best = k
    return best


def step(e, d, delta):
    e2 = list(e)
    e2[d] += delta
    if e2[d] < 0:
        return None
    return tuple(e2)


class VRRNode:
    __slots__ = ("q", "m", "step", "dir", "src", "fac")

    def __init__(self, q, m):
        self.q = q          # tuple of 4 exponent triples
        self.m = m
        self.step = None    # None for the (ss|ss)^m base
        self.dir = -1
        self.src = [None] * 8   # keys (q, m) or None
        self.fac = [0, 0, 0, 0]  # integer factors for s2/3, s4/5, s6, s7


class ClassGraph:
    """Recurrence graph for (la lb | lc ld), optionally widened for deriv1."""

    def __init__(self, la, lb, lc, ld):
        self.am = (la, lb, lc, ld)
        self.bra_j = la >= lb
        self.ket_l = lc >= ld
        self.e_lo = la if self.bra_j else lb
        self.e_hi = la + lb
        self.f_lo = lc if self.ket_l else ld
        self.f_hi = lc + ld
        self.Ltot = la + lb + lc + ld
        qmax = [0, 0, 0, 0]
        qmax[0 if self.bra_j else 1] = self.e_hi
        qmax[2 if self.ket_l else 3] = self.f_hi
        # stable sort of (3,2,1,0) by qmax (OSTEI_VRR_Algorithm_Base.cpp:463-466)
        self.order = sorted([3, 2, 1, 0], key=lambda c: qmax[c])
        self.nodes = {}      # (q, m) -> VRRNode
        self.node_list = []  # topological order (sources first)
        self.tops = []       # [(e, f)] exponent triples of the non-zero bra / ket centre
        for le in range(self.e_lo, self.e_hi + 1):
            for e in carts(le):
                for lf in range(self.f_lo, self.f_hi + 1):
                    for f in carts(lf):
                        self.tops.append((e, f))
        self.top_keys = [(self.top_quartet(e, f), 0) for (e, f) in self.tops]
        for k in self.top_keys:
            self._need(*k)

    # ------------------------------------------------------------------ VRR
    def top_quartet(self, e, f):
        z = (0, 0, 0)
        q = [z, z, z, z]
        q[0 if self.bra_j else 1] = e
        q[2 if self.ket_l else 3] = f
        return tuple(q)

    def _need(self, q, m):
        key = (q, m)
        if key in self.nodes:
            return key
        # iterative DFS to keep the recursion shallow
        stack = [key]
        while stack:
            k = stack[-1]
            if k in self.nodes:
                stack.pop()
                continue
            node = self._make(*k)
            missing = [s for s in node.src if s is not None and s not in self.nodes]
            if missing:
                stack.extend(missing)
                continue
            self.nodes[k] = node
            self.node_list.append(node)
            stack.pop()
        return key

    def _make(self, q, m):
        node = VRRNode(q, m)
        centre = None
        for c in self.order:
            if sum(q[c]) > 0:
                centre = c
                break
        if centre is None:
            return node
        d = mk_dir(q[centre])
        node.step = centre
        node.dir = d
        partner = {I: J, J: I, K: L, L: K}[centre]
        o1, o2 = (K, L) if centre in (I, J) else (I, J)

        def with_(base, c, delta):
            if base is None:
                return None
            g = step(base[c], d, delta)
            if g is None:
                return None
            b = list(base)
            b[c] = g
            return tuple(b)

        s0 = with_(q, centre, -1)
        sc2 = with_(q, centre, -2)
        spa = with_(s0, partner, -1)
        s6 = with_(s0, o1, -1)
        s7 = with_(s0, o2, -1)
        fc2 = q[centre][d] - 1
        fpa = q[partner][d]
        node.src[0] = (s0, m)
        node.src[1] = (s0, m + 1)
        first, second = (sc2, spa) if centre in (I, K) else (spa, sc2)
        node.fac[0], node.fac[1] = (fc2, fpa) if centre in (I, K) else (fpa, fc2)
        if first is not None:
            node.src[2] = (first, m)
            node.src[3] = (first, m + 1)
        if second is not None:
            node.src[4] = (second, m)
            node.src[5] = (second, m + 1)
        node.fac[2] = q[o1][d]
        node.fac[3] = q[o2][d]
        if s6 is not None:
            node.src[6] = (s6, m + 1)
        if s7 is not None:
            node.src[7] = (s7, m + 1)
        return node

    # ------------------------------------------------------------------ HRR
    def hrr_program(self):
        """Straight-line HRR: returns (bra_ops, ket_ops, final) where

        bra_ops: for each top ket Gaussian f, list of ops producing the final
                 bra doublets from the tops;
        ket_ops: for each final bra component, list of ops producing (ab|cd).
        An op is (dst, src0, src1, sign, dir): dst = src0 + sign*X[dir]*src1.
        Names are tuples so callers can map them to variables.
        """
        la, lb, lc, ld = self.am
        bra_final = [(a, b) for a in carts(la) for b in carts(lb)]
        ket_final = [(c, d) for c in carts(lc) for d in carts(ld)]

        def bra_reduce(a, b, memo, ops):
            key = (a, b)
            if key in memo:
                return memo[key]
            if self.bra_j:
                if sum(b) == 0:
                    memo[key] = ("top", a)
                    return memo[key]
                i = mk_dir(b)
                b1 = step(b, i, -1)
                a1 = step(a, i, +1)
                s0 = bra_reduce(a1, b1, memo, ops)
                s1 = bra_reduce(a, b1, memo, ops)
                ops.append((("bra", a, b), s0, s1, +1, i))
            else:
                if sum(a) == 0:
                    memo[key] = ("top", b)
                    return memo[key]
                i = mk_dir(a)
                a1 = step(a, i, -1)
                b1 = step(b, i, +1)
                s0 = bra_reduce(a1, b1, memo, ops)
                s1 = bra_reduce(a1, b, memo, ops)
                ops.append((("bra", a, b), s0, s1, -1, i))
            memo[key] = ("bra", a, b)
            return memo[key]

        bra_ops = []
        memo = {}
        for (a, b) in bra_final:
            bra_reduce(a, b, memo, bra_ops)
        bra_map = {k: memo[k] for k in bra_final}

        def ket_reduce(c, d, memo, ops):
            key = (c, d)
            if key in memo:
                return memo[key]
            if self.ket_l:
                if sum(d) == 0:
                    memo[key] = ("top", c)
                    return memo[key]
                i = mk_dir(d)
                d1 = step(d, i, -1)
                c1 = step(c, i, +1)
                s0 = ket_reduce(c1, d1, memo, ops)
                s1 = ket_reduce(c, d1, memo, ops)
                ops.append((("ket", c, d), s0, s1, +1, i))
            else:
                if sum(c) == 0:
                    memo[key] = ("top", d)
                    return memo[key]
                i = mk_dir(c)
                c1 = step(c, i, -1)
                d1 = step(d, i, +1)
                s0 = ket_reduce(c1, d1, memo, ops)
                s1 = ket_reduce(c1, d, memo, ops)
                ops.append((("ket", c, d), s0, s1, -1, i))
            memo[key] = ("ket", c, d)
            return memo[key]

        ket_ops = []
        memo = {}
        for (c, d) in ket_final:
            ket_reduce(c, d, memo, ket_ops)
        ket_map = {k: memo[k] for k in ket_final}
        return bra_ops, bra_map, ket_ops, ket_map

    def ket_tops(self):
        out = []
        for lf in range(self.f_lo, self.f_hi + 1):
            out.extend(carts(lf))
        return out

    def bra_tops(self):
        out = []
        for le in range(self.e_lo, self.e_hi + 1):
            out.extend(carts(le))
        return out

    # --------------------------------------------------------------- counts
    def counts(self):
        """Reference-convention sizes: bcont, prim, cont (worksize table)."""
        la, lb, lc, ld = self.am
        bcont = len(self.tops)
        maxm = {}
        comps = {}
        for node in self.node_list:
            cls = tuple(sum(g) for g in node.q)
            maxm[cls] = max(maxm.get(cls, -1), node.m)
            comps.setdefault(cls, set()).add(node.q)
        prim = sum(ncart(c[0]) * ncart(c[1]) * ncart(c[2]) * ncart(c[3]) * (mm + 1) for c, mm in maxm.items())
        # cont: HRR intermediates (bra-HRR classes x top ket AMs, plus final
        # bra x ket-HRR classes), excluding the final class itself
        bra_classes = set()
        ket_classes = set()
        bra_ops, _, ket_ops, _ = self.hrr_program()
        for (dst, _, _, _, _) in bra_ops:
            bra_classes.add((sum(dst[1]), sum(dst[2])))
        for (dst, _, _, _, _) in ket_ops:
            ket_classes.add((sum(dst[1]), sum(dst[2])))
        nket_top = sum(ncart(f) for f in range(self.f_lo, self.f_hi + 1))
        cont = 0
        for (a, b) in bra_classes:
            if not self.ket_hrr_needed() and (a, b) == (la, lb):
                continue
            cont += ncart(a) * ncart(b) * nket_top
        for (c, d) in ket_classes:
            if (c, d) == (lc, ld):
                continue
            cont += ncart(la) * ncart(lb) * ncart(c) * ncart(d)
        return {"bcont": bcont, "prim": prim, "cont": cont, "maxm": maxm, "comps": comps}

    def ket_hrr_needed(self):
        la, lb, lc, ld = self.am
        return (ld > 0) if self.ket_l else (lc > 0)

    def bra_hrr_needed(self):
        la, lb, lc, ld = self.am
        return (lb > 0) if self.bra_j else (la > 0)

    def flop_counts(self):
        """Reference-convention FLOPs (FMA = 2): f_vrr per primitive quartet
        (every needed component at every m up to its class maximum, as the
        reference emits it), f_setup per primitive quartet, f_boys, f_hrr per
        contracted quartet, bcont contraction adds."""
        c = self.counts()
        f_vrr = 0
        for cls, comps in c["comps"].items():
            if sum(cls) == 0:
                continue
            nm = c["maxm"][cls] + 1
            for q in comps:
                node = self.nodes.get((q, 0)) or next(self.nodes[(q, m)] for m in range(64) if (q, m) in self.nodes)
                ops = 1 + 2  # MUL + FMA(aop)
                if node.src[2] is not None:
                    ops += 4
                if node.src[4] is not None:
                    ops += 4
                if node.src[6] is not None:
                    ops += 2
                if node.src[7] is not None:
                    ops += 2
                f_vrr += ops * nm
        # integer constants k/(2p) etc (one MUL each for k > 1)
        ints = {"2p": set(), "2q": set(), "2pq": set()}
        for node in self.node_list:
            if node.step is None:
                continue
            bra = node.step in (I, J)
            if node.src[2] is not None:
                ints["2p" if bra else "2q"].add(node.fac[0])
            if node.src[4] is not None:
                ints["2p" if bra else "2q"].add(node.fac[1])
            if node.src[6] is not None:
                ints["2pq"].add(node.fac[2])
            if node.src[7] is not None:
                ints["2pq"].add(node.fac[3])
        f_const = sum(1 for s in ints.values() for k in s if k > 1)
        has_bra = any(n.step in (I, J) for n in self.node_list if n.step is not None)
        has_ket = any(n.step in (K, L) for n in self.node_list if n.step is not None)
        # setup (OSTEI_Writer.cpp:595-702): mul, sum, div, 3 sub, R2 (1+2+2),
        # alpha, 1/p, 1/q, 1/2p, 1/2q, 1/2pq, F_x, sqrt + 2 mul, (L+1) scaling
        f_setup = 1 + 1 + 1 + 3 + 5 + 1 + 1 + 1 + 1 + 1 + 1 + 1 + 3 + (self.Ltot + 1)
        if has_bra:
            f_setup += 2 + 3      # a_over_p (neg, mul), aop_PQ
        if has_ket:
            f_setup += 1 + 3 + 1  # a_over_q, aoq_PQ, neg
        L_ = self.Ltot
        f_boys = 20 * (L_ + 1) if L_ < 4 else 20 + 2 + 3 * L_
        bra_ops, _, ket_ops, _ = self.hrr_program()
        nket_top = len(self.ket_tops())
        la, lb, lc, ld = self.am
        f_hrr = 2 * (len(bra_ops) * nket_top + len(ket_ops) * ncart(la) * ncart(lb))
        return {"f_vrr": f_vrr, "f_setup": f_setup + f_const, "f_boys": f_boys, "bcont": c["bcont"],
                "f_hrr": f_hrr, "prim": c["prim"], "cont": c["cont"]}


@lru_cache(maxsize=None)
def graph(la, lb, lc, ld):
    return ClassGraph(la, lb, lc, ld)

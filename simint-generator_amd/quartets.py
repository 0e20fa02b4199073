"""Screened shell-quartet sets on the device: the caller side of the hot path.

Builds, for a list of shells, every unique shell quartet (ij|kl) with
i >= j, k >= l, ij >= kl whose Schwarz bound sqrt((ij|ij)) sqrt((kl|kl))
reaches a threshold, and runs them through simint_hip_compute_eri_prefix.

Layout.  Unique pairs are oriented so am1 >= am2 and grouped by (am1, am2)
("pair class") and by primitive-pair count, so the 64 kets of a tile share
one primitive loop.  Inside a group, pairs are sorted by their bound (the
largest diagonal element of (ij|ij), computed on the GPU from the diagonal
quartets, shell_screen.c:9-60), descending: the kets that survive the screen
for a given bra are then a prefix of each group.  Quartets between groups
are counted once: a bra meets ket groups of lower (class, group) keys fully
and its own group up to itself.

Everything a pass needs is resident in HBM; a pass only uploads the per-bra
prefix lengths and output offsets of each launch.
"""
import ctypes
import threading

import numpy as np

import simint


class Launch:
    """One class launch: bras b0.. of group `bra`, kets [lo[i], nket[i]) of
    group `ket` (lo None: from 0), quartet (b, k) at off[i] + (k - lo[i]) * nc,
    in FP64 (precision 0) or with the FP32 VRR/HRR (1).  `chunk` numbers the
    output-buffer chunk the launch writes (launches of one chunk share it)."""
    __slots__ = ("bra", "ket", "b0", "nket", "off", "nq", "flops", "nout", "cls", "lo", "precision", "chunk")

    def __init__(self, bra, ket, b0, nket, off, nq, flops, nout, cls, lo=None, precision=0, chunk=0):
        self.bra, self.ket, self.b0, self.nket, self.off = bra, ket, b0, nket, off
        self.nq, self.flops, self.nout, self.cls = nq, flops, nout, cls
        self.lo, self.precision, self.chunk = lo, precision, chunk


class Group:
    """One (pair class, primitive count) group of unique pairs, sorted by bound.
    With a device shell set the pairs are filled on the device and the
    bounds stay there too (dbounds)."""

    def __init__(self, basis, cls, nprim, pairs, bounds, dshells=None, dbounds=None, screen=simint.SCREEN_NONE):
        self.cls = cls
        self.nprim = nprim
        self.pairs = pairs
        self.bounds = bounds
        self.dbounds = dbounds
        if dshells is None:
            self.host = simint.Pairs.from_pairs(basis, pairs, screen)
            self.dev = simint.DevicePairs(self.host)
        else:
            self.host = None
            self.dev = simint.DevicePairs.from_shells(dshells, pairs, cls)


def _class_flops(info, la, lb, lc, ld):
    c = info[simint_class_name((la, lb, lc, ld))]
    per_prim = c["f_setup"] + c["f_vrr"] + c["f_boys"] + c["bcont"]
    return per_prim, c["f_hrr"], c["ncart"]


def simint_class_name(am):
    return "_".join("spdfghik"[x] for x in am)


def mixed_threshold(cls, base=1e-4):
    """Schwarz bound below which a quartet of class cls runs the FP32-VRR
    kernel.  The FP32 rounding error of the VRR outputs is relative to their
    size; the HRR transfers (a,b+1| = (a+1,b| + AB (a,b| cancel strongly for
    the far-apart pairs that have small bounds and amplify it, by about 40x
    with one HRR side and about 10^4 with both (measured on shells of the
    64-atom recipe).  The base threshold applies to classes without HRR;
    the others get it divided by 400 and 10^5, which keeps the absolute
    error of every class at or below 1e-10 on the 256-atom system (measured:
    1.9e-10 with divisors 100 and 2.5e4)."""
    la, lb, lc, ld = cls
    bra_hrr = min(la, lb) > 0
    ket_hrr = min(lc, ld) > 0
    if bra_hrr and ket_hrr:
        return base / 1e5
    if bra_hrr or ket_hrr:
        return base / 400.0
    return base


class ScreenedSet:
    """All unique quartets of `shells` with sqrt((ab|ab)(cd|cd)) >= tol.
    mixed: base threshold (see mixed_threshold) below which a quartet's
    bound sends it to the FP32-VRR kernels (None: every quartet FP64)."""

    def __init__(self, shells, tol=1e-12, maxam=2, out_doubles=1 << 29, mixed=None, deriv=0, front_end="host",
                 prim_tol=0.0):
        """front_end "host": shell pairs filled on the host (the drop-in
        simint_create_multi_shellpair2) and uploaded, bounds sorted and
        screened quartets counted in numpy; "device": pairs filled, bounds
        computed and sorted, and counts taken on the GPU from a device copy
        of the shells (only the shell-index lists go up and the per-bra
        counts come back).
        prim_tol > 0: the pairs carry Schwarz primitive bounds
        (SIMINT_SCREEN_SCHWARZ) and the kernels skip every primitive
        quartet whose bound product is below prim_tol (simint's screen_tol
        in the primitive loop); each skipped one moves an integral by at
        most prim_tol.  Host front end and ERIs only."""
        import time
        import torch
        t0 = time.perf_counter()
        self.tol = tol
        self.mixed = mixed
        self.deriv = deriv
        self.front_end = front_end
        if prim_tol < 0 or (prim_tol > 0 and (front_end != "host" or deriv)):
            raise ValueError("prim_tol: a tolerance >= 0, with the host front end and plain ERIs")
        self.prim_tol = float(prim_tol)
        screen = simint.SCREEN_SCHWARZ if prim_tol > 0 else simint.SCREEN_NONE
        self.basis = simint.Basis(shells)
        self.shells = shells
        self.out_doubles = out_doubles
        self._fun_lock = threading.Lock()
        L = simint.lib()
        L.simint_init()
        # unique pairs by class
        byclass = {}
        for i in range(len(shells)):
            for j in range(i + 1):
                a, b = (i, j) if shells[i][0] >= shells[j][0] else (j, i)
                byclass.setdefault((shells[a][0], shells[b][0]), []).append((a, b))
        # Schwarz bounds from diagonal quartets on the GPU
        self.groups = []
        self.dshells = simint.DeviceShells(self.basis) if front_end == "device" else None
        for cls in (sorted(byclass) if front_end == "device" else []):
            pairs = np.array(byclass[cls], dtype=np.int32)
            P = simint.DevicePairs.from_shells(self.dshells, pairs, cls)
            bound = torch.empty(len(pairs), dtype=torch.float64, device="cuda")
            order = torch.empty(len(pairs), dtype=torch.int32, device="cuda")
            P.bounds(bound.data_ptr())
            P.order(bound.data_ptr(), order.data_ptr())
            simint.synchronize()
            P.free()
            sb = bound.index_select(0, order.long())
            ordh = order.cpu().numpy()
            sp = pairs[ordh]
            npr = np.array([self._pair_nprim(shells, int(a), int(b)) for a, b in sp], dtype=np.int64)
            edges = np.flatnonzero(np.diff(npr)) + 1
            for s0, s1 in zip(np.r_[0, edges], np.r_[edges, len(sp)]):
                self.groups.append(Group(self.basis, cls, int(npr[s0]), [tuple(int(x) for x in q) for q in sp[s0:s1]],
                                         sb[s0:s1].cpu().numpy(), self.dshells, sb[s0:s1].contiguous()))
        for cls in (sorted(byclass) if front_end != "device" else []):
            pairs = byclass[cls]
            tmp = simint.Pairs.from_pairs(self.basis, pairs)
            dev = simint.DevicePairs(tmp)
            nab = simint.ncart(cls[0]) * simint.ncart(cls[1])
            out = torch.empty(len(pairs) * nab * nab, dtype=torch.float64, device="cuda")
            simint.compute_diag(dev, out.data_ptr())
            simint.synchronize()
            diag = out.view(len(pairs), nab, nab).diagonal(dim1=1, dim2=2).abs().amax(dim=1).cpu().numpy()
            nprim = tmp.field("nprim12")
            dev.free()
            tmp.free()
            for np_ in sorted(set(nprim.tolist()), reverse=True):
                sel = np.nonzero(nprim == np_)[0]
                order = sel[np.argsort(-diag[sel], kind="stable")]
                self.groups.append(Group(self.basis, cls, int(np_), [pairs[k] for k in order], diag[order].copy(),
                                         screen=screen))
        import json
        import os
        here = os.path.dirname(os.path.abspath(__file__))
        self.info = json.load(open(os.path.join(here, "csrc", "gen", "classes.json")))
        self.setup_pairs_s = time.perf_counter() - t0
        self._plan()
        self.setup_s = time.perf_counter() - t0

    @staticmethod
    def _pair_nprim(shells, a, b):
        na, nb = len(shells[a][2]), len(shells[b][2])
        return na * (na + 1) // 2 if a == b else na * nb

    def _counts(self, gb, gk, tol):
        """Per bra of gb: kets of gk (bound-sorted) with bound_b bound_k >=
        tol^2 (within one group at most b + 1)."""
        if self.front_end == "device":
            import torch
            out = torch.empty(len(gb.pairs), dtype=torch.int32, device="cuda")
            simint.screen_counts(len(gb.pairs), gb.dbounds.data_ptr(), len(gk.pairs), gk.dbounds.data_ptr(),
                                 gk is gb, tol, out.data_ptr())
            simint.synchronize()
            return out.cpu().numpy().astype(np.int64)
        kb = gk.bounds
        with np.errstate(divide="ignore", over="ignore"):
            need = np.where(gb.bounds > 0, tol * tol / gb.bounds, np.inf)
        # kb is descending: count of kb >= need
        n = np.searchsorted(-kb, -need, side="right").astype(np.int64)
        if gk is gb:
            n = np.minimum(n, np.arange(1, len(gb.bounds) + 1))
        return n

    def _plan(self):
        """Per (bra group, ket group): prefix lengths for every bra, chunked
        so that one launch's output fits the output buffer."""
        self.launches = []
        self.nchunk = 0
        self.total_quartets = 0
        self.total_flops = 0.0
        self.total_bytes = 0.0
        G = self.groups
        for gi, gb in enumerate(G):
            for gk in G[:gi + 1]:
                la, lb = gb.cls
                lc, ld = gk.cls
                pp, fh, nc = _class_flops(self.info, la, lb, lc, ld)
                if self.deriv:
                    nc *= 12
                # kets surviving the screen: bound_b * bound_k >= tol^2
                nket = self._counts(gb, gk, self.tol)
                K = gb.nprim * gk.nprim
                cls = (la, lb, lc, ld)
                # kets whose bound keeps them in FP64 on a mixed plan
                n64 = None if self.mixed is None else self._counts(gb, gk, mixed_threshold(cls, self.mixed))
                kb = n64
                cap = self.out_doubles
                # chunks: runs of whole bras whose output fits `cap`; a bra
                # whose own kets exceed it is cut into ket ranges
                b = 0
                nb = len(nket)
                while b < nb:
                    if nket[b] * nc > cap:
                        step = max(1, cap // nc)
                        for k0 in range(0, int(nket[b]), step):
                            k1 = min(k0 + step, int(nket[b]))
                            self._emit_chunk(gb, gk, cls, K, pp, fh, nc, kb, b, b + 1,
                                             np.array([k0], dtype=np.int64), np.array([k1], dtype=np.int64))
                        b += 1
                        continue
                    cum = np.cumsum(nket[b:] * nc)
                    e = min(nb, b + int(np.searchsorted(cum, cap, side="right")))
                    self._emit_chunk(gb, gk, cls, K, pp, fh, nc, kb, b, e, None, nket[b:e])
                    b = e
        self._check_plan()

    def _emit_chunk(self, gb, gk, cls, K, pp, fh, nc, kb, b, e, lo, hi):
        """Launches of one output chunk: bras b..e-1, kets [lo, hi) each
        (lo None: from 0), laid out densely from the chunk start."""
        lo0 = np.zeros_like(hi) if lo is None else lo
        seg = hi - lo0
        nq = int(seg.sum())
        if nq <= 0:
            return
        off = np.zeros(len(seg), dtype=np.int64)
        off[1:] = np.cumsum(seg[:-1] * nc)
        if self.mixed is None:
            parts = [(lo, hi, off, 0)]
        else:
            # kets are bound-sorted: FP64 for the leading kets with
            # bound_b * bound_k >= threshold^2 (their count per bra in kb)
            n64 = np.clip(kb[b:e], lo0, hi)
            parts = [(lo, n64, off, 0), (n64, hi, off + (n64 - lo0) * nc, 1)]
        for plo, phi, oo, prec in parts:
            cnt = phi - (plo if plo is not None else 0)
            nqp = int(cnt.sum())
            if nqp == 0:
                continue
            fl = float(nqp) * (K * pp + fh)
            self.launches.append(Launch(gb, gk, b, phi.astype(np.int32), oo, nqp, fl, nqp * nc, cls,
                                        None if plo is None else plo.astype(np.int32), prec, self.nchunk))
            self.total_flops += fl
        self.nchunk += 1
        self.total_quartets += nq
        self.total_bytes += 8.0 * nq * nc

    def _check_plan(self):
        """Every launch's output stays inside the chunk buffer (a launch with
        a region past it would write out of bounds on the device)."""
        for ln in self.launches:
            nc = ln.nout // max(ln.nq, 1)
            lo = np.zeros_like(ln.nket) if ln.lo is None else ln.lo
            cnt = ln.nket.astype(np.int64) - lo
            if len(cnt) and ((ln.off + cnt * nc).max() > self.out_doubles or cnt.min() < 0):
                raise RuntimeError("launch plan exceeds the output buffer (class %s)" % (ln.cls,))

    def chunk_launches(self, chunk):
        """Launches writing output chunk `chunk` (index built on first use)."""
        cache = self.__dict__.setdefault("_by_chunk", {})
        ent = cache.get(id(self.launches))
        if ent is None or ent[0] is not self.launches:
            by = {}
            for ln in self.launches:
                by.setdefault(ln.chunk, []).append(ln)
            ent = cache[id(self.launches)] = (self.launches, by)
        return ent[1].get(chunk, [])

    def chunk_doubles(self, chunk):
        """Extent of output chunk `chunk` in doubles."""
        n = 0
        for ln in self.chunk_launches(chunk):
            lo = 0 if ln.lo is None else int(ln.lo[-1])
            n = max(n, int(ln.off[-1]) + (int(ln.nket[-1]) - lo) * (ln.nout // max(ln.nq, 1)))
        return n

    def deriv_worksize(self):
        """Device scratch (doubles) the largest derivative launch needs."""
        return max([simint.deriv1_worksize(ln.bra.dev, ln.ket.dev, ln.nq) for ln in self.launches] + [1])

    def run_pass(self, out_ptr, events=None, chunk=None, work_ptr=None):
        """Queue every launch of one pass (or of one output chunk) on the
        library stream.  If events is a callable, it is called before and
        after each launch with the launch (to record timing events on the
        library stream)."""
        self.run_launches(out_ptr, self.launches if chunk is None else self.chunk_launches(chunk), events,
                          work_ptr)

    def run_launches(self, out_ptr, launches, events=None, work_ptr=None):
        """Queue the given launches of this set's plans (as run_pass)."""
        self.work_ptr = work_ptr
        if self.deriv and not work_ptr:
            raise ValueError("a derivative pass needs a work buffer (deriv_worksize doubles)")
        for ln in launches:
            if events:
                events(ln, True)
            if self.deriv:
                simint.compute_deriv1_range(ln.bra.dev, ln.ket.dev, ln.b0, ln.lo, ln.nket, self.work_ptr, out_ptr)
            elif ln.lo is None and ln.precision == 0:
                simint.compute_prefix(ln.bra.dev, ln.ket.dev, ln.b0, ln.nket, ln.off, out_ptr, self.prim_tol)
            else:
                simint.compute_range(ln.bra.dev, ln.ket.dev, ln.b0, ln.lo, ln.nket, ln.off, out_ptr, self.prim_tol,
                                     ln.precision)
            if events:
                events(ln, False)

    def sample(self, n, seed=0, precision=None):
        """n random quartets of the screened set (of the launches of one
        precision if given), as shell-index quadruples, weighted uniformly
        over quartets."""
        rng = np.random.default_rng(seed)
        w = np.array([ln.nq if precision is None or ln.precision == precision else 0 for ln in self.launches],
                     dtype=np.float64)
        picks = rng.choice(len(self.launches), size=n, p=w / w.sum())
        out = []
        for li in picks:
            ln = self.launches[li]
            lo = ln.lo if ln.lo is not None else np.zeros_like(ln.nket)
            cum = np.cumsum(ln.nket - lo)
            r = int(rng.integers(0, cum[-1]))
            bi = int(np.searchsorted(cum, r, side="right"))
            k = int(lo[bi]) + r - (cum[bi - 1] if bi > 0 else 0)
            bp = ln.bra.pairs[ln.b0 + bi]
            kp = ln.ket.pairs[k]
            out.append((bp[0], bp[1], kp[0], kp[1]))
        return out

    def located_sample(self, n, seed=0, precision=None, cls=None, launches=None):
        """n random quartets of the screened set (or of the given launches)
        with where a pass writes them: a list of (chunk, offset in the chunk
        buffer, doubles, (i, j, k, l)), uniform over the quartets of the
        launches of one precision and one class if given."""
        rng = np.random.default_rng(seed)
        pool = self.launches if launches is None else launches
        w = np.array([ln.nq if (precision is None or ln.precision == precision) and (cls is None or ln.cls == cls)
                      else 0 for ln in pool], dtype=np.float64)
        if not len(w) or w.sum() == 0:
            return []
        out = []
        for li in rng.choice(len(pool), size=n, p=w / w.sum()):
            ln = pool[li]
            lo = ln.lo if ln.lo is not None else np.zeros_like(ln.nket)
            cum = np.cumsum(ln.nket.astype(np.int64) - lo)
            r = int(rng.integers(0, cum[-1]))
            bi = int(np.searchsorted(cum, r, side="right"))
            j = r - (int(cum[bi - 1]) if bi > 0 else 0)
            k = int(lo[bi]) + j
            nc = ln.nout // ln.nq
            bp = ln.bra.pairs[ln.b0 + bi]
            kp = ln.ket.pairs[k]
            out.append((ln.chunk, int(ln.off[bi]) + j * nc, nc, (bp[0], bp[1], kp[0], kp[1])))
        return out

    def device_bytes(self):
        """Device memory the set holds (its shell-pair groups)."""
        return sum(g.dev.bytes for g in self.groups)

    def nbf(self):
        """Number of Cartesian basis functions of the shells."""
        return int(sum(simint.ncart(s[0]) for s in self.shells))

    def _group_fun(self, g):
        """Device array of the first basis function of both shells of every
        pair of group g (built once, on first use).  Built under a lock:
        threads running tiles of one set side by side would otherwise each
        build it, and replacing one thread's array frees it while that
        thread's digestion still holds its address."""
        with self._fun_lock:
            if getattr(g, "dfun", None) is None:
                import torch
                f0 = np.r_[0, np.cumsum([simint.ncart(s[0]) for s in self.shells])].astype(np.int32)
                fun = f0[np.asarray(g.pairs, dtype=np.int64).reshape(-1)]
                g.dfun = torch.from_numpy(np.ascontiguousarray(fun, dtype=np.int32)).cuda()
                torch.cuda.current_stream().synchronize()
            return g.dfun

    def run_jk(self, out_ptr, d_density, d_J, d_K, launches=None, events=None):
        """Queue the launches (default: the whole pass) with the J/K
        digestion of their quartets after each class kernel: J and K
        (device nbf x nbf, accumulated) receive every unique quartet of the
        set once; call simint.jk_symmetrize after the last launch."""
        if self.deriv:
            raise ValueError("J/K digestion takes plain integrals, not derivatives")
        nbf = self.nbf()
        for ln in (self.launches if launches is None else launches):
            if events:
                events(ln, True)
            jk = simint.JK(nbf, 1 if ln.bra is ln.ket else 0, self._group_fun(ln.bra).data_ptr(),
                           self._group_fun(ln.ket).data_ptr(), d_density, d_J, d_K)
            simint.compute_jk_range(ln.bra.dev, ln.ket.dev, ln.b0, ln.lo, ln.nket, ln.off, out_ptr, jk, self.prim_tol,
                                    ln.precision)
            if events:
                events(ln, False)

    def free(self):
        for g in self.groups:
            g.dev.free()
            if g.host is not None:
                g.host.free()
        if self.dshells is not None:
            self.dshells.free()
            self.dshells = None
        self.basis.free()

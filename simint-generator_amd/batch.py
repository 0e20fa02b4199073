"""Batched quartet streams on the device: any mix of classes and
contraction depths in one submission (quartet-list kernels, one launch per
class, optionally recorded into one graph).

    b = QuartetBatch(basis, quartets)          # quartets: (n, 4) shell indices
    out = torch.empty(b.out_doubles, dtype=torch.float64, device="cuda")
    b.run(out.data_ptr())                       # or: g = b.capture(ptr); g.launch()
    # quartet i's integrals: out[b.offsets[i] : b.offsets[i] + b.sizes[i]]

Setup (host, untimed): quartets are grouped by class and, within a class,
ordered by primitive-quartet count (largest first) so the lanes of a
wavefront run loops of nearly equal length; each class gets a bra and a ket
shell-pair set (simint_create_multi_shellpair2 over its quartets' pairs, in
that order) uploaded once, plus device index and output-offset arrays.
Output offsets follow the stream order of the quartets, so the result is
laid out as if the quartets were computed one by one.
"""
import ctypes

import numpy as np

import simint


def _pair_set(basis, a_idx, b_idx, screen):
    """simint_create_multi_shellpair2 over pairs (a_idx[k], b_idx[k])."""
    n = len(a_idx)
    raw = np.frombuffer(basis.arr, dtype=np.dtype((np.void, ctypes.sizeof(simint.Shell))))
    ab = np.empty(2 * n, dtype=raw.dtype)
    ab[0::2] = raw[a_idx]
    ab[1::2] = raw[b_idx]
    AB = (simint.Shell * max(2 * n, 1)).from_buffer_copy(ab.tobytes() if n else bytes(ctypes.sizeof(simint.Shell)))
    p = simint.Pairs.__new__(simint.Pairs)
    p.mp = simint.MultiShellpair()
    L = simint.lib()
    L.simint_initialize_multi_shellpair(ctypes.byref(p.mp))
    L.simint_create_multi_shellpair2(n, AB, ctypes.byref(p.mp), screen)
    p.n = n
    return p


class _ClassGroup:
    pass


class QuartetBatch:
    def __init__(self, basis, quartets, screen=simint.SCREEN_NONE, screen_tol=0.0, deriv=0):
        import torch
        q = np.ascontiguousarray(quartets, dtype=np.int64).reshape(-1, 4)
        self.n = len(q)
        self.deriv = deriv
        self.screen_tol = screen_tol
        am = np.array([basis.arr[i].am for i in range(basis.n)], dtype=np.int64)
        npr = np.array([basis.arr[i].nprim for i in range(basis.n)], dtype=np.int64)
        qa = am[q]
        nc = ((qa + 1) * (qa + 2) // 2).prod(axis=1) * (12 if deriv else 1)
        self.sizes = nc
        self.offsets = np.zeros(self.n, dtype=np.int64)
        if self.n > 1:
            self.offsets[1:] = np.cumsum(nc)[:-1]
        self.out_doubles = int(nc.sum())
        # primitive-quartet count (same-shell pairs keep the triangle)
        def pairk(x, y):
            kx, ky = npr[x], npr[y]
            return np.where(x == y, kx * (kx + 1) // 2, kx * ky)
        kq = pairk(q[:, 0], q[:, 1]) * pairk(q[:, 2], q[:, 3])
        cls = ((qa[:, 0] * 8 + qa[:, 1]) * 8 + qa[:, 2]) * 8 + qa[:, 3]
        order = np.lexsort((-kq, cls))
        bounds = np.flatnonzero(np.diff(cls[order])) + 1
        self.groups = []
        work = 0
        for idx in np.split(order, bounds) if self.n else []:
            g = _ClassGroup()
            g.cls = tuple(int(x) for x in qa[idx[0]])
            g.n = len(idx)
            g.kmax = int(kq[idx[0]])
            g.quartets = int(g.n)
            g.Pp = _pair_set(basis, q[idx, 0], q[idx, 1], screen)
            g.Qp = _pair_set(basis, q[idx, 2], q[idx, 3], screen)
            g.P = simint.DevicePairs(g.Pp)
            g.Q = simint.DevicePairs(g.Qp)
            g.Pp.free()
            g.Qp.free()
            g.idx = torch.arange(g.n, dtype=torch.int32, device="cuda")
            g.qout = torch.from_numpy(self.offsets[idx]).to("cuda")
            g.prim_quartets = int(kq[idx].sum())
            if deriv:
                work = max(work, simint.deriv1_worksize(g.P, g.Q, g.n))
            self.groups.append(g)
        self.work = torch.empty(max(work, 1), dtype=torch.float64, device="cuda") if deriv else None

    def run(self, d_out, events=None, streams=None):
        """Queue every class on the library stream (asynchronous).  events(g,
        start) is called before and after each class launch if given.
        streams: spread the classes round-robin over this many side streams,
        forked from and joined back into the library stream (classes with
        few quartets each fill only part of the GPU; side by side they fill
        it).  Inside a capture the fork and join become graph edges."""
        if streams:
            return self._run_streams(d_out, streams)
        for g in self.groups:
            if events:
                events(g, True)
            if self.deriv:
                simint.compute_deriv1_list(g.P, g.Q, g.n, g.idx.data_ptr(), g.idx.data_ptr(), g.qout.data_ptr(),
                                           self.work.data_ptr(), d_out, self.screen_tol)
            else:
                self._list(g, d_out)
            if events:
                events(g, False)

    def _list(self, g, d_out):
        """One class's list launch, with its tuned lanes per quartet (if
        any) for this launch only."""
        lanes = simint.list_split_of(g.cls)
        L = simint.lib()
        prev = L.simint_hip_set_list_split(lanes) if lanes else None
        if prev is not None and prev < 0:
            raise ValueError("tuned list split of %s: %d lanes not accepted" % (g.cls, lanes))
        try:
            simint.compute_list(g.P, g.Q, g.n, g.idx.data_ptr(), g.idx.data_ptr(), g.qout.data_ptr(), d_out,
                                self.screen_tol)
        finally:
            if prev is not None:
                L.simint_hip_set_list_split(prev)

    def _prepare(self, n):
        """Side streams and per-stream derivative scratch, allocated before
        any capture (no allocation may happen while one records)."""
        import torch
        if getattr(self, "_side", None) is None or len(self._side) != n:
            self._side = [torch.cuda.Stream() for _ in range(n)]
            self._works = [self.work] + [torch.empty_like(self.work) for _ in range(n - 1)] if self.deriv else None

    def _run_streams(self, d_out, n):
        import torch
        L = simint.lib()
        self._prepare(n)
        prev = L.simint_hip_stream()     # the stream the caller's calls queue on (restored below)
        main = torch.cuda.ExternalStream(prev)
        fork = torch.cuda.Event()
        fork.record(main)
        for s in self._side:
            s.wait_event(fork)
        # largest classes first, so the long launches start early
        order = sorted(self.groups, key=lambda g: -g.prim_quartets * (1 + sum(g.cls)) ** 3)
        try:
            for k, g in enumerate(order):
                L.simint_hip_set_stream(ctypes.c_void_p(self._side[k % n].cuda_stream))
                if self.deriv:
                    simint.compute_deriv1_list(g.P, g.Q, g.n, g.idx.data_ptr(), g.idx.data_ptr(),
                                               g.qout.data_ptr(), self._works[k % n].data_ptr(), d_out,
                                               self.screen_tol)
                else:
                    self._list(g, d_out)
        finally:
            L.simint_hip_set_stream(ctypes.c_void_p(prev))
        for s in self._side:
            e = torch.cuda.Event()
            e.record(s)
            main.wait_event(e)

    def capture(self, d_out, streams=None):
        """The whole batch recorded into one graph (replay with .launch())."""
        if streams:
            self._prepare(streams)
        with simint.Graph.capture() as gr:
            self.run(d_out, streams=streams)
        return gr

    def free(self):
        for g in self.groups:
            g.P.free()
            g.Q.free()
        self.groups = []

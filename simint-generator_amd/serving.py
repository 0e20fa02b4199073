"""Continuous-batching ERI service on one GPU.

Requests arrive as a Poisson stream.  Each asks for the Schwarz-screened
ERIs (or first derivatives) of one molecule: a synthetic one built with the
64-atom-system recipe (its own atom count and seed), or explicit shells.
The server

  * rejects a malformed request (a non-finite or non-positive exponent, a
    non-finite coefficient or centre, an angular momentum above d, an empty
    shell) before it touches the device, without disturbing the others;
  * admits a request when its estimated device footprint fits the
    remaining HBM budget, then charges what its set-up actually allocated
    (the shell-pair data it keeps resident) -- requests that arrive while
    others run are admitted as memory frees, not after the queue drains;
  * cuts every admitted request into tiles (one output chunk of its
    screened quartet set each) and lets its worker threads -- one HIP
    stream each, so their kernels overlap on the device -- take tiles across
    requests, the request with the least work left first: a large request
    no longer holds a slot for its whole life, and small ones that arrive
    behind it finish without waiting for it;
  * reduces every tile on the device into that request's per-tile digest
    (sum of values and sum of squares in FP64) and sums the per-tile digests
    in tile order at the end, so a request's result does not depend on how
    its tiles were scheduled -- it is the same alone and in the mix;
  * for a "jk" request (a Fock build's two-electron part), also digests
    every tile's integrals into the request's Coulomb and exchange matrices
    of a density (simint_hip_compute_jk_range; the density is the request's
    own or one drawn from its seed) and returns J and K -- equal alone and
    in the mix to the last bits, the atomic sums' order aside;
  * stops scheduling a request's tiles once it has run past its time limit
    and reports it as timed out.

Latency is arrival to completion; throughput is shell quartets (derivative
quartets counted once) per second of wall time.

    python simint-generator_amd/serving.py [--requests 24] [--rate 2.0] ...
"""
import argparse
import json
import math
import os
import queue
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import quartets  # noqa: E402
import simint  # noqa: E402
import workloads  # noqa: E402

MAXAM = 2   # shells the screened sets take (the recipe's s, p, d)


class Request:
    def __init__(self, rid, natom, seed, kind, arrival, shells=None, timeout=None, density=None):
        self.rid, self.natom, self.seed, self.kind, self.arrival = rid, natom, seed, kind, arrival
        self.shells = shells       # explicit shells, or None: the recipe for (natom, seed)
        self.timeout = timeout     # seconds after admission, or None
        self.density = density     # "jk": nbf x nbf symmetric density, or None: drawn from the seed
        self.J = self.K = None     # "jk": the results (host arrays)
        self.jk_dev = None
        self.t_admit = self.t_done = None
        self.quartets = 0
        self.digest = None
        self.samples = None        # [(quartet, values)] when the server probes results
        self.error = None
        # scheduling state
        self.ss = None
        self.sums = None
        self.next = 0
        self.done_tiles = 0
        self.inflight = 0
        self.charged = 0.0
        self.probes = {}
        self.kept = []
        self.left = 0.0

    def to_dict(self):
        return {"id": self.rid, "natom": self.natom, "seed": self.seed, "kind": self.kind,
                "arrival": self.arrival, "admit": self.t_admit, "done": self.t_done,
                "latency": None if self.t_done is None else self.t_done - self.arrival,
                "quartets": self.quartets, "digest": self.digest, "error": self.error}


def request_stream(n, rate, seed=0, natom_range=(10, 200), deriv_frac=0.3, jk_frac=0.0):
    """n requests: Poisson arrivals at `rate` per second, atom counts uniform
    in natom_range, 30% first-derivative requests, per-request seeds; with
    jk_frac, that share of the ERI requests asks for J and K instead of the
    digest alone."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        natom = int(rng.integers(natom_range[0], natom_range[1] + 1))
        kind = "deriv" if rng.random() < deriv_frac else "eri"
        if kind == "eri" and jk_frac and rng.random() < jk_frac:
            kind = "jk"
        out.append(Request(i, natom, int(rng.integers(1, 2 ** 31)), kind, t))
    return out


def request_density(nbf, seed):
    """The density a "jk" request without its own uses: C C^T / nbf with C
    uniform in [-1/2, 1/2), nbf x nbf/4 (positive semidefinite, symmetric)."""
    rng = np.random.default_rng(seed)
    C = rng.uniform(-0.5, 0.5, (nbf, max(2, nbf // 4)))
    return np.ascontiguousarray(C @ C.T / nbf)


def validate(shells):
    """None if the shells can be served, else the reason they cannot."""
    if not shells:
        return "no shells"
    for k, (am, xyz, al, co) in enumerate(shells):
        if not isinstance(am, (int, np.integer)) or am < 0 or am > MAXAM:
            return "shell %d: angular momentum %r (0..%d served)" % (k, am, MAXAM)
        if len(al) == 0 or len(al) != len(co):
            return "shell %d: %d exponents, %d coefficients" % (k, len(al), len(co))
        if not all(math.isfinite(float(x)) for x in xyz):
            return "shell %d: centre not finite" % k
        if not all(math.isfinite(float(a)) and float(a) > 0.0 for a in al):
            return "shell %d: exponent not finite and positive" % k
        if not all(math.isfinite(float(c)) for c in co):
            return "shell %d: coefficient not finite" % k
    return None


def footprint(nshell, kind, chunk_doubles):
    """Estimated device bytes a request will hold before its set-up has run:
    its shell pairs (linear and tiled views, ~20 doubles per primitive pair,
    6.25 primitive pairs per pair on average) and, for "jk", the density, J
    and K (at most 6 functions per shell, d).  The output chunk buffers and
    the derivative scratch belong to the workers, not to requests."""
    npair = nshell * (nshell + 1) // 2
    mats = 3 * 8 * (6 * nshell) ** 2 if kind == "jk" else 0
    return npair * 6.25 * 20 * 8 * 2 + mats


class Server:
    def __init__(self, hbm_budget=64 << 30, slots=4, tol=1e-10, chunk_doubles=1 << 27, verbose=False, probe=0,
                 timeout=None, deriv_chunk_doubles=None):
        self.verbose = verbose
        self.last_note = 0.0
        self.probe = probe         # quartets per request whose values are kept (result checks)
        self.budget = float(hbm_budget)
        self.free = float(hbm_budget)
        self.slots = slots
        self.tol = tol
        self.chunk = chunk_doubles
        # derivative tiles are larger: every derivative call runs a class
        # kernel per (a+1)/(a-1) term and the assembly, whose launch ramps
        # and tails larger tiles amortise
        self.deriv_chunk = deriv_chunk_doubles or 4 * chunk_doubles
        self.timeout = timeout     # default per-request time limit (s after admission)
        self.cv = threading.Condition()
        self.waiting = []          # arrived, not admitted (arrival order)
        self.active = []           # admitted, tiles left or in flight
        self.setting_up = 0
        self.arrivals_done = False
        self.t0 = None
        self.max_concurrent = 0
        self.tiles_run = 0

    def now(self):
        return time.perf_counter() - self.t0

    def _shells(self, r):
        return r.shells if r.shells is not None else workloads.grid_molecule(natom=r.natom, seed=r.seed)

    @staticmethod
    def _nshell(r):
        # the recipe has six shells per atom
        return len(r.shells) if r.shells is not None else 6 * r.natom

    # ------------------------------------------------------------ scheduling
    def _next_task(self):
        """Under the lock: ('tile', r, c), ('setup', r) or None.  Requests past
        their time limit stop getting tiles and end once none is in flight;
        a request larger than the whole budget is refused."""
        now = self.now()
        live = []
        for r in list(self.active):
            lim = r.timeout if r.timeout is not None else self.timeout
            if r.error is None and lim is not None and now - r.t_admit > lim:
                r.error = "timeout after %.1f s" % (now - r.t_admit)
            if r.error is not None and r.inflight == 0:
                self._finish(r)
                self._complete(r, self.done)
            elif r.error is None and r.next < r.ss.nchunk:
                live.append(r)
        for r in list(self.waiting):
            if footprint(self._nshell(r), r.kind, self.chunk) > self.budget:
                self.waiting.remove(r)
                r.error = "rejected: footprint above the HBM budget"
                r.t_admit = r.t_done = now
                self._complete(r, self.done)
        # admission comes first (one set-up at a time): a request that fits
        # the remaining budget is set up as soon as it arrives, and its
        # tiles then compete with the running requests' by remaining work
        if self.setting_up == 0 or not live:
            for r in self.waiting:
                if footprint(self._nshell(r), r.kind, self.chunk) <= self.free:
                    self.waiting.remove(r)
                    est = footprint(self._nshell(r), r.kind, self.chunk)
                    self.free -= est
                    r.charged = est
                    r.t_admit = now
                    self.setting_up += 1
                    return ("setup", r)
        if live:
            r = min(live, key=lambda x: (x.left, x.arrival))
            c = r.next
            r.next += 1
            r.inflight += 1
            return ("tile", r, c)
        return None

    def _finish(self, r):
        """Under the lock: the request has no tile left in flight."""
        self.active.remove(r)
        self.free += r.charged
        r.t_done = self.now()

    def _setup(self, r):
        import torch
        if r.kind not in ("eri", "deriv", "jk"):
            r.error = "rejected: unknown request kind %r" % (r.kind,)
            return
        shells = self._shells(r)
        bad = validate(shells)
        if bad is not None:
            r.error = "rejected: " + bad
            return
        deriv = 1 if r.kind == "deriv" else 0
        ss = quartets.ScreenedSet(shells, tol=self.tol, out_doubles=self.deriv_chunk if deriv else self.chunk,
                                  deriv=deriv)
        r.ss = ss
        r.quartets = ss.total_quartets
        r.left = float(ss.total_bytes)
        r.sums = torch.zeros((max(ss.nchunk, 1), 2), dtype=torch.float64, device="cuda")
        r.work_need = ss.deriv_worksize() if deriv else 0
        if r.kind == "jk":
            nbf = ss.nbf()
            D = r.density if r.density is not None else request_density(nbf, r.seed)
            D = np.ascontiguousarray(D, dtype=np.float64)
            if D.shape != (nbf, nbf) or not np.all(np.isfinite(D)) or np.abs(D - D.T).max() > 1e-12 * max(
                    1.0, np.abs(D).max()):
                r.error = "rejected: density must be a finite symmetric %d x %d matrix" % (nbf, nbf)
                return
            r.jk_dev = (torch.from_numpy(D).cuda(), torch.zeros((nbf, nbf), dtype=torch.float64, device="cuda"),
                        torch.zeros((nbf, nbf), dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        if self.probe:
            for c, off, nc, q in ss.located_sample(self.probe, seed=r.seed):
                r.probes.setdefault(c, []).append((off, nc, q))

    def _tile(self, r, c, out, stream, work):
        """Run output chunk c of request r on this worker's stream and reduce
        it into r's digest slot c."""
        import torch
        ss = r.ss
        n = ss.chunk_doubles(c)
        if out[0].numel() < ss.out_doubles:
            # this worker's earlier tiles ended with a stream synchronisation
            out[0] = None
            out[0] = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
        out = out[0]
        wp = None
        if ss.deriv:
            need = r.work_need
            if work[0] is None or work[0].numel() < need:
                work[0] = torch.empty(need, dtype=torch.float64, device="cuda")
            wp = work[0].data_ptr()
        with torch.cuda.stream(stream):
            if r.jk_dev is not None:
                D, J, K = r.jk_dev
                ss.run_jk(out.data_ptr(), D.data_ptr(), J.data_ptr(), K.data_ptr(), ss.chunk_launches(c))
            else:
                ss.run_pass(out.data_ptr(), chunk=c, work_ptr=wp)
            x = out[:n]
            r.sums[c, 0] = x.sum()
            r.sums[c, 1] = torch.dot(x, x)
            kept = [(q, out[off:off + nc].clone()) for off, nc, q in r.probes.get(c, [])]
            stream.synchronize()
        return [(q, v.cpu().numpy()) for q, v in kept], sum(8.0 * ln.nout for ln in ss.chunk_launches(c))

    def _worker(self, done):
        import torch
        stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())
        out = [torch.empty(self.chunk, dtype=torch.float64, device="cuda")]   # grows for derivative tiles
        work = [None]
        while True:
            with self.cv:
                while True:
                    task = self._next_task()
                    if task is not None:
                        break
                    if self.arrivals_done and not self.waiting and not self.active and not self.setting_up:
                        return
                    self.cv.wait(timeout=0.5)
            if task[0] == "setup":
                r = task[1]
                try:
                    self._setup(r)
                except Exception as e:      # reported per request, the service goes on
                    r.error = repr(e)
                with self.cv:
                    self.setting_up -= 1
                    if r.ss is not None:
                        actual = float(r.ss.device_bytes()) + r.sums.numel() * 8.0
                        if r.jk_dev is not None:
                            actual += sum(m.numel() * 8.0 for m in r.jk_dev)
                        self.free += r.charged - actual
                        r.charged = actual
                    if r.error is None and r.ss is not None and r.ss.nchunk > 0:
                        self.active.append(r)
                        self.max_concurrent = max(self.max_concurrent, len(self.active))
                    else:
                        self.active.append(r)
                        self._finish(r)
                        self._complete(r, done)
                    self.cv.notify_all()
                continue
            _, r, c = task
            kept, nbytes = [], 0.0
            err = None
            try:
                kept, nbytes = self._tile(r, c, out, stream, work)
            except Exception as e:
                err = repr(e)
            with self.cv:
                self.tiles_run += 1
                r.inflight -= 1
                r.done_tiles += 1
                if self.verbose and self.now() - self.last_note > 30.0:
                    # a progress line while long requests run
                    self.last_note = self.now()
                    print("[serve %.1fs] %d requests running; request %d: %d of %d tiles"
                          % (self.now(), len(self.active), r.rid, r.done_tiles, r.ss.nchunk), flush=True)
                r.left -= nbytes
                r.kept.extend(kept)
                if err is not None and r.error is None:
                    r.error = err
                if r.inflight == 0 and (r.done_tiles == r.ss.nchunk or r.error is not None):
                    self._finish(r)
                    self._complete(r, done)
                self.cv.notify_all()

    def _complete(self, r, done):
        """Under the lock: digest in tile order, release the request's data."""
        if r.error is None and r.sums is not None:
            r.digest = [float(v) for v in r.sums.sum(dim=0).cpu()]
            if self.probe:
                r.samples = sorted(r.kept, key=lambda qv: qv[0])
            if r.jk_dev is not None:
                # every tile's stream was synchronised when the tile ended
                D, J, K = r.jk_dev
                simint.jk_symmetrize(J.shape[0], J.data_ptr(), K.data_ptr())
                simint.synchronize()
                r.J, r.K = J.cpu().numpy(), K.cpu().numpy()
                r.digest += [float((D * J).sum()), float((D * K).sum())]
        r.jk_dev = None
        if r.ss is not None:
            r.ss.free()
            r.ss = None
        r.sums = None
        if self.verbose:
            print("[serve %.1fs] request %d (%d atoms, %s) done: %.2f s after arrival, %d quartets%s"
                  % (r.t_done, r.rid, r.natom, r.kind, r.t_done - r.arrival, r.quartets,
                     "" if r.error is None else ", error " + r.error), file=sys.stderr, flush=True)
        done.put(r)

    # ------------------------------------------------------------------ run
    def serve(self, requests):
        """Feed `requests` at their arrival times (seconds from start) and
        serve them; returns the requests with timings filled in."""
        import torch
        simint.lib().simint_init()
        torch.cuda.init()
        done = queue.Queue()
        self.done = done
        self.t0 = time.perf_counter()
        workers = [threading.Thread(target=self._worker, args=(done,), daemon=True) for _ in range(self.slots)]
        for w in workers:
            w.start()
        for r in sorted(requests, key=lambda r: r.arrival):
            dt = r.arrival - self.now()
            if dt > 0:
                time.sleep(dt)
            with self.cv:
                self.waiting.append(r)
                self.cv.notify_all()
        with self.cv:
            self.arrivals_done = True
            self.cv.notify_all()
        for w in workers:
            w.join()
        self.wall = self.now()
        return requests


def summarize(reqs, wall):
    lat = np.array([r.t_done - r.arrival for r in reqs if r.t_done is not None and r.error is None])
    q = sum(r.quartets for r in reqs if r.error is None)
    return {"requests": len(reqs), "errors": sum(1 for r in reqs if r.error),
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "mean_latency_s": float(lat.mean()) if lat.size else None,
            "quartets": int(q), "wall_s": wall, "quartets_per_s": q / wall if wall > 0 else None,
            "deriv_requests": sum(1 for r in reqs if r.kind == "deriv")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--min-atoms", type=int, default=10)
    ap.add_argument("--max-atoms", type=int, default=200)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--timeout", type=float, default=None, help="per-request limit, seconds after admission")
    ap.add_argument("--jk-frac", type=float, default=0.0,
                    help="share of the ERI requests that ask for Coulomb and exchange matrices")
    a = ap.parse_args()
    reqs = request_stream(a.requests, a.rate, a.seed, (a.min_atoms, a.max_atoms), jk_frac=a.jk_frac)
    srv = Server(hbm_budget=a.budget_gib * (1 << 30), slots=a.slots, tol=a.tol, verbose=True, timeout=a.timeout)
    srv.serve(reqs)
    res = summarize(reqs, srv.wall)
    res["max_concurrent"] = srv.max_concurrent
    res["tiles"] = srv.tiles_run
    res["requests_detail"] = [r.to_dict() for r in reqs]
    print(json.dumps(res))


if __name__ == "__main__":
    main()

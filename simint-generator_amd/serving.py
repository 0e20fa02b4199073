"""Continuous-batching ERI service on one GPU.

Requests arrive as a Poisson stream.  Each asks for the Schwarz-screened
ERIs (or first derivatives) of one synthetic molecule built with the
64-atom-system recipe (its own atom count and seed).  The server

  * admits a request when a worker slot is free and its estimated device
    footprint (shell-pair data, an output chunk buffer, derivative scratch)
    fits the remaining HBM budget -- requests that arrive while others run
    are admitted as slots and memory free up, not after the queue drains;
  * runs admitted requests concurrently, each on its own worker thread and
    so on its own HIP stream (the library keeps one stream per host
    thread), so their kernels overlap on the device;
  * keeps each request's shell pairs resident in HBM while it runs, streams
    its integrals chunk by chunk through one output buffer, and reduces every
    chunk on the device into a digest (sum of values and sum of squares in
    FP64), the stand-in for a consumer such as a Fock build.

Latency is arrival to completion; throughput is shell quartets (derivative
quartets counted once) per second of wall time.

    python simint-generator_amd/serving.py [--requests 24] [--rate 2.0] ...
"""
import argparse
import json
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


class Request:
    def __init__(self, rid, natom, seed, kind, arrival):
        self.rid, self.natom, self.seed, self.kind, self.arrival = rid, natom, seed, kind, arrival
        self.t_admit = self.t_done = None
        self.quartets = 0
        self.digest = None
        self.samples = None        # [(quartet, values)] when the server probes results
        self.error = None

    def to_dict(self):
        return {"id": self.rid, "natom": self.natom, "seed": self.seed, "kind": self.kind,
                "arrival": self.arrival, "admit": self.t_admit, "done": self.t_done,
                "latency": None if self.t_done is None else self.t_done - self.arrival,
                "quartets": self.quartets, "digest": self.digest, "error": self.error}


def request_stream(n, rate, seed=0, natom_range=(10, 200), deriv_frac=0.3):
    """n requests: Poisson arrivals at `rate` per second, atom counts uniform
    in natom_range, 30% first-derivative requests, per-request seeds."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        natom = int(rng.integers(natom_range[0], natom_range[1] + 1))
        kind = "deriv" if rng.random() < deriv_frac else "eri"
        out.append(Request(i, natom, int(rng.integers(1, 2 ** 31)), kind, t))
    return out


def footprint(natom, kind, chunk_doubles):
    """Device bytes a request may hold: pair data (linear and tiled views,
    ~20 doubles per primitive pair, 6.25 primitive pairs per pair on
    average), the output chunk and, for derivatives, the class scratch."""
    nsh = 6 * natom
    npair = nsh * (nsh + 1) // 2
    pairs = npair * 6.25 * 20 * 8 * 2
    out = 8.0 * chunk_doubles
    work = 8.0 * chunk_doubles * 4 if kind == "deriv" else 0.0
    return pairs + out + work


class Server:
    def __init__(self, hbm_budget=64 << 30, slots=4, tol=1e-10, chunk_doubles=1 << 27, verbose=False, probe=0):
        self.verbose = verbose
        self.probe = probe         # quartets per request whose values are kept (result checks)
        self.budget = float(hbm_budget)
        self.free = float(hbm_budget)
        self.slots = slots
        self.tol = tol
        self.chunk = chunk_doubles
        self.cv = threading.Condition()
        self.waiting = []          # arrived, not admitted (arrival order)
        self.arrivals_done = False
        self.t0 = None
        self.max_concurrent = 0
        self.running = 0

    def now(self):
        return time.perf_counter() - self.t0

    # ------------------------------------------------------------ admission
    def _admissible(self):
        """First waiting request (arrival order) that fits the free budget."""
        for r in self.waiting:
            if footprint(r.natom, r.kind, self.chunk) <= self.free:
                return r
        return None

    def _worker(self, done):
        import torch
        stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())
        out = torch.empty(self.chunk, dtype=torch.float64, device="cuda")
        while True:
            with self.cv:
                while True:
                    r = self._admissible()
                    if r is not None:
                        break
                    if self.arrivals_done and not self.waiting:
                        return
                    self.cv.wait()
                self.waiting.remove(r)
                fp = footprint(r.natom, r.kind, self.chunk)
                self.free -= fp
                self.running += 1
                self.max_concurrent = max(self.max_concurrent, self.running)
                r.t_admit = self.now()
            try:
                self._process(r, out, stream)
            except Exception as e:      # reported per request, the service goes on
                r.error = repr(e)
            r.t_done = self.now()
            with self.cv:
                self.free += fp
                self.running -= 1
                self.cv.notify_all()
            if self.verbose:
                print("[serve %.1fs] request %d (%d atoms, %s) done: %.2f s after arrival, %d quartets%s"
                      % (r.t_done, r.rid, r.natom, r.kind, r.t_done - r.arrival, r.quartets,
                         "" if r.error is None else ", error " + r.error), file=sys.stderr, flush=True)
            done.put(r)

    def _process(self, r, out, stream):
        import torch
        shells = workloads.config2(natom=r.natom, seed=r.seed)
        deriv = 1 if r.kind == "deriv" else 0
        ss = quartets.ScreenedSet(shells, tol=self.tol, out_doubles=self.chunk, deriv=deriv)
        work = None
        if deriv:
            work = torch.empty(ss.deriv_worksize(), dtype=torch.float64, device="cuda")
        t_note = time.perf_counter()
        probes = {}
        if self.probe:
            for c, off, nc, q in ss.located_sample(self.probe, seed=r.seed):
                probes.setdefault(c, []).append((off, nc, q))
        kept = []
        with torch.cuda.stream(stream):
            # zeroed on the library stream, so the fill is ordered before the
            # accumulations that stream runs
            s = torch.zeros(2, dtype=torch.float64, device="cuda")
            for c in range(ss.nchunk):
                if self.verbose and time.perf_counter() - t_note > 30.0:
                    print("[serve %.1fs] request %d: chunk %d of %d" % (self.now(), r.rid, c, ss.nchunk),
                          file=sys.stderr, flush=True)
                    t_note = time.perf_counter()
                n = ss.chunk_doubles(c)
                ss.run_pass(out.data_ptr(), chunk=c, work_ptr=None if work is None else work.data_ptr())
                x = out[:n]
                s[0] += x.sum()
                s[1] += torch.dot(x, x)
                for off, nc, q in probes.get(c, []):
                    kept.append((q, out[off:off + nc].clone()))
            stream.synchronize()
        r.quartets = ss.total_quartets
        r.digest = [float(v) for v in s.cpu()]
        if self.probe:
            r.samples = [(q, v.cpu().numpy()) for q, v in kept]
        ss.free()

    # ------------------------------------------------------------------ run
    def serve(self, requests):
        """Feed `requests` at their arrival times (seconds from start) and
        serve them; returns the requests with timings filled in."""
        import torch
        simint.lib().simint_init()
        torch.cuda.init()
        done = queue.Queue()
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
    a = ap.parse_args()
    reqs = request_stream(a.requests, a.rate, a.seed, (a.min_atoms, a.max_atoms))
    srv = Server(hbm_budget=a.budget_gib * (1 << 30), slots=a.slots, tol=a.tol, verbose=True)
    srv.serve(reqs)
    res = summarize(reqs, srv.wall)
    res["max_concurrent"] = srv.max_concurrent
    res["requests_detail"] = [r.to_dict() for r in reqs]
    print(json.dumps(res))


if __name__ == "__main__":
    main()

"""Continuous-batching solve server over the mixed-precision engine, and the plain FP64
one-at-a-time path it is measured against.

A request is a preprocessed mesh plus flow parameters. The server admits requests into the
engine as they arrive (from worker threads, so that planning and upload overlap the running
iterations), keeps each case on the device until it converges (residual drop) or reaches the
iteration cap, retires finished cases without draining the others and admits waiting requests
as slots and HBM free up. The engine stops a converged case on the device by itself, so the
polling interval changes latency only, never a result.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import native as N
from .solver import Batch, EulerSolver


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else float("nan")


class Server:
    def __init__(self, device: int = 0, hbm_budget: float = 200e9, max_cases: int = 256, drop: float = 1e-6,
                 max_iter: int = 5000, admit_workers: int = 4, tuning: tuple | None = None, poll_ms: float = 20.0,
                 batch=None):
        """batch: an engine with Batch's add/run/poll/result/remove methods (default: a new Batch)"""
        kw = {}
        if tuning is not None:
            kw = dict(wg_per_cu=tuning[0], xcds=tuning[1], nontemporal=tuning[2])
        self.batch = batch if batch is not None else Batch(device=device, hbm_budget=hbm_budget,
                                                           max_cases=max_cases, **kw)
        self.drop, self.max_iter = drop, max_iter
        self.admit_workers = admit_workers
        self.poll_ms = poll_ms

    def close(self):
        self.batch.close()

    def serve(self, requests: list, arrivals) -> list[dict]:
        """requests: list of (Mesh, Params, tag); arrivals: seconds from the start. Returns one record
        per request: arrival, admitted, finished (s), outcome (converged / capped / failed /
        rejected), iterations, rms drop, CL, CD, tag. A failed case (non-finite residual) and a
        request that can never fit the HBM budget are reported as such, not as solves."""
        n = len(requests)
        arrivals = np.asarray(arrivals, dtype=np.float64)
        rec = [dict(idx=i, tag=requests[i][2], arrival=float(arrivals[i])) for i in range(n)]
        lock = threading.Lock()
        waiting = []          # arrived, not admitted (engine full)
        rejected = []         # can never be admitted
        id2req = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731

        def admit(i, alone=False):
            m, p, _ = requests[i]
            try:
                cid = self.batch.add(m, p, drop=self.drop, max_iter=self.max_iter)
            except N.E3DError as e:
                with lock:
                    if e.code == N.E3D_ERR_BUSY and not alone:
                        waiting.append(i)
                    else:
                        # larger than the whole budget, or busy with nothing else resident
                        rec[i].update(finished=now(), outcome="rejected", error=str(e))
                        rejected.append(i)
                return
            with lock:
                id2req[cid] = i
                rec[i]["admitted"] = now()

        pool = ThreadPoolExecutor(self.admit_workers)
        futures = []
        fed = threading.Event()

        def feeder():
            for i in range(n):
                dt = arrivals[i] - now()
                if dt > 0:
                    time.sleep(dt)
                with lock:
                    futures.append(pool.submit(admit, i))
            fed.set()

        ft = threading.Thread(target=feeder, daemon=True)
        ft.start()
        finished = 0
        freed = False
        iters_per_run = 10
        while finished < n:
            with lock:
                resident = len(id2req)
            if resident == 0:
                time.sleep(0.0005)
            else:
                t_run = time.perf_counter()
                self.batch.run(iters_per_run)
            with lock:
                pending = list(futures)
            for f in pending:
                if f.done():
                    f.result()  # re-raise unexpected admission errors
                    with lock:
                        futures.remove(f)
            st = self.batch.poll()
            if resident:
                # size the next run to about poll_ms of device time
                el = (time.perf_counter() - t_run) * 1e3
                iters_per_run = int(min(500, max(5, iters_per_run * self.poll_ms / max(el, 1e-3))))
            for s in st:
                if not s["done"]:
                    continue
                with lock:
                    i = id2req.pop(s["id"], None)
                if i is None:
                    continue
                r = self.batch.result(s["id"])
                self.batch.remove(s["id"])
                h = r["rms"]
                rec[i].update(finished=now(), outcome=s["outcome"], iterations=s["iterations"], CL=float(r["CL"]),
                              CD=float(r["CD"]),
                              drop=float(h[-1] / h[0]) if len(h) and h[0] > 0 else float("nan"))
                finished += 1
                freed = True
            with lock:
                finished += len(rejected)
                rejected.clear()
                in_flight = len(futures)
                empty = not id2req
                # retry the waiting requests once the admissions in flight have settled, after a
                # retirement freed room, or when nothing is resident at all
                if waiting and in_flight == 0 and (freed or empty):
                    if empty:
                        # into an empty engine a request must fit, or it never will: try the first
                        # alone, the others once it has settled
                        futures.append(pool.submit(admit, waiting.pop(0), True))
                        freed = True
                    else:
                        retry, waiting[:] = list(waiting), []
                        freed = False
                        for j in retry:
                            futures.append(pool.submit(admit, j))
        ft.join()
        pool.shutdown()
        return rec


def plain_solve(mesh, params, drop: float = 1e-6, max_iter: int = 5000, chunk: int = 50) -> dict:
    """One case on the plain FP64 path (e3d_create + RK5 iterations until the residual drop or the
    cap); wall time includes plan and upload. Iterations run in chunks of at most `chunk`, sized
    from the observed convergence rate so that the chunks shrink to single iterations near the
    target; iterations, CL, CD and the drop are all those of the state the solve stops at."""
    t0 = time.perf_counter()
    n_elem = mesh.info()["n_elem"]
    with EulerSolver(mesh, params) as s:
        rms0 = None
        it = 0
        k = 1
        converged = False
        last = float("nan")
        while it < max_iter and not converged:
            k = max(1, min(k, chunk, max_iter - it))
            rms = np.sqrt(s.iterate(k) / n_elem)
            it += k
            if rms0 is None:
                rms0 = float(rms[0])
            last = float(rms[-1])
            target = drop * rms0
            converged = drop > 0 and bool(np.any(rms < target))
            if not np.isfinite(last):
                break
            # remaining iterations at the rate of this chunk; the next chunk takes half of them
            rate = (last / float(rms[0])) ** (1.0 / max(k - 1, 1)) if k > 1 and rms[0] > 0 else 0.0
            if 0.0 < rate < 1.0 and last > target > 0:
                k = int(0.5 * np.log(target / last) / np.log(rate))
            else:
                k = chunk
        cf, cm = s.aero()
    cl = cf[params.orient_cl[0]] * params.orient_cl[1]
    cd = cf[params.orient_cd[0]] * params.orient_cd[1]
    outcome = "converged" if converged else ("failed" if not np.isfinite(last) else "capped")
    return {"service": time.perf_counter() - t0, "iterations": it, "CL": float(cl), "CD": float(cd),
            "drop": float(last / rms0) if rms0 else float("nan"), "outcome": outcome}


def fifo_latencies(arrivals, services) -> tuple[np.ndarray, float]:
    """Latency of each request and the makespan for one server taking requests in arrival order."""
    t = 0.0
    lat = []
    for a, s in zip(arrivals, services):
        t = max(t, a) + s
        lat.append(t - a)
    return np.asarray(lat), t - float(arrivals[0])


def summarize(rec: list[dict]) -> dict:
    """Latency percentiles and solves/h over the requests that were solved (converged or stopped at
    the iteration cap), with the count of each outcome."""
    ok = [r for r in rec if r.get("outcome", "converged") in ("converged", "capped")]
    lat = [r["finished"] - r["arrival"] for r in ok]
    span = max(r["finished"] for r in rec) - min(r["arrival"] for r in rec)
    counts = {k: sum(1 for r in rec if r.get("outcome", "converged") == k)
              for k in ("converged", "capped", "failed", "rejected")}
    return {"requests": len(rec), "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99),
            "makespan_s": span, "solves_per_h": 3600.0 * len(ok) / span, **counts}

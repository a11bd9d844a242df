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
                 max_iter: int = 5000, admit_workers: int = 4, tuning: tuple | None = None, poll_ms: float = 20.0):
        kw = {}
        if tuning is not None:
            kw = dict(wg_per_cu=tuning[0], xcds=tuning[1], nontemporal=tuning[2])
        self.batch = Batch(device=device, hbm_budget=hbm_budget, max_cases=max_cases, **kw)
        self.drop, self.max_iter = drop, max_iter
        self.admit_workers = admit_workers
        self.poll_ms = poll_ms

    def close(self):
        self.batch.close()

    def serve(self, requests: list, arrivals) -> list[dict]:
        """requests: list of (Mesh, Params, tag); arrivals: seconds from the start. Returns one record
        per request: arrival, admitted, finished (s), iterations, rms drop, CL, CD, tag."""
        n = len(requests)
        arrivals = np.asarray(arrivals, dtype=np.float64)
        rec = [dict(idx=i, tag=requests[i][2], arrival=float(arrivals[i])) for i in range(n)]
        lock = threading.Lock()
        waiting = []          # arrived, not admitted (engine full)
        id2req = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731

        def admit(i):
            m, p, _ = requests[i]
            try:
                cid = self.batch.add(m, p, drop=self.drop, max_iter=self.max_iter)
            except N.E3DError as e:
                if e.code != N.E3D_ERR_BUSY:
                    raise
                with lock:
                    waiting.append(i)
                return
            with lock:
                id2req[cid] = i
                rec[i]["admitted"] = now()

        pool = ThreadPoolExecutor(self.admit_workers)
        futures = []

        def feeder():
            for i in range(n):
                dt = arrivals[i] - now()
                if dt > 0:
                    time.sleep(dt)
                futures.append(pool.submit(admit, i))

        ft = threading.Thread(target=feeder, daemon=True)
        ft.start()
        finished = 0
        iters_per_run = 10
        while finished < n:
            with lock:
                resident = len(id2req)
            if resident == 0:
                time.sleep(0.0005)
            else:
                t_run = time.perf_counter()
                self.batch.run(iters_per_run)
            for f in list(futures):
                if f.done():
                    f.result()  # re-raise admission errors
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
                rec[i].update(finished=now(), iterations=s["iterations"], CL=float(r["CL"]), CD=float(r["CD"]),
                              drop=float(h[-1] / h[0]) if len(h) and h[0] > 0 else float("nan"))
                finished += 1
                with lock:
                    retry, waiting[:] = list(waiting), []
                for j in retry:
                    futures.append(pool.submit(admit, j))
        ft.join()
        pool.shutdown()
        return rec


def plain_solve(mesh, params, drop: float = 1e-6, max_iter: int = 5000, chunk: int = 50) -> dict:
    """One case on the plain FP64 path (e3d_create + RK5 iterations until the residual drop or the
    cap, checked every `chunk` iterations); wall time includes plan and upload."""
    t0 = time.perf_counter()
    n_elem = mesh.info()["n_elem"]
    with EulerSolver(mesh, params) as s:
        rms0 = None
        it = 0
        done = False
        while it < max_iter and not done:
            k = min(chunk, max_iter - it)
            rms = np.sqrt(s.iterate(k) / n_elem)
            if rms0 is None:
                rms0 = rms[0]
            hit = np.nonzero(rms < drop * rms0)[0] if drop > 0 else []
            if len(hit):
                it += int(hit[0]) + 1
                done = True
            else:
                it += k
        cf, cm = s.aero()
        last = rms[-1]
    cl = cf[params.orient_cl[0]] * params.orient_cl[1]
    cd = cf[params.orient_cd[0]] * params.orient_cd[1]
    return {"service": time.perf_counter() - t0, "iterations": it, "CL": float(cl), "CD": float(cd),
            "drop": float(last / rms0) if rms0 else float("nan")}


def fifo_latencies(arrivals, services) -> tuple[np.ndarray, float]:
    """Latency of each request and the makespan for one server taking requests in arrival order."""
    t = 0.0
    lat = []
    for a, s in zip(arrivals, services):
        t = max(t, a) + s
        lat.append(t - a)
    return np.asarray(lat), t - float(arrivals[0])


def summarize(rec: list[dict]) -> dict:
    lat = [r["finished"] - r["arrival"] for r in rec]
    span = max(r["finished"] for r in rec) - min(r["arrival"] for r in rec)
    return {"requests": len(rec), "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99),
            "makespan_s": span, "solves_per_h": 3600.0 * len(rec) / span}

"""Continuous-batching solve server over the mixed-precision engine, and the plain FP64
one-at-a-time path it is measured against.

A request is a preprocessed mesh plus flow parameters. The server admits requests into the
engine as they arrive (from worker threads, so that planning and upload overlap the running
iterations), keeps each case on the device until it converges (residual drop) or reaches the
iteration cap, retires finished cases without draining the others and admits waiting requests
as slots and HBM free up. The engine stops a converged case on the device by itself, so the
polling interval changes latency only, never a result.

A case whose residual flattens at the floor of the engine's FP32 fluxes before its drop (the
engine's stall test) is finished on the FP64 path: its state is read back, the case leaves the
engine, and an FP64 solver continues from that state on its own stream, beside the engine, until
the drop or the iteration cap (counting the engine's iterations).
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import native as N
from .solver import Batch, EulerSolver


# the server's stall test: windows of 1000 iterations; stalled when a window's lowest residual is
# not below 0.9 x the previous window's while below 1e-3 x the first residual
STALL_TEST = (1000, 0.9, 1e-3)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else float("nan")


class Server:
    def __init__(self, device: int = 0, hbm_budget: float = 200e9, max_cases: int = 256, drop: float = 1e-6,
                 max_iter: int = 5000, admit_workers: int = 4, tuning: tuple | None = None, poll_ms: float = 20.0,
                 batch=None, stall: tuple | None = STALL_TEST, finish_workers: int = 2, finisher=None):
        """batch: an engine with Batch's add/run/poll/status/result/get_state/remove methods (default: a
        new Batch). stall: the engine's stall test (window, ratio, level), None: off. finisher(mesh,
        params, state, rms0, iterations, drop, max_iter) -> dict finishes a stalled case (default:
        finish_fp64 on the same device)."""
        kw = {}
        if tuning is not None:
            kw = dict(wg_per_cu=tuning[0], xcds=tuning[1], nontemporal=tuning[2])
        self.batch = batch if batch is not None else Batch(device=device, hbm_budget=hbm_budget,
                                                           max_cases=max_cases, **kw)
        self.drop, self.max_iter = drop, max_iter
        self.admit_workers = admit_workers
        self.poll_ms = poll_ms
        self.stall = stall
        self.finish_workers = finish_workers
        self.finisher = finisher if finisher is not None else (
            lambda m, p, st, r0, it, drop, cap: finish_fp64(m, p, st, r0, it, drop, cap, device=device))

    def close(self):
        self.batch.close()

    def serve(self, requests: list, arrivals) -> list[dict]:
        """requests: list of (Mesh, Params, tag); arrivals: seconds from the start. Returns one record
        per request: arrival, admitted, finished (s), outcome (converged / capped / failed /
        rejected), iterations, rms drop, CL, CD, tag; a case finished on the FP64 path also has
        stalled_at (engine iterations) and fp64_iterations. A failed case (non-finite residual) and
        a request that can never fit the HBM budget are reported as such, not as solves."""
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
                cid = self.batch.add(m, p, drop=self.drop, max_iter=self.max_iter, stall=self.stall)
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
        fpool = ThreadPoolExecutor(self.finish_workers)
        futures = []
        finishing = []        # (request, future) of stalled cases on the FP64 path
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
            for i, f in list(finishing):
                if f.done():
                    fin = f.result()  # re-raise unexpected errors of the FP64 path
                    finishing.remove((i, f))
                    rec[i].update(finished=now(), outcome=fin["outcome"], iterations=fin["iterations"],
                                  fp64_iterations=fin["fp64_iterations"], CL=fin["CL"], CD=fin["CD"],
                                  drop=fin["drop"])
                    finished += 1
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
                if s["outcome"] == "stalled":
                    m, p, _ = requests[i]
                    rms0 = self.batch.status(s["id"])["rms0"]
                    state = self.batch.get_state(s["id"])
                    self.batch.remove(s["id"])
                    rec[i].update(stalled_at=s["iterations"])
                    finishing.append((i, fpool.submit(self.finisher, m, p, state, rms0, s["iterations"], self.drop,
                                                      self.max_iter)))
                    freed = True
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
        fpool.shutdown()
        return rec


def _drive(s, n_elem: int, drop: float, max_iter: int, it: int = 0, rms0: float | None = None,
           chunk: int = 50) -> dict:
    """RK5 iterations of an FP64 solver until rms < drop x rms0 or `max_iter` iterations in all (`it`
    done before). Iterations run in chunks of at most `chunk`, sized from the observed convergence
    rate so that the chunks shrink to single iterations near the target; iterations, CL, CD and the
    drop are all those of the state the solve stops at. rms0 None: the first iteration's rms."""
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
    outcome = "converged" if converged else ("failed" if not np.isfinite(last) else "capped")
    return {"iterations": it, "outcome": outcome, "drop": float(last / rms0) if rms0 else float("nan")}


def _coeffs(s, params) -> tuple[float, float]:
    cf, cm = s.aero()
    return float(cf[params.orient_cl[0]] * params.orient_cl[1]), float(cf[params.orient_cd[0]] * params.orient_cd[1])


def plain_solve(mesh, params, drop: float = 1e-6, max_iter: int = 5000, chunk: int = 50) -> dict:
    """One case on the plain FP64 path (e3d_create + RK5 iterations until the residual drop or the
    cap); wall time includes plan and upload."""
    t0 = time.perf_counter()
    with EulerSolver(mesh, params) as s:
        r = _drive(s, mesh.info()["n_elem"], drop, max_iter, chunk=chunk)
        cl, cd = _coeffs(s, params)
    return {"service": time.perf_counter() - t0, "CL": cl, "CD": cd, **r}


def finish_fp64(mesh, params, state: dict, rms0: float, iterations: int, drop: float, max_iter: int,
                device: int = 0, chunk: int = 50) -> dict:
    """Finishing phase of a case the mixed-precision engine retired as stalled: an FP64 solver
    (e3d_create) starts from the engine's state (e3d_set_state) and iterates until rms < drop x
    rms0, rms0 being the first iteration's rms in the engine, or until max_iter iterations counting
    the engine's `iterations`. Returns outcome, total iterations, fp64_iterations, drop, CL, CD."""
    t0 = time.perf_counter()
    with EulerSolver(mesh, params, device=device) as s:
        s.set_state(state)
        r = _drive(s, mesh.info()["n_elem"], drop, max_iter, it=iterations, rms0=rms0, chunk=chunk)
        cl, cd = _coeffs(s, params)
    return {"service": time.perf_counter() - t0, "CL": cl, "CD": cd, "fp64_iterations": r["iterations"] - iterations,
            **r}


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
            "makespan_s": span, "solves_per_h": 3600.0 * len(ok) / span, **counts,
            "finished_fp64": sum(1 for r in rec if "stalled_at" in r)}

"""The serving loop's host logic (admission, waiting list, retirement, outcomes) on the CPU, with a
scripted engine in place of the device batch: every request ends in a record, whether it solves,
fails, or can never fit the HBM budget, and the loop never waits on a request nobody will retry."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from euler3d_amd import native as N
from euler3d_amd.serve import Server, summarize


class ScriptedBatch:
    """Engine stand-in: a request 'mesh' is a dict(bytes=, iters=, nan=, stall=); a case converges
    after `iters` iterations, fails at its first iteration when nan is set, or stalls after `stall`
    iterations when the stall test is on."""

    def __init__(self, budget: float, slots: int, delay: float = 0.0):
        self.budget, self.slots, self.delay = budget, slots, delay
        self.cases = {}
        self.used = 0.0
        self.next_id = 0
        self.lock = threading.Lock()
        self.max_resident = 0

    def add(self, m, p, drop=1e-6, max_iter=5000, stall=None):
        if self.delay:
            threading.Event().wait(self.delay)
        with self.lock:
            if m["bytes"] > self.budget:
                raise N.E3DError(N.E3D_ERR_ARG, "larger than the whole budget")
            if len(self.cases) >= self.slots or self.used + m["bytes"] > self.budget:
                raise N.E3DError(N.E3D_ERR_BUSY, "full")
            cid = self.next_id
            self.next_id += 1
            self.cases[cid] = dict(m=m, it=0, done=0, cap=max_iter, stall=m.get("stall") if stall else None)
            self.used += m["bytes"]
            self.max_resident = max(self.max_resident, len(self.cases))
            return cid

    def run(self, n):
        with self.lock:
            for c in self.cases.values():
                for _ in range(n):
                    if c["done"]:
                        break
                    c["it"] += 1
                    if c["m"].get("nan"):
                        c["done"] = 3
                    elif c["it"] >= c["m"]["iters"]:
                        c["done"] = 1
                    elif c["it"] >= c["cap"]:
                        c["done"] = 2
                    elif c["stall"] is not None and c["it"] >= c["stall"]:
                        c["done"] = 4

    def poll(self):
        with self.lock:
            return [{"id": k, "done": bool(c["done"]), "outcome": N.CASE_OUTCOMES[c["done"]], "iterations": c["it"]}
                    for k, c in self.cases.items()]

    def status(self, cid):
        c = self.cases[cid]
        return {"done": bool(c["done"]), "outcome": N.CASE_OUTCOMES[c["done"]], "iterations": c["it"], "rms0": 1.0}

    def get_state(self, cid):
        return {"it": self.cases[cid]["it"]}

    def result(self, cid):
        c = self.cases[cid]
        bad = c["done"] == 3
        return {"rms": np.array([1.0, float("nan") if bad else 1e-7]), "CL": float("nan") if bad else 0.3,
                "CD": float("nan") if bad else 0.01, "failed": bad}

    def remove(self, cid):
        with self.lock:
            c = self.cases.pop(cid)
            self.used -= c["m"]["bytes"]


def scripted_finisher(m, p, state, rms0, iterations, drop, max_iter):
    """FP64 path stand-in: converges `m["fp64"]` iterations after the engine's state, within the cap"""
    assert state == {"it": iterations} and rms0 == 1.0
    threading.Event().wait(0.002)
    it = min(iterations + m["fp64"], max_iter)
    return {"outcome": "converged" if iterations + m["fp64"] <= max_iter else "capped", "iterations": it,
            "fp64_iterations": it - iterations, "CL": 0.25, "CD": 0.02, "drop": 1e-7}


def serve(reqs, arrivals, budget=10.0, slots=4, delay=0.0, max_iter=1000, stall=(100, 0.9, 1e-3)):
    eng = ScriptedBatch(budget, slots, delay)
    srv = Server(batch=eng, max_iter=max_iter, admit_workers=3, poll_ms=1.0, stall=stall, finisher=scripted_finisher)
    rec = srv.serve([(m, None, "t") for m in reqs], arrivals)
    return rec, eng


@pytest.mark.timeout(60)
def test_every_request_ends_with_an_outcome():
    reqs = [dict(bytes=3.0, iters=40), dict(bytes=50.0, iters=10), dict(bytes=2.0, iters=5, nan=True),
            dict(bytes=4.0, iters=3000), dict(bytes=3.0, iters=60)]
    rec, eng = serve(reqs, [0.0, 0.0, 0.001, 0.002, 0.003])
    out = [r["outcome"] for r in rec]
    assert out == ["converged", "rejected", "failed", "capped", "converged"], out
    assert rec[0]["iterations"] == 40 and rec[3]["iterations"] == 1000
    assert "whole budget" in rec[1]["error"]
    s = summarize(rec)
    assert (s["converged"], s["capped"], s["failed"], s["rejected"]) == (2, 1, 1, 1)
    assert not eng.cases and eng.used == 0.0


@pytest.mark.timeout(60)
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("delay", [0.0, 0.003])
def test_waiting_requests_are_backfilled(delay, seed):
    """More requests than slots and budget at once, with slow admissions (so a request can be
    refused after the last retirement took its retry list): all are served, never more resident
    than fit."""
    rng = np.random.default_rng(seed)
    reqs = [dict(bytes=float(rng.uniform(1.0, 4.0)), iters=int(rng.integers(5, 80))) for _ in range(30)]
    rec, eng = serve(reqs, np.zeros(30), budget=8.0, slots=3, delay=delay)
    assert all(r["outcome"] == "converged" for r in rec)
    assert [r["iterations"] for r in rec] == [q["iters"] for q in reqs]
    assert eng.max_resident <= 3 and not eng.cases


@pytest.mark.timeout(60)
def test_stalled_cases_finish_on_the_fp64_path():
    """A stalled case leaves the engine (its slot and bytes free up for a waiting request) and its
    record is the finisher's: outcome, total iterations counting the engine's, fp64_iterations."""
    reqs = [dict(bytes=5.0, iters=900, stall=300, fp64=50), dict(bytes=5.0, iters=40),
            dict(bytes=5.0, iters=30), dict(bytes=4.0, iters=2000, stall=200, fp64=5000)]
    rec, eng = serve(reqs, [0.0, 0.0, 0.0, 0.0], budget=10.0, slots=2)
    assert [r["outcome"] for r in rec] == ["converged", "converged", "converged", "capped"]
    assert (rec[0]["stalled_at"], rec[0]["iterations"], rec[0]["fp64_iterations"]) == (300, 350, 50)
    assert (rec[3]["stalled_at"], rec[3]["iterations"], rec[3]["fp64_iterations"]) == (200, 1000, 800)
    assert "stalled_at" not in rec[1] and rec[1]["iterations"] == 40
    assert not eng.cases and eng.used == 0.0
    # with the stall test off the engine keeps the cases to the cap
    rec, _ = serve(reqs, [0.0, 0.0, 0.0, 0.0], budget=10.0, slots=2, stall=None)
    assert [r["outcome"] for r in rec] == ["converged", "converged", "converged", "capped"]
    assert rec[0]["iterations"] == 900 and "stalled_at" not in rec[0]

"""The mixed-precision engine's stall test and the FP64 finishing phase.

The engine's residual flattens where the FP32 rounding of the flux perturbations dominates, 5.5 to
6.5 orders below the first residual on the served workloads' cases; the FP64 path goes on falling.
A case asked for a deeper drop than that floor:
* without the stall test stops at its iteration cap, its residual never below the target;
* with it, retires as stalled at exactly the iteration the rule gives on its own residual history
  (replayed here on the host);
* finished on the FP64 path from the engine's state, converges to the target within the cap, with
  CL and CD within 1e-6 of the plain FP64 solve to the same drop;
* through the server, ends "converged" with the engine's and the FP64 path's iterations recorded.
Cases: two of the served request mix (seed 0): a 17,496-cell bump channel and a 29,792-cell wing.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from euler3d_amd import Batch
from euler3d_amd.serve import Server, finish_fp64, plain_solve
from euler3d_amd.workloads import request_mix

pytestmark = pytest.mark.gpu

STALL = (500, 0.9, 1e-3)
CASES = {"bump": (21, 1e-8, 12000), "wing": (15, 1e-7, 16000)}  # request index, drop, iteration cap


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def stall_iteration(h, window, ratio, level, drop):
    """The stall rule on a residual history: iterations done when it fires (None: never), or -1 when
    the drop is reached first."""
    r0 = h[0]
    start, prev, wmin = 0, 0.0, 0.0
    for it, r in enumerate(h):
        if r < drop * r0:
            return -1
        wmin = r if it == start else min(wmin, r)
        if it + 1 - start == window:
            if prev > 0.0 and not (wmin < ratio * prev) and wmin < level * r0:
                return it + 1
            prev, start = wmin, it + 1
    return None


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    idx, drop, cap = CASES[request.param]
    c = request_mix(24, seed=0, max_cells=2e6)[idx]
    m, p = c.build()
    say(f"[gpu] finishing case {request.param}: {m.info()['n_elem']} cells, drop {drop}, cap {cap}")
    return request.param, m, p, drop, cap


@pytest.mark.timeout(300)
def test_gpu_stall_then_fp64_finish(case):
    kind, m, p, drop, cap = case
    with Batch() as b:
        i = b.add(m, p, drop=drop, max_iter=cap)
        b.run(cap)
        st0 = b.status(i)
        h0 = b.result(i)["rms"]
    say(f"[gpu] {kind}: no stall test: {st0['outcome']} at {st0['iterations']}, "
        f"lowest drop {np.min(h0) / h0[0]:.2e}")
    assert st0["outcome"] == "capped" and st0["iterations"] == cap
    assert np.min(h0) >= drop * h0[0]

    with Batch() as b:
        i = b.add(m, p, drop=drop, max_iter=cap, stall=STALL)
        b.run(cap)
        st = b.status(i)
        h = b.result(i)["rms"]
        state = b.get_state(i)
    it = st["iterations"]
    say(f"[gpu] {kind}: stall test: {st['outcome']} at {it}")
    assert st["outcome"] == "stalled"
    # the same iterations as without the test up to the stall, and the rule fires exactly there
    assert np.array_equal(h.view(np.uint64), h0[:it].view(np.uint64))
    assert stall_iteration(h0, *STALL, drop) == it

    fin = finish_fp64(m, p, state, st["rms0"], it, drop, cap)
    plain = plain_solve(m, p, drop=drop, max_iter=cap)
    say(f"[gpu] {kind}: FP64 finish {fin['outcome']} after {fin['fp64_iterations']} more iterations "
        f"(total {fin['iterations']}, drop {fin['drop']:.2e}); plain FP64 {plain['outcome']} at "
        f"{plain['iterations']}; dCL {fin['CL'] - plain['CL']:.2e} dCD {fin['CD'] - plain['CD']:.2e}")
    assert fin["outcome"] == "converged" and fin["drop"] < drop and fin["iterations"] <= cap
    assert fin["fp64_iterations"] == fin["iterations"] - it > 0
    assert plain["outcome"] == "converged"
    assert abs(fin["CL"] - plain["CL"]) <= 1e-6 and abs(fin["CD"] - plain["CD"]) <= 1e-6


@pytest.mark.timeout(300)
def test_gpu_server_finishes_stalled_cases(case):
    kind, m, p, drop, cap = case
    srv = Server(hbm_budget=20e9, max_cases=4, drop=drop, max_iter=cap, stall=STALL, poll_ms=5.0)
    try:
        rec = srv.serve([(m, p, kind)], [0.0])
    finally:
        srv.close()
    r = rec[0]
    say(f"[gpu] {kind}: served {r['outcome']}, stalled at {r.get('stalled_at')}, "
        f"{r.get('fp64_iterations')} FP64 iterations")
    assert r["outcome"] == "converged" and r["stalled_at"] > 0 and r["fp64_iterations"] > 0
    assert r["iterations"] == r["stalled_at"] + r["fp64_iterations"] <= cap
    assert r["drop"] < drop

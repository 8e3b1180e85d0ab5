"""The served-stream workload and its load generator: Poisson arrivals of walk
requests into :class:`~dreams_rns.server.WalkServer`.

Requests follow the serve5 generator (:func:`dreams_rns.synth.serve5_request`):
matrix order m uniform in {2, 3, 4} (2x2 PCFs of the pcf2 shape, 3x3 / 4x4
CMFs of 10-400 ops), depth log-uniform in [1000, 100000], K uniform in
{32, 64, 128}. Compiling a CMF takes longer than walking it, so the programs
come from a seeded pool drawn with that generator and each request draws
its program from the pool and its depth and K afresh. A request's latency is
the server's own clock from submission to the CRT integers of p and q in
host memory.
"""

from __future__ import annotations

import math
import time
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from . import synth
from .engine import WalkSpec, cmf_spec, pcf_spec

DEPTH_RANGE = (1000, 100000)
K_CHOICES = (32, 64, 128)


def _compile_cmf(mat, m, dim, axes, dirs):
    from .compiler import compile_cmf_from_dict

    return compile_cmf_from_dict(mat, m, dim, axes, dirs)


def request_spec(req: Dict[str, Any]) -> WalkSpec:
    """The walk of one serve5 request: a PCF from A = [[1, a(0)], [0, 1]]
    with shift 1, or a CMF from the identity at its shift."""
    if req["kind"] == "pcf":
        from .runner import compile_pcf, pcf_initial_values

        prog = req.get("program") or compile_pcf(req["a"], req["b"])
        a0 = req["a0"] if "a0" in req else pcf_initial_values(req["a"])
        return pcf_spec(prog, a0, req["depth"], req["K"], 1, target=req["L"])
    return cmf_spec(req["program"], req["depth"], req["K"], req["shift"], target=req["L"])


def program_pool(n_per_order: int = 16, seed: int = synth.SEED_BASE + 5) -> Dict[int, List[Dict[str, Any]]]:
    """Seeded requests of each matrix order from serve5_request, with their
    programs compiled: {2: [...], 3: [...], 4: [...]}."""
    from .runner import compile_pcf, pcf_initial_values

    rng = np.random.default_rng(seed)
    pool: Dict[int, List[Dict[str, Any]]] = {2: [], 3: [], 4: []}
    while min(len(v) for v in pool.values()) < n_per_order:
        req = synth.serve5_request(rng, _compile_cmf)
        m = 2 if req["kind"] == "pcf" else int(req["m"])
        if len(pool[m]) >= n_per_order:
            continue
        if req["kind"] == "pcf":
            req["program"] = compile_pcf(req["a"], req["b"])
            req["a0"] = pcf_initial_values(req["a"])
        pool[m].append(req)
    return pool


def draw_request(rng, pool: Dict[int, List[Dict[str, Any]]]) -> Dict[str, Any]:
    """One request: m uniform in {2, 3, 4}, a program of that order from the
    pool, depth log-uniform in DEPTH_RANGE, K uniform in K_CHOICES."""
    m = int(rng.choice([2, 3, 4]))
    base = pool[m][int(rng.integers(0, len(pool[m])))]
    req = dict(base)
    req["depth"] = int(round(math.exp(rng.uniform(math.log(DEPTH_RANGE[0]), math.log(DEPTH_RANGE[1])))))
    req["K"] = int(rng.choice(K_CHOICES))
    req["L"] = float(rng.uniform(-5, 5))
    return req


def run_stream(server, specs: Sequence[WalkSpec], arrivals: Sequence[float], timeout: float = 300.0,
               poll_every: float = 0.002) -> Dict[str, Any]:
    """Submit specs[i] at arrivals[i] seconds (Poisson arrival times), poll
    while waiting, then drain. Returns the per-request results (in ticket
    order) and the stream's statistics: latency p50 / p99 (ms), served
    residue-steps/s over the window from the first submission to the last
    result, and the offered load."""
    results: Dict[int, Dict[str, Any]] = {}
    tickets: List[int] = []
    t0 = time.perf_counter()
    next_poll = 0.0
    late = 0.0
    for t, spec in zip(arrivals, specs):
        while True:
            now = time.perf_counter() - t0
            if now >= t:
                break
            if now >= next_poll:
                for r in server.poll():
                    results[r["ticket"]] = r
                next_poll = now + poll_every
            else:
                time.sleep(min(t - now, next_poll - now, 0.0005))
        late = max(late, time.perf_counter() - t0 - t)
        tickets.append(server.submit(spec))
    t_sub = time.perf_counter() - t0
    deadline = time.perf_counter() + timeout
    while len(results) < len(tickets):
        got = server.poll()
        for r in got:
            results[r["ticket"]] = r
        if not got:
            if time.perf_counter() > deadline:
                raise TimeoutError(f"{len(tickets) - len(results)} served requests still running")
            time.sleep(0.0005)
    ordered = [results[t] for t in tickets]
    ok = [r for r in ordered if r["status"] == 0]
    lat = np.array([r["latency_s"] for r in ok]) if ok else np.zeros(0)
    work = float(sum(s.depth * s.K for s in specs))
    served = float(sum(r["spec"].depth * r["spec"].K for r in ok))
    if ok:
        span = max(r["t_done"] for r in ok) - min(r["t_submit"] for r in ordered)
    else:
        span = float("nan")
    duration = float(arrivals[-1]) if len(arrivals) else 0.0
    stats = {
        "requests": len(ordered), "ok": len(ok),
        "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "served_residue_steps_per_s": served / span if ok and span > 0 else None,
        "offered_residue_steps_per_s": work / duration if duration > 0 else None,
        "window_s": span, "submit_s": t_sub, "max_submit_lag_ms": late * 1e3,
        "mean_residue_steps_per_request": work / len(specs) if len(specs) else None,
    }
    return {"results": ordered, "stats": stats}

"""The served stream (Poisson arrivals of serve5 requests: m in {2, 3, 4},
depth log-uniform in [1000, 100000], K in {32, 64, 128}) through the
continuous-batching server: every result -- residues, CRT integers of p and
q, FP64 shadow bits and delta_float -- equals the same walk run alone in its
own batch, and the cheapest requests equal the C oracle's walk and the
reference's CRT."""

import math

import numpy as np
import pytest

import dreams_rns as D
from dreams_rns import serve5, synth
from dreams_rns.engine import BatchPlan
from dreams_rns.server import WalkServer
from oracle import cref
from oracle import walk as ow

pytestmark = pytest.mark.gpu


def _requests(seed=1005):
    # drawn with the generator itself until the stream holds a K = 128 and a
    # depth >= 50000 request
    rng = np.random.default_rng(seed)
    reqs = []
    while not (len(reqs) >= 12 and any(r["K"] == 128 for r in reqs) and any(r["depth"] >= 50000 for r in reqs)):
        reqs.append(synth.serve5_request(rng, serve5._compile_cmf))
    return reqs


def _solo(spec):
    with BatchPlan([spec], crt=True) as plan:
        r = plan.run()
        return r.walk_dict(0), float(r.out[0]["delta_float"]), r.crt_values(0)


def test_served_stream_equals_solo_and_oracle(gpu_ctx):
    reqs = _requests()
    specs = [serve5.request_spec(r) for r in reqs]
    assert {2, 3, 4} <= {s.program.m for s in specs} and max(s.depth for s in specs) >= 50000
    rng = np.random.default_rng(7)
    arrivals = np.sort(rng.uniform(0.0, 0.25, size=len(specs)))  # a Poisson stream of len(specs) arrivals
    print("serve5 stream", [(s.program.m, s.depth, s.K) for s in specs], flush=True)
    with WalkServer(round_steps=1000, max_slots=8) as srv:
        out = serve5.run_stream(srv, specs, arrivals, timeout=240)
    print("serve5 stats", out["stats"], flush=True)
    results = out["results"]
    assert [r["spec"] for r in results] == specs and out["stats"]["ok"] == len(specs)
    for i, (r, s) in enumerate(zip(results, specs)):
        ref, dref, (pb, qb) = _solo(s)
        assert r["p_residues"].tolist() == ref["p_residues"].tolist(), i
        assert r["q_residues"].tolist() == ref["q_residues"].tolist(), i
        assert (r["p_big"], r["q_big"]) == (pb, qb), i
        for key in ("p_float", "q_float", "log_scale"):
            assert float(r[key]).hex() == float(ref[key]).hex(), (i, key)
        assert float(r["delta_float"]).hex() == float(dref).hex(), i
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]
    # the three cheapest requests against the C oracle
    cost = [s.depth * s.K * (1 + len(s.program.opcodes)) for s in specs]
    for i in sorted(range(len(specs)), key=lambda j: cost[j])[:3]:
        s, r = specs[i], results[i]
        m = s.program.m
        primes = D.generate_rns_primes(s.K)
        if s.init == "pcf":
            rr, pf, ls = cref.pcf_walk(s.program, s.a0, s.depth, primes)
        else:
            rr, pf, ls = cref.walk(s.program, s.depth, primes, s.axes)
        assert r["p_residues"].tolist() == rr[0, m - 1].tolist(), i
        assert r["q_residues"].tolist() == rr[m - 1, m - 1].tolist(), i
        pl = [int(p) for p in primes]
        xp, M = ow.crt_reconstruct(rr[0, m - 1].tolist(), pl)
        xq, _ = ow.crt_reconstruct(rr[m - 1, m - 1].tolist(), pl)
        assert (r["p_big"], r["q_big"]) == (ow.centered(xp, M), ow.centered(xq, M)), i
        assert r["p_float"] == pf[0, m - 1] and r["q_float"] == pf[m - 1, m - 1], i
        assert abs(r["log_scale"] - ls) <= 1e-12 * max(1.0, abs(ls)), i
        d, _ = D.compute_dreams_delta_float(pf[0, m - 1], pf[m - 1, m - 1], ls, s.target)
        assert r["delta_float"] == d or abs(r["delta_float"] - d) <= 1e-12 * abs(d), i


def test_served_stream_statistics(gpu_ctx):
    pool = serve5.program_pool(4)
    rng = np.random.default_rng(3)
    specs = [serve5.request_spec(serve5.draw_request(rng, pool)) for _ in range(40)]
    arrivals = synth.serve5_arrivals(400.0, 0.1)[: len(specs)]
    specs = specs[: len(arrivals)]
    with WalkServer(round_steps=1000) as srv:
        st = serve5.run_stream(srv, specs, arrivals, timeout=240)["stats"]
    print("serve5 small stream", st, flush=True)
    assert st["ok"] == st["requests"] == len(specs) > 10
    assert 0 < st["p50_ms"] <= st["p99_ms"] and st["served_residue_steps_per_s"] > 0
    assert all(s.K in serve5.K_CHOICES and serve5.DEPTH_RANGE[0] <= s.depth <= serve5.DEPTH_RANGE[1] for s in specs)

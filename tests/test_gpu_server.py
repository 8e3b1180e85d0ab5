"""Continuous-batching server on the GPU: every request's result is the same
alone (its own batch) and in the served mix, bit for bit, whatever rounds it
shared and however its depth was split into rounds."""

import time

import numpy as np
import pytest

import dreams_rns as D
from dreams_rns import synth
from dreams_rns.engine import BatchPlan, cmf_spec, pcf_spec
from dreams_rns.server import WalkServer

pytestmark = pytest.mark.gpu


def _mixed_specs(n_pcf=24, n_cmf=8, seed=5):
    rng = np.random.default_rng(seed)
    specs = []
    cfg = synth.pcf2(n_walks=n_pcf, seed=seed)
    for w in cfg["walks"]:
        depth = int(rng.integers(50, 3000))
        K = int(rng.choice([16, 32, 64]))
        specs.append(pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), depth, K, 1,
                              target=w["L"]))
    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    for i in range(n_cmf):
        m = 3 if i % 2 == 0 else 4
        spec = synth.cmf_walk_spec(rng, m, cf, 10, 200, False)
        specs.append(cmf_spec(spec["program"], int(rng.integers(20, 400)), int(rng.choice([16, 32])),
                              spec["shift"], None, target=0.5))
    return specs


def _alone(spec):
    with BatchPlan([spec], crt=True) as plan:
        r = plan.run()
        return r.walk_dict(0), float(r.out[0]["delta_float"]), r.crt_values(0)


def test_served_results_equal_alone(gpu_ctx):
    specs = _mixed_specs()
    alone = [_alone(s) for s in specs]
    rng = np.random.default_rng(11)
    with WalkServer(round_steps=97, max_slots=12) as srv:
        tickets = {}
        for i, s in enumerate(specs):
            tickets[srv.submit(s)] = i
            if rng.random() < 0.3:
                time.sleep(0.002)  # some requests arrive while others are on the GPU
            q, r = srv.pending()
            assert r <= 12
        results = srv.drain(timeout=120)
    assert len(results) == len(specs)
    for r in results:
        assert r["status"] == 0
        i = tickets[r["ticket"]]
        ref, dref, crt = alone[i]
        assert r["p_residues"].tolist() == ref["p_residues"].tolist(), i
        assert r["q_residues"].tolist() == ref["q_residues"].tolist(), i
        assert (r["p_big"], r["q_big"]) == crt, i
        for key in ("p_float", "q_float", "log_scale"):
            assert float(r[key]).hex() == float(ref[key]).hex(), (i, key)
        assert float(r["delta_float"]).hex() == float(dref).hex(), i
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]


def test_server_admits_by_memory_budget(gpu_ctx):
    specs = _mixed_specs(n_pcf=10, n_cmf=0, seed=9)
    # a budget for a few slots: the rest wait in the queue
    with WalkServer(round_steps=500, mem_budget=600_000) as srv:
        cap = srv.capacity()
        assert 1 <= cap < len(specs)
        for s in specs:
            srv.submit(s)
        peak = 0
        results = []
        t0 = time.time()
        while len(results) < len(specs) and time.time() - t0 < 120:
            results += srv.poll()
            peak = max(peak, srv.pending()[1])
        assert len(results) == len(specs)
        assert 1 <= peak <= cap

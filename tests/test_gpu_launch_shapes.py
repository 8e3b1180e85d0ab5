"""Launch shapes do not change results: walks per workgroup (lane groups of
one workgroup on different walks' step ranges) and the prime stripe of wide
walks give residues, CRT integers, FP64 results and top-k bit-identical to
the default shape, on every walk-kernel family -- the PCF and companion
forward-difference kernels, the general one and the interpreter (whose
lane groups share a program) -- in batched and persistent modes; a few
walks of each batch also against the C oracle."""

import numpy as np
import pytest

import dreams_rns as D
from dreams_rns import synth
from dreams_rns.engine import BatchPlan, cmf_spec, pcf_spec
from oracle import cref

pytestmark = pytest.mark.gpu


def _cf(mat, m, dim, axes, dirs):
    return D.compile_cmf_from_dict(mat, m, dim, axes, dirs)


def _copy(res):
    return {"residues": res.residues.copy(), "out": res.out.copy(), "limbs": res.limbs.copy(),
            "neg": res.neg.copy(), "topk": res.topk.copy()}


def _same(a, b, what):
    assert np.array_equal(a["residues"], b["residues"]), what
    assert np.array_equal(a["limbs"], b["limbs"]) and np.array_equal(a["neg"], b["neg"]), what
    for key in ("p_float", "q_float", "log_scale", "delta_float"):
        assert [float(v).hex() for v in a["out"][key]] == [float(v).hex() for v in b["out"][key]], (what, key)
    assert a["topk"].tolist() == b["topk"].tolist(), what


def _run(specs, top_k, **kw):
    print("launch shape", kw, len(specs), "walks", flush=True)
    with BatchPlan(specs, crt=True, top_k=top_k, **kw) as plan:
        return _copy(plan.run())


def _mix(n, seed):
    cfg = synth.mix4(_cf, n_walks=n, seed=seed, top_k=8)
    specs = []
    for w in cfg["walks"]:
        if w["kind"] == "pcf":
            p = D.compile_pcf(w["a"], w["b"])
            specs.append(pcf_spec(p, D.pcf_initial_values(w["a"]), w["depth"], w["K"], 1, target=w["L"]))
        else:
            specs.append(cmf_spec(w["program"], min(w["depth"], 3000), w["K"], w["shift"], target=w["L"]))
    return specs


def test_walks_per_workgroup_on_an_uneven_mix(gpu_ctx):
    # 2x2 / 3x3 / 4x4, K 16..128, uneven depths: every family in one batch
    specs = _mix(48, seed=2207)
    ref = _run(specs, 8, threads=256, chunk_steps=512)
    for kw in ({"threads": 256, "chunk_steps": 512, "walks_per_wg": 2},
               {"threads": 256, "chunk_steps": 512, "walks_per_wg": 16},
               {"threads": 128, "chunk_steps": 1024, "walks_per_wg": 4, "slice": 32},
               {"threads": 256, "chunk_steps": 512, "walks_per_wg": 4, "mode": "persistent"}):
        _same(ref, _run(specs, 8, **kw), kw)


def test_walks_per_workgroup_sharing_programs(gpu_ctx):
    # interpreted 3x3 CMFs: two programs, eight shifts each, K = 16: lane
    # groups of one workgroup run different walks of one program
    cfg = synth.cmf3(_cf, n_walks=2)
    specs = [cmf_spec(w["program"], 700, 16, [(v + sh) % 6 for v in w["shift"]], target=0.25)
             for w in cfg["walks"] for sh in range(8)]
    ref = _run(specs, 4, threads=256, chunk_steps=512)
    for kw in ({"threads": 256, "chunk_steps": 512, "walks_per_wg": 16},
               {"threads": 64, "chunk_steps": 256, "walks_per_wg": 4}):
        _same(ref, _run(specs, 4, **kw), kw)
    primes = D.generate_rns_primes(16)
    with BatchPlan(specs[:3], crt=True, threads=256, walks_per_wg=16) as plan:
        res = plan.run()
        for i, s in enumerate(specs[:3]):
            rr, _, _ = cref.walk(s.program, s.depth, primes, list(s.axes))
            d = res.walk_dict(i)
            m = s.program.m
            assert d["p_residues"].tolist() == rr[0, m - 1].tolist(), i
            assert d["q_residues"].tolist() == rr[m - 1, m - 1].tolist(), i


def test_walks_per_workgroup_on_pcf_shapes(gpu_ctx):
    # pcf2 PCFs (49 shapes of deg b, deg a): lane groups share a shape
    cfg = synth.pcf2(n_walks=160, seed=77)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 900, 32, 1, target=w["L"])
             for w in cfg["walks"]]
    ref = _run(specs, 16, threads=64, chunk_steps=4096)
    for kw in ({"threads": 64, "chunk_steps": 4096, "walks_per_wg": 2},
               {"threads": 256, "chunk_steps": 300, "walks_per_wg": 8}):
        _same(ref, _run(specs, 16, **kw), kw)
    primes = D.generate_rns_primes(32)
    for i in (0, 81, 159):
        rr, _, _ = cref.pcf_walk(specs[i].program, specs[i].a0, 900, primes)
        with BatchPlan([specs[i]], crt=True, threads=64, walks_per_wg=2) as plan:
            d = plan.run().walk_dict(0)
        assert d["p_residues"].tolist() == rr[0, 1].tolist() and d["q_residues"].tolist() == rr[1, 1].tolist(), i


def test_prime_stripes_of_wide_walks(gpu_ctx):
    # 300 primes as stripes of 64 or 128 (with two walks per workgroup):
    # the slices of a step range write their own columns of one product
    gpu_ctx.ensure_primes(300)
    cfg = synth.pcf2(n_walks=6, seed=91)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 700, 300, 1, target=w["L"])
             for w in cfg["walks"]]
    progs = synth.cmf3(_cf, n_walks=2)["walks"]
    specs += [cmf_spec(w["program"], 400, 300, w["shift"], target=0.25) for w in progs]
    ref = _run(specs, 2)
    for kw in ({"slice": 64}, {"slice": 128, "walks_per_wg": 2}, {"slice": 32, "threads": 256, "walks_per_wg": 4}):
        _same(ref, _run(specs, 2, **kw), kw)


def test_bad_launch_shapes_are_refused(gpu_ctx):
    specs = _mix(4, seed=5)
    for kw in ({"slice": 48}, {"slice": 512}, {"walks_per_wg": 3}):
        with pytest.raises(Exception):
            BatchPlan(specs, **kw)


def test_graph_replays_equal_stream_launch(gpu_ctx):
    # the captured launch sequence (three M_f / shadow windows, the first
    # ahead of the seeds, top-k on the shadow's stream) replayed four times,
    # with stream-launched work between the replays and both the copying and
    # the resident run: every replay gives the stream launch's outputs
    cfg = synth.pcf2(n_walks=256, seed=131)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 1200, 32, 1, target=w["L"])
             for w in cfg["walks"]]
    kw = {"crt": True, "top_k": 16, "threads": 64, "chunk_steps": 4096}
    with BatchPlan(specs, **kw) as ps:
        ref = _copy(ps.run())
        with BatchPlan(specs, use_graph=True, **kw) as pg:
            for rep in range(4):
                print("graph replay", rep, flush=True)
                if rep % 2:
                    pg.run(fetch=False)
                    got = _copy(pg.fetch())
                else:
                    got = _copy(pg.run())
                _same(ref, got, rep)
                ps.run()


def test_copy_back_into_pageable_sinks(gpu_ctx, monkeypatch):
    # page-locked sinks take the outputs as soon as each is final (residues
    # and CRT limbs beside the shadow chain); pageable ones get every copy
    # after the join: both give the same outputs, with and without a graph,
    # also when the sinks are reused run after run
    from dreams_rns import _native as N
    cfg = synth.pcf2(n_walks=192, seed=137)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 900, 32, 1, target=w["L"])
             for w in cfg["walks"]]
    kw = {"crt": True, "top_k": 16, "threads": 64, "chunk_steps": 4096}
    with BatchPlan(specs, **kw) as pin:
        ref = _copy(pin.run())
        monkeypatch.setattr(N, "pinned_zeros", lambda shape, dtype: np.zeros(shape, dtype))
        for graph in (False, True):
            print("pageable sinks, graph", graph, flush=True)
            with BatchPlan(specs, use_graph=graph, **kw) as pg:
                for rep in range(2):
                    _same(ref, _copy(pg.run()), ("pageable", graph, rep))
                    _same(ref, _copy(pin.run()), ("page-locked", graph, rep))

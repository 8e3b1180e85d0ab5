"""GPU parity on the shapes of the other workloads: 3x3 CMFs with INV at the
cmf3 depth and prime count, the uneven mix of 2x2 / 3x3 / 4x4 walks (mix4),
and the full 4096-walk pcf2 batch through size-independent properties
(every CRT integer reduces back to its residues and lies in the centred
range; the top-k is the order of the returned deltas)."""

import math

import numpy as np
import pytest

import dreams_rns as D
from dreams_rns import synth
from dreams_rns.engine import BatchPlan, cmf_spec, pcf_spec
from oracle import cref
from oracle import walk as ow

pytestmark = pytest.mark.gpu
REL = 1e-12


def _close(a, b, rel=REL):
    a, b = float(a), float(b)
    return a == b or abs(a - b) <= rel * max(abs(a), abs(b))


def _cf(mat, m, dim, axes, dirs):
    return D.compile_cmf_from_dict(mat, m, dim, axes, dirs)


def _copy(res):
    return {"residues": res.residues.copy(), "out": res.out.copy(),
            "limbs": None if res.limbs is None else res.limbs.copy(),
            "neg": None if res.neg is None else res.neg.copy(),
            "topk": None if res.topk is None else res.topk.copy()}


def _oracle_check(res, i, spec, primes):
    m = spec.program.m
    init = (None, None)
    if spec.init == "pcf":
        init = (np.array([[1] * spec.K, [spec.a0 % int(q) for q in primes[: spec.K]], [0] * spec.K, [1] * spec.K]),
                np.array([1.0, float(spec.a0), 0.0, 1.0]))
    rr, Pf, ls = cref.walk(spec.program, spec.depth, primes[: spec.K], list(spec.axes), *init)
    d = res.walk_dict(i)
    assert d["p_residues"].tolist() == rr[0, m - 1].tolist(), i
    assert d["q_residues"].tolist() == rr[m - 1, m - 1].tolist(), i
    assert _close(d["p_float"], Pf[0, m - 1]) and _close(d["q_float"], Pf[m - 1, m - 1]), i
    assert _close(d["log_scale"], ls), i
    pb, qb = res.crt_values(i)
    pl = [int(q) for q in primes[: spec.K]]
    xp, M = ow.crt_reconstruct(rr[0, m - 1].tolist(), pl)
    xq, _ = ow.crt_reconstruct(rr[m - 1, m - 1].tolist(), pl)
    assert (pb, qb) == (ow.centered(xp, M), ow.centered(xq, M)), i


def test_cmf3_shape_full_depth_against_c_oracle(gpu_ctx):
    # 3x3 rational CMFs compiled to 100-250 ops with INV, depth 10000, K = 64
    cfg = synth.cmf3(_cf, n_walks=6)
    specs = [cmf_spec(w["program"], cfg["depth"], cfg["K"], w["shift"], target=0.25) for w in cfg["walks"]]
    with BatchPlan(specs, crt=True, top_k=3) as plan:
        res = plan.run()
        primes = D.generate_rns_primes(64)
        for i, s in enumerate(specs):
            print("cmf3 walk", i, len(s.program.opcodes), "ops", flush=True)
            _oracle_check(res, i, s, primes)
        dl = res.out["delta_float"].tolist()
        assert res.topk.tolist() == sorted(range(len(dl)), key=lambda j: (-dl[j], j))[:3]


def _mix_specs(n, seed=1004):
    cfg = synth.mix4(_cf, n_walks=n, seed=seed, top_k=16)
    specs = []
    for w in cfg["walks"]:
        if w["kind"] == "pcf":
            p = D.compile_pcf(w["a"], w["b"])
            specs.append(pcf_spec(p, D.pcf_initial_values(w["a"]), w["depth"], w["K"], 1, target=w["L"]))
        else:
            specs.append(cmf_spec(w["program"], w["depth"], w["K"], w["shift"], target=w["L"]))
    return specs


def test_mix4_shape_batched_equals_plain_and_oracle(gpu_ctx):
    # uneven depths (500..50000), orders 2/3/4, K up to 128: one batched plan
    # (chunked tasks, longest first) equals one launch sequence per walk, and
    # the cheapest walks match the C restatement
    specs = _mix_specs(40)
    with BatchPlan(specs, crt=True, top_k=16) as plan:
        a = _copy(plan.run())
        res = plan.result
        primes = D.generate_rns_primes(128)
        cost = sorted(range(len(specs)), key=lambda i: specs[i].depth * specs[i].K * len(specs[i].program.opcodes))
        for i in cost[:5]:
            print("mix4 oracle walk", i, specs[i].program.m, specs[i].depth, specs[i].K, flush=True)
            _oracle_check(res, i, specs[i], primes)
    with BatchPlan(specs, mode="plain", crt=True, top_k=16) as plan:
        b = _copy(plan.run())
    # persistent mode: task-queue walk kernels, the whole sequence one graph,
    # replayed three times with host work between the replays
    with BatchPlan(specs, mode="persistent", crt=True, top_k=16) as plan:
        c = [_copy(plan.run()) for _ in range(3)]
    for other in [b] + c:
        assert np.array_equal(a["residues"], other["residues"])
        assert np.array_equal(a["limbs"], other["limbs"]) and np.array_equal(a["neg"], other["neg"])
        for key in ("p_float", "q_float", "log_scale", "delta_float"):
            assert [float(v).hex() for v in a["out"][key]] == [float(v).hex() for v in other["out"][key]], key
        assert a["topk"].tolist() == other["topk"].tolist()
    dl = a["out"]["delta_float"].tolist()
    assert a["topk"].tolist() == sorted(range(len(dl)), key=lambda j: (-dl[j], j))[:16]


def _family(spec):
    # the walk kernel family the library picks for this program
    from dreams_rns._native import pack_program

    meta, m = pack_program(spec.program)[2], spec.program.m
    if meta["fd_ok"] and m == 2 and meta["fd_degree"] <= 6:
        k = meta["kinds"]
        return "fd_pcf" if (k & 3) == 0 and ((k >> 4) & 3) == 1 else "fd"
    if meta["fd_ok"] and m in (3, 4) and meta["fd_degree"] <= 3 and all(
            ((meta["kinds"] >> (2 * (i * m + j))) & 3) == (1 if i == j + 1 else 0)
            for i in range(m) for j in range(m - 1)):
        return "fd_comp"
    if meta["fd_ok"] and m == 3 and meta["fd_degree"] <= 3:
        return "fd"
    return "interp"


def test_mix4_costliest_walks_against_c_oracle(gpu_ctx):
    # for every (matrix order, kernel family) of a 400-walk mix4 batch, the
    # deepest walk at the highest K (depth up to 50000, K 128, interpreted
    # 4x4 programs with INV) against the C restatement, in the batch
    specs = _mix_specs(400)
    groups = {}
    for i, s in enumerate(specs):
        groups.setdefault((s.program.m, _family(s)), []).append(i)
    picks = {key: max(idx, key=lambda j: (specs[j].depth * specs[j].K, j)) for key, idx in groups.items()}
    assert {(2, "fd_pcf"), (3, "interp"), (4, "interp")} <= set(picks)
    primes = D.generate_rns_primes(128)
    with BatchPlan(specs, crt=True, top_k=16, threads=256, chunk_steps=512) as plan:
        res = plan.run()
        for key, i in sorted(picks.items()):
            s = specs[i]
            print("mix4 costliest", key, "walk", i, "depth", s.depth, "K", s.K, len(s.program.opcodes), "ops",
                  flush=True)
            _oracle_check(res, i, s, primes)


def test_pcf2_full_batch_size_independent_properties(gpu_ctx):
    cfg = synth.pcf2()
    depth, K = cfg["depth"], cfg["K"]
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), depth, K, 1, target=w["L"])
             for w in cfg["walks"]]
    primes = [int(p) for p in D.generate_rns_primes(K)]
    M = math.prod(primes)
    with BatchPlan(specs, crt=True, top_k=cfg["top_k"]) as plan:
        res = plan.run()
        print("pcf2 batch ran", len(specs), "walks", flush=True)
        for i in range(len(specs)):
            for row in (0, 1):
                mag = int.from_bytes(res.limbs[i, row, :K].tobytes(), "little")
                v = -mag if res.neg[i, row] else mag
                assert 2 * abs(v) <= M, (i, row)
                assert [v % p for p in primes] == res.residues[i, row, :K].tolist(), (i, row)
        dl = res.out["delta_float"].tolist()
        assert res.topk.tolist() == sorted(range(len(dl)), key=lambda j: (-dl[j], j))[: cfg["top_k"]]
        rng = np.random.default_rng(17)
        for i in rng.choice(len(specs), size=12, replace=False):
            _oracle_check(res, int(i), specs[int(i)], np.array(primes, dtype=np.int64))

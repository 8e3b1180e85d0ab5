"""GPU parity of the pFq CMF sweep (a100_sweep.py:336-528 over the
reference's sweep_data subset) and of walks wider than one workgroup's 256
primes: residues, full matrices and CRT integers bit-exact against the C
oracle, the FP64 shadow within 1e-12 relative."""

import json
import os

import numpy as np
import pytest

import dreams_rns as D
from dreams_rns.cmf_walk import compile_cmf_spec
from dreams_rns.engine import BatchPlan, cmf_spec
from dreams_rns.sweep import CmfSweep, sweep_offsets
from oracle import cref
from oracle import walk as ow

pytestmark = pytest.mark.gpu
REL = 1e-12
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sweep_3f2.json")


def _close(a, b, rel=REL):
    a, b = float(a), float(b)
    return a == b or abs(a - b) <= rel * max(abs(a), abs(b))


@pytest.fixture(scope="module")
def sweep_fix():
    with open(FIX) as f:
        return json.load(f)


def _axes(offsets, traj):
    return [(s, 1, t, 1) for s, t in zip(offsets, traj)]


def test_sweep_subset_against_c_oracle(gpu_ctx, sweep_fix):
    specs = [sweep_fix["specs"][0], sweep_fix["specs"][37]]
    trajs = [sweep_fix["trajectories"][i] for i in (0, 4, 9, 21)]
    shifts = sweep_fix["shifts"]
    depth, K = 120, 32
    primes = D.generate_rns_primes(K)
    sw = CmfSweep(specs, trajs, shifts, depth=depth, K=K, max_walks=100, crt=True)
    n_rows = n_rec = 0
    for batch in sw:
        print("sweep batch", len(batch), flush=True)
        for i in range(len(batch)):
            si, ti = int(batch.spec_idx[i]), int(batch.traj_idx[i])
            prog = compile_cmf_spec(specs[si], trajs[ti])
            rr, Pf, ls = cref.walk(prog, depth, primes, _axes(batch.offsets[i], trajs[ti]))
            d = batch.walk_dict(i)
            m = prog.m
            assert d["p_residues"].tolist() == rr[0, m - 1].tolist(), (si, ti, batch.offsets[i])
            assert d["q_residues"].tolist() == rr[m - 1, m - 1].tolist(), (si, ti, batch.offsets[i])
            assert _close(d["p_float"], Pf[0, m - 1]) and _close(d["q_float"], Pf[m - 1, m - 1])
            assert _close(d["log_scale"], ls)
            xp, M = ow.crt_reconstruct(rr[0, m - 1].tolist(), primes.tolist())
            xq, _ = ow.crt_reconstruct(rr[m - 1, m - 1].tolist(), primes.tolist())
            assert (d["p_crt"], d["q_crt"]) == (ow.centered(xp, M), ow.centered(xq, M))
        recs = list(batch.records())
        for r in recs:
            assert r["shift_vals"] == sweep_offsets([shifts[r["shift_idx"]]], 3, 5)[0]
        n_rows += len(batch)
        n_rec += len(recs)
    assert sw.skipped == []
    assert n_rec == len(specs) * len(trajs) * len(shifts)
    assert n_rows < n_rec  # shifts of equal offsets walked once


@pytest.mark.parametrize("K", [257, 300, 1000, 1024])
def test_prime_slices_match_c_oracle(gpu_ctx, sweep_fix, K):
    # walks of more than 256 primes run as prime slices; several step ranges
    # per walk (chunk_steps 32) make every slice combine in order
    spec = sweep_fix["specs"][5]
    cases = [([1, 2, 1, 3, 1], sweep_fix["trajectories"][7]), ([4, 1, 2, 1, 1], sweep_fix["trajectories"][20]),
             ([2, 2, 2, 2, 2], [1, 1, 1, 1, 1])]
    depth = 100
    primes = D.generate_rns_primes(K)
    specs, refs = [], []
    for off, traj in cases:
        prog = compile_cmf_spec(spec, traj)
        specs.append(cmf_spec(prog, depth, K, off))
        refs.append(cref.walk(prog, depth, primes, _axes(off, traj)))
    print("slices K", K, flush=True)
    with BatchPlan(specs, full=True, crt=True, chunk_steps=32) as plan:
        res = plan.run()
    for i, (rr, Pf, ls) in enumerate(refs):
        assert np.array_equal(res.full_matrix(i, 4), rr), (K, i)
        d = res.walk_dict(i)
        assert d["primes"].tolist() == primes.tolist()
        assert _close(d["p_float"], Pf[0, 3]) and _close(d["q_float"], Pf[3, 3]) and _close(d["log_scale"], ls)
        xp, M = ow.crt_reconstruct(rr[0, 3].tolist(), primes.tolist())
        xq, _ = ow.crt_reconstruct(rr[3, 3].tolist(), primes.tolist())
        assert res.crt_values(i) == (ow.centered(xp, M), ow.centered(xq, M))


def test_more_primes_than_one_context_holds_is_refused(gpu_ctx, sweep_fix):
    # 8192 primes is the limit of a matrix walk: 1025 runs (five prime slices)
    # and equals the C oracle, one more than 8192 is refused before anything
    # runs, with the reference's exception type for bad arguments
    prog = compile_cmf_spec(sweep_fix["specs"][5], [1, 1, 1, 1, 1])
    primes = D.generate_rns_primes(1025)
    with BatchPlan([cmf_spec(prog, 10, 1025, [1, 1, 1, 1, 1])], full=True, crt=True) as plan:
        res = plan.run()
    rr, _, _ = cref.walk(prog, 10, primes, _axes([1, 1, 1, 1, 1], [1, 1, 1, 1, 1]))
    assert np.array_equal(res.full_matrix(0, 4), rr)
    xq, M = ow.crt_reconstruct(rr[3, 3].tolist(), primes.tolist())
    assert res.crt_values(0)[1] == ow.centered(xq, M)
    with pytest.raises(ValueError, match="8193"):
        BatchPlan([cmf_spec(prog, 10, 8193, [1, 1, 1, 1, 1])])


def test_mixed_widths_in_one_batch(gpu_ctx, sweep_fix):
    # walks of 64 and of 600 primes, 4x4 and 2x2, share one batch
    spec = sweep_fix["specs"][11]
    prog4 = compile_cmf_spec(spec, sweep_fix["trajectories"][3])
    prog2 = compile_cmf_spec({"p": 1, "q": 1, "a_params": ["1/2"], "b_params": ["3/2"]}, [1, 2])
    plan_in = [(prog4, 64, [1, 2, 3, 1, 2]), (prog4, 600, [2, 1, 1, 4, 1]), (prog2, 600, [3, 1]),
               (prog2, 64, [1, 1]), (prog4, 600, [1, 1, 1, 1, 1])]
    depth = 90
    kmax = max(k for _, k, _ in plan_in)
    primes = D.generate_rns_primes(kmax)
    specs = [cmf_spec(p, depth, k, off) for p, k, off in plan_in]
    with BatchPlan(specs, full=True) as plan:
        res = plan.run()
    for i, (p, k, off) in enumerate(plan_in):
        rr, Pf, ls = cref.walk(p, depth, primes[:k], _axes(off, p.directions))
        assert np.array_equal(res.full_matrix(i, p.m), rr), i
        d = res.walk_dict(i)
        m = p.m
        assert _close(d["p_float"], Pf[0, m - 1]) and _close(d["q_float"], Pf[m - 1, m - 1]), i


def test_batch_api_with_sweep_offsets(gpu_ctx, sweep_fix):
    spec = sweep_fix["specs"][20]
    traj = sweep_fix["trajectories"][12]
    prog = compile_cmf_spec(spec, traj)
    offs = sweep_offsets(sweep_fix["shifts"][:9], 3, 5)
    got = D.run_cmf_walk_batch_gpu(prog, 80, 48, shift_val_list=offs, batch_size=4)
    assert len(got) == len(offs)
    for sv, g in zip(offs, got):
        one = D.run_cmf_walk(prog, 80, 48, sv)
        assert g["shift_vals"] == sv
        assert g["p_residues"].tolist() == one["p_residues"].tolist()
        assert g["q_residues"].tolist() == one["q_residues"].tolist()
        assert g["p_float"].hex() == one["p_float"].hex() and g["q_float"].hex() == one["q_float"].hex()


def test_companion_forward_differences_equal_interpreter(gpu_ctx):
    # pFq companions (3x3 and 4x4, last columns of degree 0..3 with static
    # ones) run on the companion forward-difference kernel; forcing the
    # interpreter must give the same residues, shadow and CRT, bit for bit,
    # across step ranges (chunk_steps 64) and prime slices (K 300)
    pfq = [(2, 1, ["1/2", "-1/3"], ["3/2"]), (0, 2, [], ["1/2", "1/3"]), (2, 2, ["1", "2"], ["-1/2", "5/4"]),
           (3, 2, ["-1/4", "-1/6", "1/6"], ["-1/3", "1/4"]), (3, 1, ["1/2", "1/2", "1"], ["2"]),
           (1, 2, ["1/3"], ["1", "1/2"]), (3, 3, ["1", "1", "1"], ["2", "2", "2"])]
    rng = np.random.default_rng(12)
    specs = []
    for (p, q, a, b), depth, K in zip(pfq * 2, [1, 2, 7, 100, 333, 1000, 5000] * 2, [16, 32, 300, 64, 300, 31, 128] * 2):
        spec = {"p": p, "q": q, "a_params": a, "b_params": b}
        traj = [int(v) for v in rng.integers(-2, 3, p + q)]
        prog = compile_cmf_spec(spec, traj)
        off = [int(v) for v in rng.integers(1, 5, p + q)]
        specs.append(cmf_spec(prog, depth, K, off))
    print("companion fd vs interpreter", len(specs), flush=True)
    a = BatchPlan(specs, full=True, crt=True, chunk_steps=64).run()
    b = BatchPlan(specs, full=True, crt=True, chunk_steps=64, interp_only=True).run()
    assert np.array_equal(a.full, b.full)
    assert np.array_equal(a.full_f, b.full_f)
    assert np.array_equal(a.limbs, b.limbs) and np.array_equal(a.neg, b.neg)
    primes = D.generate_rns_primes(300)
    for i in (0, 3, 4, 6, 9, 12):
        s = specs[i]
        rr, Pf, ls = cref.walk(s.program, s.depth, primes[: s.K], list(s.axes))
        assert np.array_equal(a.full_matrix(i, s.program.m), rr), i


def test_companion_rows_only_path(gpu_ctx):
    # without the full state, a companion walk from the identity in one step
    # range carries rows 0 and m - 1 only when its workgroup holds one
    # sub-chunk (256 primes per slice); p, q, the shadow and CRT must equal
    # the interpreter's, and walks of two ranges (depth > 4096) the full path
    pfq = [(3, 2, ["-1/4", "-1/6", "1/6"], ["-1/3", "1/4"]), (2, 1, ["1/2", "-1/3"], ["3/2"]),
           (0, 2, [], ["1/2", "1/3"]), (3, 3, ["1", "1", "1"], ["2", "2", "2"])]
    rng = np.random.default_rng(21)
    specs = []
    for (p, q, a, b), depth, K in zip(pfq * 3, [1, 9, 500, 2000, 4096, 5000, 777, 3, 64, 1500, 4097, 250],
                                      [256, 300, 1000, 512, 256, 257, 64, 1000, 256, 700, 300, 32]):
        spec = {"p": p, "q": q, "a_params": a, "b_params": b}
        prog = compile_cmf_spec(spec, [int(v) for v in rng.integers(-2, 3, p + q)])
        specs.append(cmf_spec(prog, depth, K, [int(v) for v in rng.integers(1, 5, p + q)]))
    print("companion rows-only", len(specs), flush=True)
    a = BatchPlan(specs, crt=True).run()
    b = BatchPlan(specs, crt=True, interp_only=True).run()
    assert np.array_equal(a.residues, b.residues)
    assert np.array_equal(a.limbs, b.limbs) and np.array_equal(a.neg, b.neg)
    for key in ("p_float", "q_float", "log_scale"):
        assert [x.hex() for x in a.out[key].tolist()] == [x.hex() for x in b.out[key].tolist()], key
    primes = D.generate_rns_primes(1000)
    for i in (2, 3, 5, 7):
        s = specs[i]
        rr, Pf, ls = cref.walk(s.program, s.depth, primes[: s.K], list(s.axes))
        d = a.walk_dict(i)
        m = s.program.m
        assert d["p_residues"].tolist() == rr[0, m - 1].tolist() and d["q_residues"].tolist() == rr[m - 1, m - 1].tolist(), i


def test_sweep_hits_on_device_batches(gpu_ctx, sweep_fix):
    # the vectorised hit selection over real device batches equals the
    # reference loop run walk by walk (oracle/postwalk.py), for overflowed
    # CRTs (3F2, depth 300, K 32) and exact ones (an integer 2F1, depth 25)
    from dreams_rns.constants import load_constants
    from dreams_rns.postwalk import sweep_hits
    from oracle import postwalk as opw

    bank = load_constants(dps=60)
    cases = [
        ([sweep_fix["specs"][0], sweep_fix["specs"][44]], [sweep_fix["trajectories"][i] for i in (0, 3, 17)],
         sweep_fix["shifts"][:24], 300, 32),
        ([{"name": "2F1(1,1;2)", "spec_hash": "int", "rank": 3, "p": 2, "q": 1, "dim": 3, "a_params": ["1", "1"],
           "b_params": ["2"]}], [[1, 1, 1], [1, 0, 1], [2, 1, 1]], sweep_fix["shifts"][:16], 25, 32),
    ]
    n_hits = 0
    n_exact = 0
    for specs, trajs, shifts, depth, K in cases:
        for batch in CmfSweep(specs, trajs, shifts, depth=depth, K=K, crt=True):
            print("hits batch", len(batch), flush=True)
            got = sweep_hits(batch, specs, trajs, bank, proximity=1e-3, dps=60)
            ref = []
            for i in range(len(batch)):
                d = batch.result.walk_dict(i)
                h = opw.walk_record(d, bank, 1e-3, 60)
                p_big, q_big = batch.result.crt_values(i)
                if q_big != 0 and max(abs(p_big), abs(q_big)).bit_length() < 1000:
                    n_exact += 1
                if h is None:
                    continue
                si, ti = int(batch.spec_idx[i]), int(batch.traj_idx[i])
                for s in batch.shift_rows[i]:
                    ref.append((si, ti, s, {"cmf_name": specs[si]["name"], "traj_idx": ti, "shift_idx": s, **h}))
            ref.sort(key=lambda t: t[:3])
            ref = [t[3] for t in ref]
            assert len(got) == len(ref)
            for g, r in zip(got, ref):
                assert {k: g[k] for k in r} == r
            n_hits += len(got)
    assert n_exact > 0


def test_served_companion_walks_equal_batched(gpu_ctx, sweep_fix):
    # 3x3 / 4x4 companions run on the companion forward-difference kernel in
    # the server too (Newton tables seeded on a walk's first round, carried
    # across rounds); each result equals its batched run bit for bit
    from dreams_rns.server import WalkServer

    rng = np.random.default_rng(8)
    specs = []
    for i in range(10):
        spec = sweep_fix["specs"][3 * i] if i % 2 == 0 else {"p": 2, "q": 1, "a_params": ["1/2", "1/3"],
                                                                "b_params": ["5/4"]}
        dim = 5 if i % 2 == 0 else 3
        prog = compile_cmf_spec(spec, [int(v) for v in rng.integers(-2, 3, dim)])
        specs.append(cmf_spec(prog, int(rng.integers(30, 700)), int(rng.choice([16, 32, 64])),
                              [int(v) for v in rng.integers(1, 5, dim)]))
    with BatchPlan(specs) as plan:
        ref = plan.run()
        ref = [ref.walk_dict(i) for i in range(len(specs))]
    print("served companions", len(specs), flush=True)
    with WalkServer(round_steps=53, max_slots=6) as srv:
        tickets = {srv.submit(s): i for i, s in enumerate(specs)}
        results = srv.drain(timeout=120)
    assert len(results) == len(specs)
    for r in results:
        assert r["status"] == 0
        i = tickets[r["ticket"]]
        assert r["p_residues"].tolist() == ref[i]["p_residues"].tolist(), i
        assert r["q_residues"].tolist() == ref[i]["q_residues"].tolist(), i
        for key in ("p_float", "q_float", "log_scale"):
            assert float(r[key]).hex() == float(ref[i][key]).hex(), (i, key)


def test_seed_launch_shapes_equal_interpreter(gpu_ctx):
    # Newton-table seeds split over point groups (small batches) and one
    # workgroup per seed task (launches of 2^20 lanes and more: 4200 walks x
    # 256 primes) give the tables the program's steps imply: residues and
    # CRT bit-exact against the interpreter
    spec = {"p": 3, "q": 2, "a_params": ["-1/4", "-1/6", "1/6"], "b_params": ["-1/3", "1/4"]}
    rng = np.random.default_rng(33)
    progs = [compile_cmf_spec(spec, [int(v) for v in rng.integers(-2, 3, 5)]) for _ in range(8)]
    for n, K in ((64, 32), (4200, 256)):
        specs = [cmf_spec(progs[i % 8], 40, K, [int(v) for v in rng.integers(1, 5, 5)]) for i in range(n)]
        print("seed shapes", n, K, flush=True)
        a = BatchPlan(specs, crt=True).run()
        b = BatchPlan(specs, crt=True, interp_only=True).run()
        assert np.array_equal(a.residues, b.residues), (n, K)
        assert np.array_equal(a.limbs, b.limbs) and np.array_equal(a.neg, b.neg), (n, K)

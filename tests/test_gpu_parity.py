"""GPU parity: the HIP path (through the C ABI) against the reference's golden
vectors and the oracle, bit-exact for residues and CRT integers; the FP64
shadow bit-exact where the reference's floating-point operations are exact
or correctly rounded, else within the contract tolerance: delta_float within
1e-12 relative."""

import math

import numpy as np
import pytest

from conftest import program_from_json

import dreams_rns as D
from dreams_rns import synth
from dreams_rns.engine import BatchPlan, cmf_spec, pcf_spec
from oracle import cref, exact
from oracle import walk as ow

pytestmark = pytest.mark.gpu
REL = 1e-12


def _hex(x):
    return float(x).hex()


def _close(a, b, rel=REL):
    a, b = float(a), float(b)
    return a == b or abs(a - b) <= rel * max(abs(a), abs(b))


def test_pcf_golden_vectors(gpu_ctx, golden):
    for w in golden["run_pcf_walk"]:
        p = program_from_json(w["program"])
        got = D.run_pcf_walk(p, w["a0"], w["depth"], w["K"], w["shift_val"])
        o = w["out"]
        print("pcf", w["a"], "|", w["b"], w["depth"], w["K"], flush=True)
        assert got["p_residues"].tolist() == o["p_residues"]
        assert got["q_residues"].tolist() == o["q_residues"]
        assert got["primes"].tolist() == o["primes"]
        assert got["p_residues"].dtype == np.int64 and isinstance(got["p_float"], np.float64)
        assert _hex(got["p_float"]) == o["p_float"] and _hex(got["q_float"]) == o["q_float"]
        assert _close(got["log_scale"], float.fromhex(o["log_scale"]))


def test_cmf_rational_axes_whole_shadow(gpu_ctx, golden):
    # 3x3 CMF with rational shift and trajectory: the whole shadow matrix,
    # step by step, against the C restatement (P_f after 1 step is M_f(0))
    c = golden["run_cmf_walk"][1]
    p = program_from_json(c["program"])
    shift = [tuple(v) for v in c["shift"]]
    traj = [tuple(v) for v in c["traj"]]
    axes = [(s[0], s[1], t[0], t[1]) for s, t in zip(shift, traj)]
    for depth in (1, 2, 3, 5, 40):
        with BatchPlan([cmf_spec(p, depth, 8, shift, traj)], full=True) as plan:
            res = plan.run()
        rr, Pf, ls = cref.walk(p, depth, D.generate_rns_primes(8), axes)
        got = res.full_shadow(0, 3)
        print("depth", depth, "\ngpu\n", got, "\noracle\n", Pf, flush=True)
        assert np.array_equal(res.full_matrix(0, 3), rr)
        for a, b in zip(got.reshape(-1), Pf.reshape(-1)):
            assert _close(a, b), (depth, a, b)


def test_cmf_golden_vectors(gpu_ctx, golden):
    for c in golden["run_cmf_walk"]:
        p = program_from_json(c["program"])
        shift = [tuple(v) for v in c["shift"]]
        traj = None if c["traj"] is None else [tuple(v) for v in c["traj"]]
        got = D.run_cmf_walk(p, c["depth"], c["K"], shift, traj)
        o = c["out"]
        print("cmf m=%d ops=%d" % (c["m"], len(c["program"]["ops"])), flush=True)
        assert got["p_residues"].tolist() == o["p_residues"]
        assert got["q_residues"].tolist() == o["q_residues"]
        for key in ("p_float", "q_float", "log_scale"):
            assert _close(got[key], float.fromhex(o[key])), key


def test_verify_pcf_readme_case(gpu_ctx, golden):
    g = golden["verify_pcf"][0]
    got = D.verify_pcf(*g["args"])
    ref = g["out"]
    assert got["p_bits"] == ref["p_bits"] == 992
    assert _close(got["delta_float"], float.fromhex(ref["delta_float"]))
    assert _close(got["delta_exact"], float.fromhex(ref["delta_exact"]))
    assert _close(got["est_float"], float.fromhex(ref["est_float"]))
    assert _close(got["log_q"], float.fromhex(ref["log_q"]))


def test_crt_matches_textbook_garner(gpu_ctx):
    rng = np.random.default_rng(7)
    for K in (1, 2, 8, 32, 63, 64, 65, 128, 129, 200, 256, 300):
        pr = [int(p) for p in D.generate_rns_primes(K)]
        for _ in range(5):
            res = [int(rng.integers(0, p)) for p in pr]
            x, M = D.crt_reconstruct(res, pr)
            assert (x, M) == ow.crt_reconstruct(res, pr) == exact.garner(res, pr)
        # round trip and signed values around the centre
        M = math.prod(pr)
        for v in (0, 1, M // 2, M // 2 + 1, M - 1, 12345678901234567890 % M):
            x, _ = D.crt_reconstruct([v % p for p in pr], pr)
            assert x == v
            assert D.centered(x, M) == (v - M if v > M // 2 else v)


def test_batch_pcf2_subset_against_c_oracle(gpu_ctx):
    cfg = synth.pcf2(n_walks=48)
    progs = [D.compile_pcf(w["a"], w["b"]) for w in cfg["walks"]]
    a0s = [D.pcf_initial_values(w["a"]) for w in cfg["walks"]]
    specs = [pcf_spec(p, a0, 2000, 32, 1, target=w["L"]) for p, a0, w in zip(progs, a0s, cfg["walks"])]
    with BatchPlan(specs, crt=True, top_k=16) as plan:
        res = plan.run()
    primes = D.generate_rns_primes(32)
    deltas = []
    for i, (p, a0, w) in enumerate(zip(progs, a0s, cfg["walks"])):
        rr, Pf, ls = cref.pcf_walk(p, a0, 2000, primes)
        d = res.walk_dict(i)
        assert d["p_residues"].tolist() == rr[0, 1].tolist(), i
        assert d["q_residues"].tolist() == rr[1, 1].tolist(), i
        assert _hex(d["p_float"]) == _hex(Pf[0, 1]) and _hex(d["q_float"]) == _hex(Pf[1, 1]), i
        assert _close(d["log_scale"], ls), i
        pb, qb = res.crt_values(i)
        xp, M = ow.crt_reconstruct(rr[0, 1].tolist(), primes.tolist())
        xq, _ = ow.crt_reconstruct(rr[1, 1].tolist(), primes.tolist())
        assert (pb, qb) == (ow.centered(xp, M), ow.centered(xq, M)), i
        dl, lq = ow.compute_dreams_delta_float(Pf[0, 1], Pf[1, 1], ls, w["L"])
        assert _close(res.out[i]["delta_float"], dl), i
        deltas.append(dl)
    order = sorted(range(len(deltas)), key=lambda j: (-deltas[j], j))[:16]
    assert res.topk.tolist() == order


def test_plain_mode_matches_batched(gpu_ctx):
    cfg = synth.pcf2(n_walks=6)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 700, 32)
             for w in cfg["walks"]]
    a = BatchPlan(specs, mode="plain").run()
    b = BatchPlan(specs, mode="batched", chunk_steps=256, threads=128).run()
    c = BatchPlan(specs, mode="batched", use_graph=True).run()
    for r in (b, c):
        assert np.array_equal(a.residues, r.residues)
        assert [_hex(v) for v in a.out["p_float"]] == [_hex(v) for v in r.out["p_float"]]
        assert [_hex(v) for v in a.out["log_scale"]] == [_hex(v) for v in r.out["log_scale"]]


def test_resident_run_then_fetch_equals_run(gpu_ctx):
    # drx_plan_run_resident leaves the outputs in HBM: the host buffers keep
    # their old contents until drx_plan_fetch, which then gives what
    # drx_plan_run gives (residues, CRT limbs and signs, FP64 results, top-k)
    cfg = synth.pcf2(n_walks=40)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 500, 32)
             for w in cfg["walks"]]
    ref = BatchPlan(specs, crt=True, top_k=8).run()
    with BatchPlan(specs, crt=True, top_k=8) as plan:
        r = plan.result
        r.residues[...] = -1
        plan.run(fetch=False)
        assert (r.residues == -1).all()
        plan.fetch()
        assert np.array_equal(r.residues, ref.residues)
        assert np.array_equal(r.limbs, ref.limbs) and np.array_equal(r.neg, ref.neg)
        assert np.array_equal(r.topk, ref.topk)
        assert [_hex(v) for v in r.out["delta_float"]] == [_hex(v) for v in ref.out["delta_float"]]
        r.residues[...] = -1
        plan.run()  # the one-call form still copies back
        assert np.array_equal(r.residues, ref.residues)


def test_short_and_ragged_depths_in_one_batch(gpu_ctx):
    # depth 0 (the initial matrix itself), 1, 2 and depths that end inside
    # a table-reduction period, beside a full-length walk, through the
    # forward-difference PCF kernel and the interpreter; each equals its
    # single-walk oracle run (residues bit-exact, FP64 bits equal)
    pcfs = [("3*n+1", "-n**2"), ("2", "n**2"), ("n**3 - 2*n", "5*n**6 + 1"), ("7", "-3*n**2+n")]
    depths = [0, 1, 2, 3, 5, 9, 13, 1000]
    specs, refs = [], []
    for i, d in enumerate(depths):
        a, b = pcfs[i % len(pcfs)]
        prog = D.compile_pcf(a, b)
        specs.append(pcf_spec(prog, D.pcf_initial_values(a), d, 16))
        refs.append(ow.run_pcf_walk(prog, D.pcf_initial_values(a), d, 16, 1))
    for interp in (False, True):
        res = BatchPlan(specs, interp_only=interp).run()
        for i, ref in enumerate(refs):
            d = res.walk_dict(i)
            assert d["p_residues"].tolist() == ref["p_residues"].tolist(), (interp, depths[i])
            assert d["q_residues"].tolist() == ref["q_residues"].tolist(), (interp, depths[i])
            assert _hex(d["p_float"]) == _hex(ref["p_float"]) and _hex(d["q_float"]) == _hex(ref["q_float"])


def test_inv_of_zero_is_zero(gpu_ctx):
    # b(n) = -1/(n+1) at shift -3: n + 1 = 0 at step 2 (INV(0) = 0 in residues, 0.0 in floats)
    p = D.compile_pcf("3", "-1/(n+1)")
    got = D.run_pcf_walk(p, 3, 60, 8, -3)
    ref = ow.run_pcf_walk(p, 3, 60, 8, -3)
    assert got["p_residues"].tolist() == ref["p_residues"].tolist()
    assert _hex(got["q_float"]) == _hex(ref["q_float"])


def test_shared_inversion_with_zero_arguments(gpu_ctx):
    # several step-dependent INVs share one inversion on the device; their
    # arguments vanish at steps 2 (x - 2), 7 (y + 3) and 4 (x*y + 24): INV(0)
    # must stay 0 for that argument only. The constant 1/3 is folded.
    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    mat = {(0, 0): "1/(x - 2)", (0, 1): "y/3 + 1/(y + 3)", (0, 2): "1",
           (1, 0): "x/(x*y + 24)", (1, 1): "0", (1, 2): "x - y",
           (2, 0): "1/(x - 2) + 1/(y + 3)", (2, 1): "1", (2, 2): "x*y/(x + 100)"}
    p = cf(mat, 3, 2, ["x", "y"], [1, 1])
    words, _, meta = D._native.pack_program(p)
    assert meta["inv_batched"] >= 3
    primes = D.generate_rns_primes(64)
    specs = [cmf_spec(p, depth, 64, [0, -10]) for depth in (3, 8, 300)]
    with BatchPlan(specs, full=True) as plan:
        res = plan.run()
    for i, s in enumerate(specs):
        rr, Pf, ls = cref.walk(p, s.depth, primes, [(0, 1, 1, 1), (-10, 1, 1, 1)])
        assert np.array_equal(res.full_matrix(i, 3), rr), i
        d = res.walk_dict(i)
        assert _close(d["p_float"], Pf[0, 2]) and _close(d["q_float"], Pf[2, 2]), i
        assert _close(d["log_scale"], ls), i


def test_folded_constant_inverse_runs_forward_differences(gpu_ctx):
    # a(n) = n/2 + 1/3, b(n) = -n^2/5: INVs of constants only, so the
    # program is a polynomial in n after folding and takes the
    # forward-difference kernel; it must equal the interpreter and the oracle
    p = D.compile_pcf("n/2 + 1/3", "-n**2/5")
    meta = D._native.pack_program(p)[2]
    assert meta["fd_ok"] == 1 and meta["fd_degree"] == 2 and meta["derived"]
    specs = [pcf_spec(p, D.pcf_initial_values("n/2 + 1/3"), d, 32) for d in (1, 9, 777)]
    a = BatchPlan(specs, full=True).run()
    b = BatchPlan(specs, full=True, interp_only=True).run()
    assert np.array_equal(a.full, b.full) and np.array_equal(a.full_f, b.full_f)
    ref = ow.run_pcf_walk(p, specs[2].a0, 777, 32, 1)
    d = a.walk_dict(2)
    assert d["p_residues"].tolist() == ref["p_residues"].tolist()
    assert d["q_residues"].tolist() == ref["q_residues"].tolist()
    assert _hex(d["p_float"]) == _hex(ref["p_float"]) and _hex(d["q_float"]) == _hex(ref["q_float"])


def test_cmf3_shapes_against_c_oracle(gpu_ctx):
    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    rng = np.random.default_rng(99)
    specs, refs = [], []
    primes = D.generate_rns_primes(64)
    for m in (3, 3, 4):
        spec = synth.cmf_walk_spec(rng, m, cf, 100, 250, True) if m == 3 else synth.cmf_walk_spec(rng, m, cf, 10, 400, False)
        p = spec["program"]
        specs.append(cmf_spec(p, 150, 64, spec["shift"]))
        axes = [(s, 1, d, 1) for s, d in zip(spec["shift"], spec["directions"])]
        refs.append(cref.walk(p, 150, primes, axes))
    with BatchPlan(specs, full=True) as plan:
        res = plan.run()
    for i, (rr, Pf, ls) in enumerate(refs):
        m = specs[i].program.m
        assert np.array_equal(res.full_matrix(i, m), rr), i
        d = res.walk_dict(i)
        assert _close(d["p_float"], Pf[0, m - 1]) and _close(d["q_float"], Pf[m - 1, m - 1]), i
        assert _close(d["log_scale"], ls), i


def test_forward_difference_path_equals_interpreter(gpu_ctx):
    # polynomial programs run on the forward-difference kernel; forcing the
    # interpreter must give the same residues and shadow, bit for bit
    cfg = synth.pcf2(n_walks=40, seed=77)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), d, K, 1, target=w["L"])
             for w, d, K in zip(cfg["walks"], [1, 2, 3, 5, 7, 64, 333, 2000] * 5, [8, 16, 32, 64, 31] * 8)]
    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    rng = np.random.default_rng(3)
    for _ in range(6):  # polynomial 3x3 CMFs (degree <= 3)
        mat = {(i, j): synth._mpoly_str(rng, ("x", "y", "z"), 1, 2, -9, 9) for i in range(3) for j in range(3)}
        p = cf(mat, 3, 3, ["x", "y", "z"], [1, -1, 2])
        specs.append(cmf_spec(p, int(rng.integers(10, 900)), 32, [(1, 3), 2, 0], [(1, 2), 1, -1]))
    a = BatchPlan(specs, full=True, crt=True).run()
    b = BatchPlan(specs, full=True, crt=True, interp_only=True).run()
    assert np.array_equal(a.full, b.full)
    assert np.array_equal(a.full_f, b.full_f)
    assert np.array_equal(a.limbs, b.limbs) and np.array_equal(a.neg, b.neg)
    primes = D.generate_rns_primes(64)
    for i in (0, 5, 7, 39, 40, 45):
        s = specs[i]
        rr, Pf, ls = cref.walk(s.program, s.depth, primes[: s.K], list(s.axes),
                               *((np.array([[1] * s.K, [s.a0 % int(q) for q in primes[: s.K]], [0] * s.K,
                                            [1] * s.K]), np.array([1.0, float(s.a0), 0.0, 1.0]))
                                 if s.init == "pcf" else (None, None)))
        m = s.program.m
        assert np.array_equal(a.full_matrix(i, m), rr), i


def _topk_ref(v, k):
    key = [(1, 0.0, i) if math.isnan(x) else (0, -x, i) for i, x in enumerate(v)]
    return [t[2] for t in sorted(key)[:k]]


def test_topk_tournament_and_rank_paths(gpu_ctx):
    # k <= 64: placement and merge rounds (one to three); larger k: several
    # tournament rounds, and above the tournament's reach the rank kernel;
    # NaN, +-0 and ties throughout
    rng = np.random.default_rng(21)
    for n, k in ((1, 1), (5, 5), (256, 64), (257, 64), (300, 7), (777, 64), (2048, 64), (2049, 1), (4095, 33),
                 (4096, 64), (4097, 64), (5000, 17),
                 (20000, 33), (100000, 1), (4096, 65), (4096, 1024), (70000, 1024), (300000, 64), (300000, 128),
                 (5000, 1025), (3000, 3000)):
        v = rng.normal(size=n).round(2)          # many exact ties
        v[rng.integers(0, n, size=max(1, n // 50))] = np.nan
        v[rng.integers(0, n, size=max(1, n // 100))] = -0.0
        v[rng.integers(0, n, size=max(1, n // 100))] = 0.0
        if n > 10:
            v[3] = np.inf
            v[7] = -np.inf
        got = gpu_ctx.topk(v, k).tolist()
        print("topk", n, k, flush=True)
        assert got == _topk_ref(v.tolist(), k), (n, k)


def test_plan_built_before_release_is_refused(gpu_ctx):
    cfg = synth.pcf2(n_walks=3, seed=31)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 50, 8) for w in cfg["walks"]]
    plan = BatchPlan(specs)
    first = plan.run().residues.copy()  # run() refills the plan's result arrays
    # a later plan with new programs grows the shared tables: the first plan still runs
    more = synth.pcf2(n_walks=40, seed=32)
    BatchPlan([pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 20, 8)
               for w in more["walks"]]).run()
    plan.result.residues[:] = 0
    assert np.array_equal(plan.run().residues, first)
    gpu_ctx.release_programs()
    with pytest.raises(Exception, match="changed since this plan was built"):
        plan.run()
    plan.close()
    assert np.array_equal(BatchPlan(specs).run().residues, first)


def _poly(rng, deg, lo=-20, hi=20, avoid01=False):
    """Polynomial string in n of exact degree `deg` (leading coefficient != 0;
    a constant avoids 0 and 1 when `avoid01`, so the entry stays general)."""
    while True:
        c = [int(v) for v in rng.integers(lo, hi + 1, size=deg + 1)]
        if c[-1] != 0 and not (avoid01 and deg == 0 and c[0] in (0, 1)):
            break
    return " + ".join(f"({v})*n**{i}" for i, v in enumerate(c))


def test_forward_difference_every_pcf_shape(gpu_ctx):
    # every (deg b, deg a) body of the companion-PCF kernel, with depths that
    # leave sub-chunks empty, short and unrolled-plus-remainder
    rng = np.random.default_rng(404)
    specs = []
    depths = [0, 1, 2, 7, 9, 31, 250, 1001]
    for db in range(7):
        for da in range(7):
            a, b = _poly(rng, da, avoid01=True), _poly(rng, db, avoid01=True)
            p = D.compile_pcf(a, b)
            specs.append(pcf_spec(p, D.pcf_initial_values(a), depths[(db * 7 + da) % len(depths)],
                                  int(rng.choice([8, 31, 32, 64])), int(rng.integers(-3, 4)), target=0.25))
    # companion PCFs with a static 0 or 1 entry (degree-0 tables)
    for a, b in (("0", "n**2 + 1"), ("n + 1", "1"), ("1", "0"), ("0", "1"), ("2*n - 3", "-1")):
        p = D.compile_pcf(a, b)
        specs.append(pcf_spec(p, D.pcf_initial_values(a), 333, 32, 2, target=0.5))
    # a 2x2 polynomial program that is not a companion PCF (generic kernel)
    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    mat = {(0, 0): "x**2 + 3", (0, 1): "2*x - 1", (1, 0): "x**3 - x", (1, 1): "5"}
    specs.append(cmf_spec(cf(mat, 2, 1, ["x"], [1]), 777, 32, [2]))
    a = BatchPlan(specs, full=True, crt=True).run()
    b = BatchPlan(specs, full=True, crt=True, interp_only=True).run()
    assert np.array_equal(a.full, b.full)
    assert np.array_equal(a.full_f, b.full_f)
    assert np.array_equal(a.limbs, b.limbs) and np.array_equal(a.neg, b.neg)
    primes = D.generate_rns_primes(64)
    for i in range(0, len(specs), 5):
        s = specs[i]
        init = None
        if s.init == "pcf":
            init = (np.array([[1] * s.K, [s.a0 % int(q) for q in primes[: s.K]], [0] * s.K, [1] * s.K]),
                    np.array([1.0, float(s.a0), 0.0, 1.0]))
        rr, Pf, ls = cref.walk(s.program, s.depth, primes[: s.K], list(s.axes), *(init or (None, None)))
        assert np.array_equal(a.full_matrix(i, s.program.m), rr), i


def test_autotune_choice_is_valid_and_results_unchanged(gpu_ctx):
    from dreams_rns.engine import autotune

    cfg = synth.pcf2(n_walks=300, seed=55)
    specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), 1500, 32, 1, target=w["L"])
             for w in cfg["walks"]]
    t = autotune(specs, max_walks=128, budget_s=10.0)
    assert t["threads"] in (64, 128, 256) and t["chunk_steps"] in (128, 256, 512, 1024, 2048, 4096, 8192, 16384)
    assert t["walks_per_wg"] in (1, 2, 4, 8) and t["walks_per_wg"] * 32 <= t["threads"]
    assert t["slice"] == 256  # 32-prime walks have nothing to slice
    assert t["sample_walk_ms"] > 0
    assert autotune(specs) is t  # cached by the batch signature
    with BatchPlan(specs, crt=True, top_k=8) as p0:
        r0 = p0.run()
        ref = (r0.residues.copy(), r0.limbs.copy(), r0.out["log_scale"].copy(), r0.topk.copy())
    with BatchPlan(specs, crt=True, top_k=8, tune=True) as p1:
        assert (p1.threads, p1.chunk_steps, p1.walks_per_wg) == (t["threads"], t["chunk_steps"], t["walks_per_wg"])
        r1 = p1.run()
        assert np.array_equal(r1.residues, ref[0]) and np.array_equal(r1.limbs, ref[1])
        assert [v.hex() for v in r1.out["log_scale"].tolist()] == [v.hex() for v in ref[2].tolist()]
        assert np.array_equal(r1.topk, ref[3])

"""Benchmark of the MI355X walk engine on its headline workload.

One step = one pass of the whole pcf2 batch through the GPU path: 4096 seeded
PCFs (deg a <= 3, deg b <= 6, coefficients in [-20, 20]), depth 2000, K = 32
primes, FP64 shadow with delta_float against a seeded target per walk, GPU
Garner CRT of p and q for every walk and top-64 by delta_float. Inputs are
packed and resident in HBM before the timed region; each step re-runs the
batch from the initial matrices.

Printed (one JSON line): residue-steps/s of the batched path (value), the
same with the copy of every output to the host, the plain
one-walk-per-launch GPU path and the CPU reference restatement on bounded
samples, the served stream (Poisson arrivals of serve5 requests swept over
offered loads: p50/p99 latency to CRT integers on the host, served rate,
saturation), the cmf3 and mix4 batches (one warm-up and two timed steps
each), and the roofline of the dominant kernel.
"""

from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "dreams-rns-cuda_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "residue-steps/sec batched+served vs plain (bit-exact CRT); p50/p99 latency"
# FP64 VALU lane-ops per second: 256 CUs x 4 SIMDs x 16 FP64 lanes x 2.4 GHz
# (rocminfo: 256 CUs, 2400 MHz), i.e. the 78.6 TFLOP/s FP64 vector peak
# counted in operations (an FMA is one)
PEAK_LANE_OPS = 256 * 64 * 2.4e9
# the GPU box gives one GPU a share of 16 host cores
BOX_CPU_SHARE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------- work model
def program_work(prog, vector: bool = False) -> float:
    """Bytecode-equivalent ops per residue-step: W = 6*MM + 2*MA with
    MM = #MUL + #POW2 + 2*#POW3 + 56*#INV + m^3 and MA = #ADD + #SUB + #NEG +
    m^2 (m - 1), counted on the walk's bytecode as the reference evaluates it
    every step; a state-vector walk's product is a matrix-vector product
    (m^2, m (m - 1)). Instructions whose operands are all constants give the
    same value every step and are not counted."""
    cnt = {}
    const_regs = set()
    for ins in prog.opcodes:
        op, a0, a1, a2 = int(ins.op), ins.arg0, ins.arg1, ins.arg2
        if op == 2:  # LOAD_C
            const_regs.add(a1)
            continue
        if op in (4, 5, 6):  # ADD SUB MUL
            if a0 in const_regs and a1 in const_regs:
                const_regs.add(a2)
                continue
            const_regs.discard(a2)
        elif op in (7, 8, 9, 10):  # NEG POW2 POW3 INV
            if a0 in const_regs:
                const_regs.add(a1)
                continue
            const_regs.discard(a1)
        elif op == 1:  # LOAD_X
            const_regs.discard(a1)
        elif op == 3:  # LOAD_N
            const_regs.discard(a0)
        cnt[op] = cnt.get(op, 0) + 1
    m = prog.m
    mm = cnt.get(6, 0) + cnt.get(8, 0) + 2 * cnt.get(9, 0) + 56 * cnt.get(10, 0) + (m * m if vector else m ** 3)
    ma = cnt.get(4, 0) + cnt.get(5, 0) + cnt.get(7, 0) + (m * (m - 1) if vector else m * m * (m - 1))
    return 6.0 * mm + 2.0 * ma


def walk_kernel_name(progs) -> str:
    """The walk kernel the library picks for most of these programs, with the
    others and their walk counts."""
    from collections import Counter

    from dreams_rns._native import pack_program

    names = Counter()
    for p in progs:
        meta = pack_program(p)[2]
        if meta["fd_ok"] and p.m == 2 and meta["fd_degree"] <= 6:
            k = meta["kinds"]
            pcf = (k & 3) == 0 and ((k >> 4) & 3) == 1  # entry (0,0) zero, (1,0) one
            names["k_walk_fd_pcf" if pcf else "k_walk_fd<2>"] += 1
        elif meta["fd_ok"] and p.m in (3, 4) and meta["fd_degree"] <= 3 and all(
                ((meta["kinds"] >> (2 * (i * p.m + j))) & 3) == (1 if i == j + 1 else 0)
                for i in range(p.m) for j in range(p.m - 1)):
            names[f"k_walk_fd_comp<{p.m}>"] += 1  # companion: sub-diagonal ones, polynomial last column
        elif meta["fd_ok"] and p.m == 3 and meta["fd_degree"] <= 3:
            names["k_walk_fd<3>"] += 1
        else:
            names[f"k_walk_chunk<{p.m}>"] += 1
    (top, _), rest = names.most_common(1)[0], names.most_common()[1:]
    return top + "".join(f" (+{k} for {v} walks)" for k, v in rest)


def measured_counters(kernel: str, workload: str):
    """Counter summary of `kernel` on `workload` committed beside the rocprofv3
    summaries: HBM bytes per launch (FETCH_SIZE / WRITE_SIZE scaled by the
    copy calibration) and VALU instructions per launch, or None."""
    name = "walk_kernel_counters.json" if workload == "pcf2" else f"walk_kernel_counters_{workload}.json"
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel.split(" ")[0] or t.get("workload") != workload:
        return None
    return t


def roofline(stats, walk_ms, equiv_work, kname, workload):
    """The walk kernel against the FP64 VALU peak. `achieved` counts the
    arithmetic the chosen kernels execute, from the library's own operation
    count of their code (drx_plan_stats: 6 FP64 ops per modular multiply --
    product, error term, quotient, rounding, two corrections -- and one per
    lazy addition), over the kernels' device time measured with HIP events on
    their stream. The committed counter summary of the same kernel and
    workload (when present) gives HBM traffic and SQ_INSTS_VALU x 64 per
    launch as a cross-check. `equivalent` is the bytecode convention W of
    the reference's per-step evaluation, which forward differences skip."""
    if not walk_ms == walk_ms or walk_ms <= 0:
        return None
    executed = 6.0 * float(stats["modmul"]) + float(stats["modadd"])
    ach = executed / (walk_ms * 1e-3)
    cnt = measured_counters(kname, workload)
    out = {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK_LANE_OPS / 1e12, "unit": "TOP/s",
           "frac": ach / PEAK_LANE_OPS, "traffic": float(cnt["traffic_bytes"]) if cnt else None,
           "kernel": kname, "kernel_ms": walk_ms, "work_per_launch": executed,
           "work_model": "executed FP64 lane-ops of the chosen kernels: 6 per modular multiply, 1 per lazy "
                         "addition (library count, drx_plan_stats)",
           "equivalent": {"achieved": equiv_work / (walk_ms * 1e-3) / 1e12,
                          "frac": equiv_work / (walk_ms * 1e-3) / PEAK_LANE_OPS, "work_per_launch": equiv_work,
                          "model": "W = 6*MM + 2*MA per residue-step of per-step bytecode evaluation (the "
                                   "reference's algorithm); forward differences skip that evaluation, so this "
                                   "rate can pass the peak"}}
    if cnt:
        v = float(cnt["valu_insts"]) * 64
        out["counters"] = {"valu_lane_ops_per_launch": v, "frac": v / (walk_ms * 1e-3) / PEAK_LANE_OPS,
                           "source": cnt["source"], "commit": cnt.get("commit")}
    return out


# ------------------------------------------------------------ CPU baseline
def cpu_share():
    """Cores this process may run on, at most the box's share for one GPU."""
    try:
        visible = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        visible = os.cpu_count() or 1
    return min(BOX_CPU_SHARE, visible), visible


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_walk(args):
    import dreams_rns as D
    from oracle import walk as ow

    a, b, depth, K = args
    prog = D.compile_pcf(a, b)
    res = ow.run_pcf_walk(prog, D.pcf_initial_values(a), depth, K)
    primes = [int(p) for p in res["primes"]]
    for key in ("p_residues", "q_residues"):
        x, M = ow.crt_reconstruct([int(v) for v in res[key]], primes)
        ow.centered(x, M)
    return depth * K


def cpu_baseline(walks, depth, K, n_one, n_all, cores):
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    jobs = [(w["a"], w["b"], depth, K) for w in walks]
    t0 = time.perf_counter()
    done = sum(_cpu_walk(j) for j in jobs[:n_one])
    one = done / (time.perf_counter() - t0)
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        pool.map(_cpu_walk, jobs[:cores])  # warm the workers
        t0 = time.perf_counter()
        done = sum(pool.map(_cpu_walk, jobs[:n_all], chunksize=1))
        allc = done / (time.perf_counter() - t0)
    return one, allc


# ------------------------------------------------------------------ served
def served_sweep(rates, seconds, mem_budget, seed_base=1005, threads=0, round_steps=1000):
    """The served stream: Poisson arrivals of serve5 requests (m in
    {2, 3, 4}, depth log-U[1000, 100000], K in {32, 64, 128}) at each offered
    rate for `seconds`, then the queue drained. Per rate: p50/p99 latency
    (submission to CRT integers on the host), served and offered
    residue-steps/s."""
    from dreams_rns import serve5, synth
    from dreams_rns.server import WalkServer

    pool = serve5.program_pool(16)
    per_rate = []
    with WalkServer(round_steps=round_steps, threads=threads, mem_budget=mem_budget) as srv:
        slots = srv.capacity()
        warm = [serve5.request_spec(dict(r, depth=1000, K=128, L=0.0)) for m in pool for r in pool[m]]
        serve5.run_stream(srv, warm, np.linspace(0.0, 0.05, len(warm)), timeout=120)  # uploads every program
        for rate in rates:
            arr = synth.serve5_arrivals(rate, seconds)
            rng = np.random.default_rng(seed_base + int(rate))
            specs = [serve5.request_spec(serve5.draw_request(rng, pool)) for _ in arr]
            st = serve5.run_stream(srv, specs, arr, timeout=300)["stats"]
            st["rate_req_s"] = float(rate)
            per_rate.append(st)
            log(f"served {rate:g} req/s: p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms, served "
                f"{st['served_residue_steps_per_s']:.3e} of {st['offered_residue_steps_per_s']:.3e} residue-steps/s")
    keep = [s for s in per_rate if s["served_residue_steps_per_s"] >= 0.9 * s["offered_residue_steps_per_s"]]
    top = max(per_rate, key=lambda s: s["rate_req_s"])
    cap = top["served_residue_steps_per_s"]
    return {
        "workload": "serve5: Poisson arrivals, m in {2,3,4} (pcf2-shape PCFs, 3x3/4x4 CMFs of 10-400 ops), depth "
                    "log-U[1000,100000], K in {32,64,128}; programs from a seeded pool of 48 generator draws",
        "seconds_per_rate": seconds, "slots": slots, "threads": threads or 256, "round_steps": round_steps,
        "rates": per_rate,
        "saturation_req_s": max((s["rate_req_s"] for s in keep), default=None),
        "capacity_residue_steps_per_s": cap,
        "capacity_req_s": cap / top["mean_residue_steps_per_request"],
        "saturation_rule": "highest offered rate whose served rate is within 10% of the offered one; capacity = "
                           "served rate at the highest offered rate (queue never empty)",
    }


# ------------------------------------------------------- batch workloads
def build_workload(name, n_walks=None):
    """(specs, top_k, description, extra) of a batch workload."""
    import dreams_rns as D
    from dreams_rns import synth
    from dreams_rns.engine import cmf_spec, pcf_spec

    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    extra = {}
    if name == "pcf2":
        cfg = synth.pcf2(n_walks=n_walks or 4096)
        specs = [pcf_spec(D.compile_pcf(w["a"], w["b"]), D.pcf_initial_values(w["a"]), cfg["depth"], cfg["K"], 1,
                          target=w["L"]) for w in cfg["walks"]]
        extra["cfg"] = cfg
        desc = (f"pcf2: {len(specs)} seeded PCFs, deg a<=3, deg b<=6, coef [-20,20], depth {cfg['depth']}, "
                f"K={cfg['K']}, FP64 shadow + delta_float, GPU Garner CRT, top-{cfg['top_k']}")
        return specs, cfg["top_k"], desc, extra
    if name == "zeta5":
        from dreams_rns.engine import vec_spec

        cfg = synth.zeta5_sweep(cf)
        specs = [vec_spec(cfg["program"], cfg["depth"], cfg["K"], sh, cfg["v0"], cfg["acc"], cfg["const"], tr)
                 for sh in cfg["shifts"] for tr in cfg["trajs"]]
        extra["cfg"] = cfg
        desc = (f"zeta5: {len(specs)} state-vector walks of the zeta(5) CMF (4x4, 6 axes), "
                f"{len(cfg['shifts'])} shifts x {len(cfg['trajs'])} trajectories, depth {cfg['depth']}, K={cfg['K']}")
        return specs, 64, desc, extra
    if name == "sweep3f2":
        from dreams_rns.sweep import CmfSweep

        with open(os.path.join(ROOT, "tests", "golden", "sweep_3f2.json")) as f:
            fix = json.load(f)
        spec = fix["specs"][0]
        depth, K = 2000, 1000  # the sweep driver's defaults for CMF mode
        sw = CmfSweep([spec], fix["bench_trajectories"], fix["bench_shifts"], depth=depth, K=K)
        rows = list(sw._rows())
        specs = [cmf_spec(prog, depth, K, so) for (_, _, prog, so, _) in rows]
        extra.update(fix=fix, spec=spec, rows=rows, sweep_walks=sw.total_walks)
        desc = (f"sweep3f2: {spec['name']} x {len(fix['bench_trajectories'])} trajectories x "
                f"{len(fix['bench_shifts'])} shifts ({sw.total_walks} sweep walks, {len(specs)} distinct starts), "
                f"4x4 per-axis companion, depth {depth}, K={K}")
        return specs, 0, desc, extra
    if name == "cmf3":
        cfg = synth.cmf3(cf, n_walks=n_walks or 1024)
        specs = [cmf_spec(w["program"], cfg["depth"], cfg["K"], w["shift"], target=0.25) for w in cfg["walks"]]
        desc = f"cmf3: {len(specs)} 3x3 CMFs (100-250 ops with INV), depth {cfg['depth']}, K={cfg['K']}"
        return specs, 64, desc, extra
    cfg = synth.mix4(cf, n_walks=n_walks or 16384)
    specs = []
    for w in cfg["walks"]:
        if w["kind"] == "pcf":
            p = D.compile_pcf(w["a"], w["b"])
            specs.append(pcf_spec(p, D.pcf_initial_values(w["a"]), w["depth"], w["K"], 1, target=w["L"]))
        else:
            specs.append(cmf_spec(w["program"], w["depth"], w["K"], w["shift"], target=w["L"]))
    desc = f"mix4: {len(specs)} uneven walks (2x2/3x3/4x4, depth 500..50000, K 16..128), top-{cfg['top_k']}"
    return specs, cfg["top_k"], desc, extra


def timed_steps(plan, steps, warmup, label):
    """Whole-batch runs with the outputs left in HBM: step wall times and
    walk-kernel device times (ms; none for a graph replay)."""
    for _ in range(warmup):
        plan.run()
    step_ms, walk_ms = [], []
    for i in range(steps):
        t = time.perf_counter()
        plan.run(fetch=False)
        step_ms.append((time.perf_counter() - t) * 1e3)
        tm = plan.ctx.last_timing()
        if plan.use_graph:
            log(f"{label} step {i}: {step_ms[-1]:.2f} ms (graph replay, device {tm[4]:.3f} ms)")
            continue
        walk_ms.append(tm[0])
        log(f"{label} step {i}: {step_ms[-1]:.2f} ms (walk kernels {tm[0]:.3f} ms, combine {tm[1]:.3f}, "
            f"CRT and shadow wait {tm[2]:.3f}, tail {tm[3]:.3f}, device {tm[4]:.3f})")
    return step_ms, walk_ms


def timed_with_copy(plan, steps: int) -> list:
    """Wall time (ms) of whole runs that also copy every output to pinned
    host memory: the PCIe-inclusive rate, reported beside the HBM-resident
    one."""
    out = []
    for _ in range(max(1, min(steps, 5))):
        t = time.perf_counter()
        plan.run()
        out.append((time.perf_counter() - t) * 1e3)
    log(f"with the copy back: {np.mean(out):.3f} ms per step")
    return out


def with_copy(rsteps: float, copy_ms: list) -> dict:
    ms = float(np.mean(copy_ms))
    return {"ms_per_step": ms, "value": rsteps / (ms * 1e-3), "unit": "residue-steps/s",
            "what": "the same step plus the copy of every output (residues, CRT limbs, FP64 results, top-k) "
                    "to pinned host memory; such a run launches on streams, the residues and limbs copied "
                    "while the shadow chain still runs"}


def batch_rate(name, steps=2, warmup=1, threads=256, chunk=256):
    """One batch workload timed in the default line: residue-steps/s with
    the outputs in HBM (plan built before the timed region), at the launch
    shape the autotuner picks for cmf3 and mix4 (256-lane workgroups,
    256-step tasks)."""
    from dreams_rns.engine import BatchPlan

    t0 = time.perf_counter()
    specs, top_k, desc, _ = build_workload(name)
    gen_s = time.perf_counter() - t0
    rsteps = float(sum(s.depth * s.K for s in specs))
    equiv = sum(program_work(s.program, s.vector) * s.depth * s.K for s in specs)
    with BatchPlan(specs, crt=True, top_k=top_k, threads=threads, chunk_steps=chunk) as plan:
        step_ms, walk_ms = timed_steps(plan, steps, warmup, name)
        st = plan.stats()
    ms = float(np.mean(step_ms))
    kname = walk_kernel_name([s.program for s in specs])
    return {"workload": desc, "value": rsteps / (ms * 1e-3), "unit": "residue-steps/s", "ms_per_step": ms,
            "steps": steps, "warmup": warmup, "threads": threads, "chunk_steps": chunk, "generate_s": gen_s,
            "roofline": roofline(st, float(np.mean(walk_ms)), equiv, kname, name)}


def workload_line(args):
    """--workload cmf3 | mix4 | zeta5 | sweep3f2: one JSON line in the same
    form (a CPU leg for zeta5 / sweep3f2)."""
    from dreams_rns.engine import BatchPlan, autotune

    t0 = time.perf_counter()
    specs, top_k, desc, extra = build_workload(args.workload, args.walks)
    log(f"generated {len(specs)} walks in {time.perf_counter() - t0:.1f}s")
    cpu = None
    if not args.no_cpu and args.workload == "zeta5":
        from oracle import walk as ow

        cfg = extra["cfg"]
        t = time.perf_counter()
        nw = 0
        for sh, tr in zip(cfg["shifts"], cfg["trajs"]):
            ow.run_cmf_walk_vec(cfg["program"], cfg["depth"], cfg["K"], sh, cfg["v0"], cfg["acc"], cfg["const"], tr)
            nw += 1
            if time.perf_counter() - t > 10.0:
                break
        el = time.perf_counter() - t
        cpu = {"value": nw * cfg["depth"] * cfg["K"] / el, "unit": "residue-steps/s", "cores": 1, "kind": "port",
               "cpu_model": cpu_model(), "sample": f"{nw} walks of the sweep through the numpy restatement, one core"}
    elif not args.no_cpu and args.workload == "sweep3f2":
        from oracle import walk as ow

        K = specs[0].K
        t = time.perf_counter()
        nw = 0
        for (_, _, prog, so, _) in extra["rows"]:
            ow.run_cmf_walk(prog, 50, K, so)
            nw += 1
            if time.perf_counter() - t > 10.0:
                break
        el = time.perf_counter() - t
        cpu = {"value": nw * 50 * K / el, "unit": "residue-steps/s", "cores": 1, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"{nw} walks of the sweep at depth 50 (the per-step cost does not depend on the depth) "
                         f"through the numpy restatement, one core"}
    if cpu:
        log(f"cpu: {cpu['value']:.3e} residue-steps/s")
    tuned = None
    if args.workload in ("zeta5", "sweep3f2"):
        args.threads = args.threads or 256
        args.chunk = args.chunk or 4096
    if not (args.threads or args.chunk):
        t = time.perf_counter()
        tuned = autotune(specs, budget_s=60.0)
        args.threads, args.chunk = tuned["threads"], tuned["chunk_steps"]
        args.walks_per_wg, args.slice = tuned["walks_per_wg"], tuned["slice"]
        log(f"autotuned in {time.perf_counter() - t:.1f}s: threads {args.threads}, chunk {args.chunk}, "
            f"walks per workgroup {args.walks_per_wg}, slice {args.slice}")
    rsteps = float(sum(s.depth * s.K for s in specs))
    equiv = sum(program_work(s.program, s.vector) * s.depth * s.K for s in specs)
    plan = BatchPlan(specs, crt=True, top_k=top_k, threads=args.threads, chunk_steps=args.chunk,
                     use_graph=args.graph, mode=args.mode, walks_per_wg=args.walks_per_wg, slice=args.slice)
    log(f"plan built: {time.perf_counter() - t0:.1f}s since start")
    step_ms, walk_ms = timed_steps(plan, args.steps, args.warmup, args.workload)
    copy_ms = timed_with_copy(plan, args.steps)
    post = None
    if args.workload == "sweep3f2":
        # the step after the walk: the sweep's hit selection over this batch
        from dreams_rns.constants import load_constants
        from dreams_rns.postwalk import sweep_hits
        from dreams_rns.sweep import SweepBatch

        rows, fix, spec = extra["rows"], extra["fix"], extra["spec"]
        bank = load_constants(dps=200)
        sb = SweepBatch(spec_idx=np.array([r[0] for r in rows], dtype=np.int32),
                        traj_idx=np.array([r[1] for r in rows], dtype=np.int32), offsets=[r[3] for r in rows],
                        shift_rows=[r[4] for r in rows], result=plan.run(), crt=True)
        t = time.perf_counter()
        hits = sweep_hits(sb, [spec], fix["bench_trajectories"], bank, proximity=1e-3, dps=200)
        post = {"ms": (time.perf_counter() - t) * 1e3, "hits": len(hits), "walks": extra["sweep_walks"],
                "what": "a100_sweep.py:426-505 hit selection (CRT check, proximity, delta against the 40-constant "
                        "bank) over the batch's results, on the host"}
        log(f"hit selection: {post['ms']:.1f} ms, {len(hits)} hits")
        if not args.no_cpu:
            from oracle import postwalk as opw

            t = time.perf_counter()
            ns = 0
            while ns < len(rows) and time.perf_counter() - t < 5.0:
                opw.walk_record(sb.result.walk_dict(ns), bank, 1e-3, 200)
                ns += 1
            per = (time.perf_counter() - t) / ns
            post["cpu_loop_ms_per_walk"] = per * 1e3
            post["cpu_loop_sample"] = (f"{ns} walks of the batch through the walk-by-walk restatement of the "
                                       "reference loop (Python CRT, mpmath), one core")
    st = plan.stats()
    plan.close()
    ms = float(np.mean(step_ms))
    walk_avg = float(np.mean(walk_ms)) if not args.graph else float("nan")
    kname = "k_walk_vec" if args.workload == "zeta5" else walk_kernel_name([s.program for s in specs])
    print(json.dumps({
        "metric": METRIC, "value": rsteps / (ms * 1e-3), "unit": "residue-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc, "walks": len(specs), "top_k": top_k, "graph": bool(args.graph), "mode": args.mode,
                   "threads": args.threads, "chunk_steps": args.chunk, "walks_per_wg": args.walks_per_wg or 1,
                   "slice": args.slice or 256, "autotuned": tuned is not None},
        "latency_batch_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99))},
        "with_copy_back": with_copy(rsteps, copy_ms),
        "roofline": roofline(st, walk_avg, equiv, kname, args.workload),
        "cpu_baseline": cpu,
        **({"sweep_walks_per_s": extra["sweep_walks"] / (ms * 1e-3), "hit_selection": post}
           if args.workload == "sweep3f2" else {}),
    }))


# -------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--walks", type=int, default=0)
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--cpu-one", type=int, default=32)
    ap.add_argument("--cpu-all", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the cmf3 / mix4 batches of the default line")
    ap.add_argument("--serve-rates", default="100,1000,3000,10000")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-budget-gib", type=float, default=4.0)
    ap.add_argument("--serve-threads", type=int, default=0, help="server walk workgroup size (0: 256)")
    ap.add_argument("--serve-round-steps", type=int, default=1000)
    ap.add_argument("--serve-only", action="store_true", help="only the served stream (no batch lines)")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--mode", choices=("batched", "persistent"), default="batched")
    ap.add_argument("--graph", action="store_true",
                    help="replay the batch as a hipGraph (the default of the pcf2 line; other workloads launch "
                         "on streams unless given)")
    ap.add_argument("--no-graph", action="store_true", help="pcf2: launch on streams instead of a graph replay")
    ap.add_argument("--walks-per-wg", type=int, default=0, help="walks per workgroup (0: one, or tuned)")
    ap.add_argument("--slice", type=int, default=0, help="primes per task of wide walks (0: 256, or tuned)")
    ap.add_argument("--chain-cus", type=int, default=0,
                    help="CUs reserved for the FP64 M_f pass and shadow (CU-masked streams); 0 = shared")
    ap.add_argument("--tune", action="store_true", help="pcf2: autotune the launch shape first")
    ap.add_argument("--workload", choices=("pcf2", "cmf3", "mix4", "zeta5", "sweep3f2"), default="pcf2",
                    help="pcf2 (the headline line); cmf3 / mix4 / zeta5 / sweep3f2 measure the other batch shapes")
    ap.add_argument("--rates", action="store_true",
                    help="print the VALU issue rates (drx_valu_rate) behind the FP64 residue arithmetic and exit")
    ap.add_argument("--profile", action="store_true",
                    help="for counter collection: run the batch --steps times and a 512 MiB device copy "
                         "(the calibration kernel for FETCH_SIZE / WRITE_SIZE), print nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(walks shard trivially across GPUs, but only the one-GPU path is built)", file=sys.stderr)
        sys.exit(2)

    import dreams_rns as D
    from dreams_rns.engine import BatchPlan

    t0 = time.perf_counter()
    if args.rates:
        from dreams_rns import _native as N

        ctx = N.context()
        names = ["FP64 FMA", "32-bit multiply-high", "32x32+64-bit multiply-add", "residue modmul mm()"]
        for op, name in enumerate(names):
            r, pk = ctx.valu_rate(op)
            ops = 6 if op == 3 else 1
            print(json.dumps({"op": name, "per_s": r, "lane_ops_per_s": r * ops, "peak_lane_ops_per_s": pk,
                              "frac_of_lane_peak": r * ops / pk, "lane_ops_each": ops}))
        return
    if args.serve_only:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        print(json.dumps(served_sweep(rates, args.serve_seconds, int(args.serve_budget_gib * (1 << 30)),
                                      threads=args.serve_threads, round_steps=args.serve_round_steps)))
        return
    if args.workload != "pcf2" and not args.profile:
        return workload_line(args)
    if args.profile and args.workload != "pcf2":
        specs, top_k, _, _ = build_workload(args.workload, args.walks)
        with BatchPlan(specs, crt=True, top_k=top_k, threads=args.threads or 256, chunk_steps=args.chunk or 512,
                       mode=args.mode) as plan:
            for _ in range(args.steps):
                plan.run()
            ms = plan.ctx.copy_calibrate(512 << 20, reps=1)
        log("profile run done: %s x%d, copy 512 MiB (%.3f ms)" % (args.workload, args.steps, ms))
        return
    specs, top_k, desc, extra = build_workload("pcf2", args.walks)
    cfg = extra["cfg"]
    depth, K = cfg["depth"], cfg["K"]
    progs = [s.program for s in specs]
    equiv = sum(program_work(p) * depth * K for p in progs)
    rsteps = float(len(specs) * depth * K)
    log(f"compiled {len(progs)} PCFs in {time.perf_counter() - t0:.1f}s")

    tuned = None
    if args.tune and not (args.threads or args.chunk):
        from dreams_rns.engine import autotune

        t = time.perf_counter()
        tuned = autotune(specs)
        args.threads, args.chunk = tuned["threads"], tuned["chunk_steps"]
        args.walks_per_wg, args.slice = tuned["walks_per_wg"], tuned["slice"]
        log(f"autotuned in {time.perf_counter() - t:.1f}s: threads {args.threads}, chunk {args.chunk}, "
            f"walks per workgroup {args.walks_per_wg}")
    # without --tune the line uses the shape the autotuner picks for this
    # batch (64 lanes, 4096 steps per task: the faster 256-lane walk kernel
    # crowds out the FP64 M_f / shadow chain that sets the step time), fixed
    # so that the walk kernel's time matches its committed kernel trace
    args.threads, args.chunk = args.threads or 64, args.chunk or 4096
    # the step is replayed as one hipGraph (0.65-0.67 ms against 0.68 launched
    # on streams); CU-masked streams are a stream-launch option
    args.graph = not args.no_graph and args.chain_cus == 0
    plan = BatchPlan(specs, crt=True, top_k=top_k, threads=args.threads, chunk_steps=args.chunk,
                     use_graph=args.graph, mode=args.mode, chain_cus=args.chain_cus,
                     walks_per_wg=args.walks_per_wg, slice=args.slice)
    log(f"plan built: {time.perf_counter() - t0:.1f}s since start")
    ctx = plan.ctx
    if args.profile:
        for _ in range(args.steps):
            plan.run()
        plan.close()
        ms = ctx.copy_calibrate(512 << 20, reps=1)
        log("profile run done: batch x%d, copy 512 MiB (%.3f ms)" % (args.steps, ms))
        return
    step_ms, walk_ms = timed_steps(plan, args.steps, args.warmup, "pcf2")
    copy_ms = timed_with_copy(plan, args.steps)
    res = plan.result
    ms = float(np.mean(step_ms))
    value = rsteps / (ms * 1e-3)
    if args.graph:
        # a graph replay has no events between its kernels: the walk kernel's
        # launch time for the roofline comes from the same batch launched on
        # streams (the same kernels, the same shapes)
        with BatchPlan(specs, crt=True, top_k=top_k, threads=args.threads, chunk_steps=args.chunk,
                       mode=args.mode, walks_per_wg=args.walks_per_wg, slice=args.slice) as splan:
            _, walk_ms = timed_steps(splan, min(args.steps, 5), 1, "pcf2 (streams, for the kernel time)")
    walk_avg = float(np.mean(walk_ms))
    st = plan.stats()
    # bit-exactness spot check of the timed output against the C restatement
    from oracle import cref

    primes = D.generate_rns_primes(K)
    check = [0, len(specs) // 2, len(specs) - 1]
    for i in check:
        rr, _, _ = cref.pcf_walk(specs[i].program, specs[i].a0, depth, primes)
        assert res.residues[i, 0, :K].tolist() == rr[0, 1].tolist(), f"walk {i} p residues differ"
        assert res.residues[i, 1, :K].tolist() == rr[1, 1].tolist(), f"walk {i} q residues differ"
    log(f"spot-checked walks {check} against the C restatement: bit-exact")
    plan.close()

    # plain comparator: one walk per launch sequence, synchronous
    n_plain = min(args.plain_sample, len(specs))
    pplan = BatchPlan(specs[:n_plain], mode="plain", crt=True)
    pplan.run()
    t = time.perf_counter()
    pplan.run()
    plain_s = time.perf_counter() - t
    plain = n_plain * depth * K / plain_s
    pplan.close()
    log(f"plain: {n_plain} walks in {plain_s * 1e3:.1f} ms -> {plain:.3e} residue-steps/s")

    served = None
    if not args.no_serve:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        served = served_sweep(rates, args.serve_seconds, int(args.serve_budget_gib * (1 << 30)),
                              threads=args.serve_threads, round_steps=args.serve_round_steps)

    extra_lines = {}
    if not args.no_extra:
        for name in ("cmf3", "mix4"):
            extra_lines[name] = batch_rate(name)
            log(f"{name}: {extra_lines[name]['value']:.3e} residue-steps/s, {extra_lines[name]['ms_per_step']:.1f} ms")

    cpu = None
    if not args.no_cpu:
        cores, visible = cpu_share()
        one, allc = cpu_baseline(cfg["walks"], depth, K, args.cpu_one, args.cpu_all, cores)
        cpu = {"value": allc, "unit": "residue-steps/s", "cores": cores, "kind": "port", "one_core": one,
               "cpu_model": cpu_model(), "cpus_visible": visible,
               "sample": f"{args.cpu_all} pcf2 walks ({cores} processes, one thread each) and {args.cpu_one} (one "
                         f"core) at depth {depth}, K={K}: numpy restatement of run_pcf_walk + crt_reconstruct + "
                         f"centered"}
        log(f"cpu ({cpu['cpu_model']}): one core {one:.3e}, {cores} cores {allc:.3e} residue-steps/s")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "residue-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "walks": len(specs), "depth": depth, "K": K, "top_k": top_k,
                   "threads": args.threads, "chunk_steps": args.chunk, "graph": bool(args.graph),
                   "mode": args.mode, "chain_cus": args.chain_cus, "walks_per_wg": args.walks_per_wg or 1,
                   "slice": args.slice or 256, "autotuned": tuned is not None},
        "latency_batch_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99))},
        "with_copy_back": with_copy(rsteps, copy_ms),
        "plain": {"value": plain, "sample_walks": n_plain, "x": value / plain},
        "served": served,
        **extra_lines,
        "roofline": roofline(st, walk_avg, equiv, walk_kernel_name(progs), "pcf2"),
        "cpu_baseline": cpu,
    }
    if cpu:
        line["x_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()

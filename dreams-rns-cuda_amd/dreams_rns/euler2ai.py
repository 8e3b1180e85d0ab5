"""Batch verification of PCF lists against their stated limits, with the
Euler2AI CSV report (python/euler2ai_full_verify.py).

The reference walks one PCF at a time on the host; here every distinct PCF
of the list is walked in device batches (walk, shadow, CRT on the GPU) and
only the per-PCF delta (mpmath) and the report rows are host work. Rows,
their order, their formatting and the caching of repeated PCFs follow
euler2ai_full_verify.py:186-420, the CSV columns are its ``fieldnames``.
"""

from __future__ import annotations

import csv
import json
import math
import os
import time
from typing import Any, Dict, List, Optional, Sequence, Tuple

from . import _native as N
from .engine import BatchPlan, pcf_spec
from .postwalk import compute_delta, compute_delta_float, crt_overflowed, parse_limit
from .runner import compile_pcf, pcf_initial_values

__all__ = ["FIELDNAMES", "estimate_K", "load_cmf_pcfs", "expand_tasks", "verify_pcf_records", "write_csv"]

# The report's columns (euler2ai_full_verify.py:244-250).
FIELDNAMES = ["task_idx", "pcf_idx", "source_idx", "a", "b", "limit", "trajectory", "shift", "delta_ref_cpu",
              "convergence_rate_ref", "delta_cuda_rns", "delta_method", "est_float", "delta_diff", "match",
              "p_bits", "depth", "K"]


def _poly_size(exprs, n):
    """(degree, coefficient bound) of the PCF's polynomials in ``n``: the
    largest degree (at least 1) and the largest |integer coefficient|, with
    100 standing in for an expression whose coefficients cannot be read as
    integers (the reference's rule). A degree that cannot be taken raises."""
    import sympy as sp

    deg = max([1] + [int(sp.degree(e, n)) for e in exprs if e.has(n)])
    bound = 1
    for e in exprs:
        try:
            for c in sp.Poly(e, n).all_coeffs():
                bound = max(bound, abs(int(c)))
        except Exception:
            bound = max(bound, 100)
    return deg, bound


def estimate_K(a_str: str, b_str: str, depth: int, safety: float = 1.5) -> int:
    """Primes for an exact CRT of a depth-``depth`` PCF walk: the same number
    euler2ai_full_verify.py:40-92 computes.

    |q_N| grows like prod_n |M(n)|, i.e. about N!^d * c^N for polynomials of
    degree d with coefficients up to c, so the walk needs about
    d * log2(N!) + N * log2(c + 1) bits (log2(N!) by Stirling's series to the
    1/2 log(2 pi N) term; a coefficient bound of 1 adds nothing). That times
    ``safety``, plus 100 bits, in 31-bit primes, at least 64. When the
    polynomials cannot be read, a degree-4 bound of at least 1000 primes."""
    import sympy as sp

    N = depth
    try:
        d, c = _poly_size((sp.sympify(a_str), sp.sympify(b_str)), sp.Symbol("n"))
        log2_fact = N * math.log2(N / math.e) + 0.5 * math.log2(2 * math.pi * N) if N > 1 else 1
        coeff_bits = math.log2(c + 1) if c > 1 else 0
        bits = int((d * log2_fact + N * coeff_bits) * safety) + 100
        return max(bits // 31 + 1, 64)
    except Exception:
        return max(int(4 * N * math.log2(max(N, 2) / math.e) * safety / 31) + 1, 1000)


def load_cmf_pcfs(path: str) -> List[Dict[str, Any]]:
    """A JSON array or JSONL file of PCF records (euler2ai_full_verify.py:170-183)."""
    with open(path) as f:
        first = f.read(1)
        f.seek(0)
        if first == "[":
            return json.load(f)
        return [json.loads(ln) for ln in (x.strip() for x in f) if ln]


def expand_tasks(records: Sequence[Dict[str, Any]], max_tasks: int = 0) -> List[Dict[str, Any]]:
    """One task per (record, source): a record's ``sources`` entries are
    [trajectory, shift, ...] lists, a record without sources is one task
    (euler2ai_full_verify.py:213-233)."""
    tasks = []
    for pi, rec in enumerate(records):
        sources = rec.get("sources", [None]) or [None]
        for si, src in enumerate(sources):
            tasks.append({
                "pcf_idx": pi, "source_idx": si, "a": rec["a"], "b": rec["b"], "limit": rec.get("limit", ""),
                "delta_ref": rec.get("delta", None), "conv_rate": rec.get("convergence_rate", None),
                "trajectory": src[0] if src and isinstance(src, list) else None,
                "shift": src[1] if src and isinstance(src, list) and len(src) > 1 else None,
            })
    return tasks[:max_tasks] if max_tasks > 0 else tasks


def _fmt(x: Optional[float]) -> str:
    return f"{x:.5f}" if x is not None else ""


def _measure(pcf, res, i, depth, dps, delta_tol):
    """est, delta, method, p_bits, diff, match of one walked PCF
    (euler2ai_full_verify.py:330-383)."""
    K = int(res.Ks[i])
    d = res.walk_dict(i)
    p_big, q_big = res.crt_values(i)
    M = math.prod(int(p) for p in d["primes"])
    est = d["p_float"] / d["q_float"] if abs(d["q_float"]) > 1e-300 else float("nan")
    overflow = crt_overflowed(p_big, q_big, est)
    target = pcf["target"]
    if target is not None:
        if overflow:
            delta, method = compute_delta_float(est, float(target), d.get("log_scale", 0.0), d["q_float"]), "float"
        else:
            delta, method = compute_delta(p_big, q_big, target, dps), "crt"
    else:
        delta, method = float("nan"), "none"
    ref = pcf["delta_ref"]
    if ref is not None and math.isfinite(delta):
        diff = abs(delta - ref)
        tol = delta_tol
        if ref < -0.5:
            tol = max(tol, 1.0)
        elif ref < 0:
            tol = max(tol, 0.2)
        match = diff < tol
    else:
        diff, match = float("nan"), False
    return (float(est), delta, method, M.bit_length(), K, diff, match)


def verify_pcf_records(records: Sequence[Dict[str, Any]], depth: int = 2000, K: int = 0, dps: int = 200,
                       delta_tol: float = 0.05, max_tasks: int = 0,
                       max_walks: int = 4096) -> Tuple[List[Dict[str, Any]], Dict[str, Any]]:
    """Report rows of every task (euler2ai_full_verify.py:186-420) and a
    summary {tasks, unique_pcfs, verified, matches, skipped, errors,
    elapsed_s}. K = 0 takes ``estimate_K`` per PCF. A PCF that does not
    compile (None: imaginary) gives SKIP rows; one whose compilation raises
    gives no row on its first task and SKIP rows after, as in the reference;
    one whose walk or measurement fails (K above the engine's 8192 primes
    included) gives an ERROR row, then SKIP rows; the other PCFs of its
    device batch are unaffected."""
    t0 = time.perf_counter()
    tasks = expand_tasks(records, max_tasks)
    pcfs: Dict[Tuple[str, str], Dict[str, Any]] = {}
    for t in tasks:
        key = (t["a"], t["b"])
        if key in pcfs:
            continue
        entry = {"state": "ok", "delta_ref": t["delta_ref"], "target": None, "err": ""}
        try:
            prog = compile_pcf(t["a"], t["b"])
            if prog is None:
                entry["state"] = "imaginary"
            else:
                entry["prog"], entry["a0"] = prog, pcf_initial_values(t["a"])
                entry["K"] = K if K > 0 else estimate_K(t["a"], t["b"], depth)
                entry["target"] = parse_limit(t["limit"], dps) if t["limit"] else None
                if entry["K"] > N.MAX_K:
                    entry["state"], entry["err"] = "walk_error", f"K={entry['K']} above {N.MAX_K} primes"
        except Exception as e:
            entry["state"], entry["err"] = "compile_error", str(e)
        pcfs[key] = entry
    todo = [k for k, e in pcfs.items() if e["state"] == "ok"]

    def walk_and_measure(keys):
        # one device batch; a PCF whose measurement raises fails alone
        specs = [pcf_spec(pcfs[k]["prog"], pcfs[k]["a0"], depth, pcfs[k]["K"], 1) for k in keys]
        with BatchPlan(specs, crt=True) as plan:
            res = plan.run()
            for i, k in enumerate(keys):
                try:
                    pcfs[k]["result"] = _measure(pcfs[k], res, i, depth, dps, delta_tol)
                except Exception as e:
                    pcfs[k]["state"], pcfs[k]["err"] = "walk_error", str(e)

    for b0 in range(0, len(todo), max_walks):
        chunk = todo[b0:b0 + max_walks]
        try:
            walk_and_measure(chunk)
        except Exception:
            # the batch was refused or failed as a whole: walk its PCFs one at
            # a time so that only the failing ones get the error row, as the
            # reference's per-PCF loop does (euler2ai_full_verify.py:380-394)
            for k in chunk:
                try:
                    walk_and_measure([k])
                except Exception as e:
                    pcfs[k]["state"], pcfs[k]["err"] = "walk_error", str(e)
    rows: List[Dict[str, Any]] = []
    seen = set()
    n_total = n_match = n_skip = n_error = 0
    for ti, t in enumerate(tasks):
        key = (t["a"], t["b"])
        e = pcfs[key]
        first = key not in seen
        seen.add(key)
        base = {"task_idx": ti, "pcf_idx": t["pcf_idx"], "source_idx": t["source_idx"], "a": t["a"], "b": t["b"],
                "limit": t["limit"], "trajectory": str(t["trajectory"]) if t["trajectory"] else "",
                "shift": str(t["shift"]) if t["shift"] else ""}
        if e["state"] != "ok":
            if first and e["state"] == "compile_error":
                n_error += 1
                continue
            if first and e["state"] == "walk_error":
                n_error += 1
                rows.append({**base, "delta_ref_cpu": t["delta_ref"], "convergence_rate_ref": t["conv_rate"],
                             "delta_cuda_rns": "ERROR", "delta_method": "", "est_float": "", "delta_diff": "",
                             "match": "ERROR", "p_bits": "", "depth": depth, "K": ""})
                continue
            n_skip += 1
            rows.append({**base, "delta_ref_cpu": t["delta_ref"], "convergence_rate_ref": t["conv_rate"],
                         "delta_cuda_rns": "SKIP", "delta_method": "", "est_float": "", "delta_diff": "",
                         "match": "SKIP", "p_bits": "", "depth": depth, "K": ""})
            continue
        est, delta, method, p_bits, K_use, diff, match = e["result"]
        n_total += 1
        n_match += bool(match)
        rows.append({**base, "delta_ref_cpu": _fmt(t["delta_ref"]), "convergence_rate_ref": _fmt(t["conv_rate"]),
                     "delta_cuda_rns": f"{delta:.5f}" if math.isfinite(delta) else str(delta),
                     "delta_method": method, "est_float": f"{est:.10f}" if math.isfinite(est) else str(est),
                     "delta_diff": f"{diff:.6f}" if math.isfinite(diff) else "", "match": "YES" if match else "NO",
                     "p_bits": p_bits, "depth": depth, "K": K_use})
    summary = {"tasks": len(tasks), "unique_pcfs": len(pcfs), "verified": n_total, "matches": n_match,
               "skipped": n_skip, "errors": n_error, "elapsed_s": time.perf_counter() - t0}
    return rows, summary


def write_csv(rows: Sequence[Dict[str, Any]], path: str) -> None:
    """The report as CSV with the reference's columns."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDNAMES)
        w.writeheader()
        for r in rows:
            w.writerow(r)

"""Euler2AI batch verification host logic (no GPU): the K estimate and the
task expansion of euler2ai_full_verify.py."""

import json
import math

from dreams_rns.euler2ai import FIELDNAMES, estimate_K, expand_tasks, load_cmf_pcfs


def test_estimate_k_closed_form():
    # README PCF at depth 2000: max degree 2, max coefficient 2:
    # bits = (2 log2(2000!) + 2000 log2 3) * 1.5 + 100, K = bits // 31 + 1
    lf = math.lgamma(2001) / math.log(2)  # Stirling's form agrees to 1e-7 here
    bits = int((2 * lf + 2000 * math.log2(3)) * 1.5) + 100
    assert estimate_K("2", "n**2", 2000) == bits // 31 + 1 == 2001
    assert estimate_K("3*n+1", "-2*n**2+5", 500) == 431
    assert estimate_K("n", "n**3-7*n", 50) == 64      # the floor
    assert estimate_K("1", "1", 1) == 64
    assert estimate_K("(", "n", 100) == 1000          # does not parse: the fallback
    # depth 20000 does not parse: 4 * log2 of N! by N log2(N/e), 1.5 safety
    assert estimate_K("(", "n", 20000) == int(4 * 20000 * math.log2(20000 / math.e) * 1.5 / 31) + 1
    # coefficients that are not integers count as 100, the degree still counts
    lf = 100 * math.log2(100 / math.e) + 0.5 * math.log2(2 * math.pi * 100)
    assert estimate_K("n/2 + x", "n**4", 100) == max((int((4 * lf + 100 * math.log2(101)) * 1.5) + 100) // 31 + 1, 64)
    # a constant numerator and a 7-term b: the largest coefficient (12) and degree (6)
    lf = 3000 * math.log2(3000 / math.e) + 0.5 * math.log2(2 * math.pi * 3000)
    assert estimate_K("5", "n**6 - 12*n + 3", 3000) == (int((6 * lf + 3000 * math.log2(13)) * 1.5) + 100) // 31 + 1


def test_task_expansion_and_loader(tmp_path):
    recs = [{"a": "2", "b": "n**2", "limit": "2/(4 - pi)", "delta": -1.0, "convergence_rate": 0.5,
             "sources": [[[1, 1], [0, 0]], [[1, 2]], None]},
            {"a": "n", "b": "1", "limit": "", "sources": []},
            {"a": "1", "b": "n"}]
    t = expand_tasks(recs)
    assert [(x["pcf_idx"], x["source_idx"]) for x in t] == [(0, 0), (0, 1), (0, 2), (1, 0), (2, 0)]
    assert t[0]["trajectory"] == [1, 1] and t[0]["shift"] == [0, 0]
    assert t[1]["trajectory"] == [1, 2] and t[1]["shift"] is None
    assert t[2]["trajectory"] is None and t[3]["limit"] == "" and t[4]["delta_ref"] is None
    assert len(expand_tasks(recs, max_tasks=2)) == 2
    p1 = tmp_path / "a.json"
    p1.write_text(json.dumps(recs))
    p2 = tmp_path / "b.jsonl"
    p2.write_text("\n".join(json.dumps(r) for r in recs) + "\n\n")
    assert load_cmf_pcfs(str(p1)) == recs == load_cmf_pcfs(str(p2))
    assert FIELDNAMES[0] == "task_idx" and FIELDNAMES[-1] == "K" and len(FIELDNAMES) == 18

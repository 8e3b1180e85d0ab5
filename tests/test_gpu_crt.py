"""CRT on the GPU for wide prime sets (the matrix-core path, K >= 257) and
at the switch from the Garner kernel: batches of rows through drx_crt
against the textbook CRT, bit-exact, including the values around +-(M-1)/2
where the rounded quotient sits next to a half-integer."""

import math

import numpy as np
import pytest

import dreams_rns as D
from oracle import walk as ow

pytestmark = pytest.mark.gpu


def _crt_rows(ctx, res, K):
    n = res.shape[0]
    limbs = np.zeros((n, K), dtype=np.uint32)
    neg = np.zeros(n, dtype=np.uint8)
    r = np.ascontiguousarray(res, dtype=np.int64)
    ctx._check(ctx.lib.drx_crt(ctx.h, r.ctypes.data, n, K, K, limbs.ctypes.data, neg.ctypes.data))
    out = []
    for i in range(n):
        mag = int.from_bytes(limbs[i].tobytes(), "little")
        out.append(-mag if neg[i] else mag)
    return out


@pytest.mark.parametrize("K", [255, 256, 257, 300, 512, 1000, 1024])
def test_crt_batch_equals_textbook(gpu_ctx, K):
    gpu_ctx.ensure_primes(K)
    pr = [int(p) for p in gpu_ctx.primes[:K]]
    M = math.prod(pr)
    H = (M - 1) // 2
    rng = np.random.default_rng(K)
    vals = [0, 1, -1, H, -H, H - 1, -(H - 1), M // 3, -(M // 7), 2 ** 31, -(2 ** 200)]
    # residues of random x in [0, M): the centred value is the answer
    rows = [[v % p for p in pr] for v in vals]
    for _ in range(150):
        rows.append([int(rng.integers(0, p)) for p in pr])
    res = np.array(rows, dtype=np.int64)
    print("crt K", K, "rows", len(rows), flush=True)
    got = _crt_rows(gpu_ctx, res, K)
    for i, row in enumerate(rows):
        x, MM = ow.crt_reconstruct(row, pr)
        assert MM == M
        assert got[i] == ow.centered(x, M), (K, i)
    assert got[: len(vals)] == vals


def test_plan_crt_mixed_widths(gpu_ctx):
    # one batch with walks of 64, 300 and 700 primes: the Garner rows and two
    # matrix-core groups write their own rows of the same output
    from dreams_rns.engine import BatchPlan, pcf_spec

    specs = []
    for i, (a, b) in enumerate([("3*n+1", "-n**2"), ("2", "n**2"), ("n+5", "2*n**3"), ("7", "-3*n**2+n")] * 3):
        specs.append(pcf_spec(D.compile_pcf(a, b), D.pcf_initial_values(a), 40 + 7 * i, [64, 300, 700][i % 3], 1))
    with BatchPlan(specs, crt=True) as plan:
        res = plan.run()
        for i, s in enumerate(specs):
            d = res.walk_dict(i)
            pr = [int(p) for p in d["primes"]]
            M = math.prod(pr)
            xp, _ = ow.crt_reconstruct([int(v) for v in d["p_residues"]], pr)
            xq, _ = ow.crt_reconstruct([int(v) for v in d["q_residues"]], pr)
            assert res.crt_values(i) == (ow.centered(xp, M), ow.centered(xq, M)), i

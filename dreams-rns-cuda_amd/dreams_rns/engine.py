"""Batched walk engine over libdreams_hip (host side of the C ABI).

A batch is a list of :class:`WalkSpec`. :class:`BatchPlan` packs it once
(programs, initial matrices, axis tables -> device), then ``run()`` executes
the whole batch on the GPU -- interpreter + chunked P*M products + ordered LDS
tree combine, FP64 shadow with delta_float, CRT (Garner up to 256 primes,
matrix cores above) and top-k -- and copies the results back (``run(fetch=False)``
leaves them in device memory until ``fetch()``). A plan can be run repeatedly; every run restarts from the
initial matrices.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

INT64_MIN, INT64_MAX = -(1 << 63), (1 << 63) - 1


@dataclass
class WalkSpec:
    """One walk: P0 * M(0) * ... * M(depth-1) mod the first K primes, or with
    ``vector`` the state-vector walk v(depth) = M(depth-1) ... M(0) v(0)
    (cmf_walk.py:445-531) with p = v[out_p], q = v[out_q].

    ``axes`` holds (sn, sd, tn, td) per axis: axis i at step s equals
    (sn*td + s*tn*sd) / (sd*td) (python/dreams_rns/cmf_walk.py:206-224).
    ``init`` is "pcf" (A = [[1, a0], [0, 1]], runner.py:376-382), "identity"
    (cmf_walk.py:396-402) or "given" (init_res [m*m][K] / init_f [m*m]; for a
    vector walk [m][K] / [m]).
    """

    program: Any
    depth: int
    K: int
    axes: Sequence[Tuple[int, int, int, int]] = ()
    init: str = "identity"
    a0: int = 0
    init_res: Optional[np.ndarray] = None
    init_f: Optional[np.ndarray] = None
    target: float = 0.0
    vector: bool = False
    out_p: int = 0
    out_q: int = 0


def pcf_spec(program, a0: int, depth: int, K: int = 32, shift_val: int = 1, target: float = 0.0) -> WalkSpec:
    """run_pcf_walk's walk: axis value shift_val + step * direction
    (runner.py:224-227)."""
    axes = [(int(shift_val), 1, int(d), 1) for d in program.directions[: program.dim]]
    if len(axes) < program.dim:
        raise IndexError("program has fewer directions than axes")
    return WalkSpec(program=program, depth=int(depth), K=int(K), axes=axes, init="pcf", a0=int(a0),
                    target=float(target))


def parse_rational_list(vals) -> Tuple[List[int], List[int]]:
    """int / Fraction / (num, den) -> (numerators, denominators)
    (cmf_walk.py:332-345)."""
    from fractions import Fraction

    nums, dens = [], []
    for v in vals:
        if isinstance(v, Fraction):
            nums.append(v.numerator)
            dens.append(v.denominator)
        elif isinstance(v, tuple):
            nums.append(int(v[0]))
            dens.append(int(v[1]))
        else:
            nums.append(int(v))
            dens.append(1)
    return nums, dens


def _axes(program, shift_vals, trajectory_vals):
    sn, sd = parse_rational_list(shift_vals)
    if trajectory_vals is not None:
        tn, td = parse_rational_list(trajectory_vals)
    else:
        tn, td = [int(d) for d in program.directions], [1] * len(program.directions)
    dim = program.dim
    if min(len(sn), len(tn)) < dim:
        raise IndexError(f"program has {dim} axes but {len(sn)} shifts / {len(tn)} trajectory values")
    return [(sn[i], sd[i], tn[i], td[i]) for i in range(dim)]


def cmf_spec(program, depth: int, K: int, shift_vals, trajectory_vals=None, target: float = 0.0) -> WalkSpec:
    """run_cmf_walk's walk from the identity (cmf_walk.py:363-431)."""
    return WalkSpec(program=program, depth=int(depth), K=int(K), axes=_axes(program, shift_vals, trajectory_vals),
                    init="identity", target=float(target))


def vec_spec(program, depth: int, K: int, shift_vals, initial_state, acc_idx: int, const_idx: int,
             trajectory_vals=None, primes=None, target: float = 0.0) -> WalkSpec:
    """run_cmf_walk_vec's walk (cmf_walk.py:445-531): v(0) from exact
    rationals (residues num * den^(p-2) mod p, cmf_walk.py:490-502; FP64
    float(f)); p = v[acc_idx], q = v[const_idx]. ``primes`` (the first K
    primes) fixes the residues of v(0)."""
    from fractions import Fraction
    from .primes import generate_rns_primes

    m = program.m
    if len(initial_state) != m:
        raise ValueError(f"initial state has {len(initial_state)} entries, the program's order is {m}")
    if not (0 <= acc_idx < m and 0 <= const_idx < m):
        raise IndexError("accumulator / constant index outside the state vector")
    pr = [int(p) for p in (primes if primes is not None else generate_rns_primes(K))[:K]]
    res = np.zeros((m, K), dtype=np.uint32)
    f = np.zeros(m, dtype=np.float64)
    for i, x in enumerate(initial_state):
        x = Fraction(x)
        f[i] = float(x)
        if x != 0:
            for k, p in enumerate(pr):
                num = x.numerator % p
                res[i, k] = num if x.denominator == 1 else num * pow(x.denominator, p - 2, p) % p
    return WalkSpec(program=program, depth=int(depth), K=int(K), axes=_axes(program, shift_vals, trajectory_vals),
                    init="given", init_res=res, init_f=f, target=float(target), vector=True,
                    out_p=int(acc_idx), out_q=int(const_idx))


@dataclass
class BatchResult:
    residues: np.ndarray                 # (n, 2, stride) int64: p then q
    out: np.ndarray                      # (n,) OUT_DTYPE
    primes: np.ndarray
    Ks: np.ndarray
    limbs: Optional[np.ndarray] = None   # (n, 2, stride) uint32
    neg: Optional[np.ndarray] = None     # (n, 2) uint8
    full: Optional[np.ndarray] = None    # (n, 16, stride) int64
    full_f: Optional[np.ndarray] = None  # (n, 16) float64: final shadow matrix
    topk: Optional[np.ndarray] = None    # (top_k,) int32

    def walk_dict(self, i: int) -> Dict[str, Any]:
        """The reference's result dict (runner.py:406-413)."""
        K = int(self.Ks[i])
        o = self.out[i]
        return {
            "p_residues": self.residues[i, 0, :K].copy(),
            "q_residues": self.residues[i, 1, :K].copy(),
            "p_float": np.float64(o["p_float"]),
            "q_float": np.float64(o["q_float"]),
            "log_scale": float(o["log_scale"]),
            "primes": self.primes[:K].copy(),
        }

    def crt_values(self, i: int) -> Tuple[int, int]:
        """Centred CRT integers (p, q) of walk i from the GPU CRT limbs."""
        if self.limbs is None:
            raise RuntimeError("plan was built without crt=True")
        K = int(self.Ks[i])
        vals = []
        for v in range(2):
            mag = int.from_bytes(self.limbs[i, v, :K].tobytes(), "little")
            vals.append(-mag if self.neg[i, v] else mag)
        return vals[0], vals[1]

    def full_matrix(self, i: int, m: int) -> np.ndarray:
        if self.full is None:
            raise RuntimeError("plan was built without full=True")
        K = int(self.Ks[i])
        return self.full[i, : m * m, :K].reshape(m, m, K).copy()

    def full_shadow(self, i: int, m: int) -> np.ndarray:
        if self.full_f is None:
            raise RuntimeError("plan was built without full=True")
        return self.full_f[i, : m * m].reshape(m, m).copy()


def _pack_walks(specs: Sequence[WalkSpec], ctx: N.Context):
    """drx_walk records and the axis / initial-state pools of a batch."""
    kmax = max(int(s.K) for s in specs)
    ctx.ensure_primes(kmax)
    walks = np.zeros(len(specs), dtype=N.WALK_DTYPE)
    axis: List[int] = []
    init_res: List[np.ndarray] = []
    init_f: List[float] = []
    res_off = 0
    for i, s in enumerate(specs):
        w = walks[i]
        w["prog"] = ctx.program_id(s.program, N.DRX_WALK_VECTOR if s.vector else N.DRX_WALK_MATRIX, int(s.K))
        w["K"] = s.K
        w["depth"] = s.depth
        w["target"] = s.target
        w["axis_off"] = len(axis)
        w["mode"] = N.DRX_WALK_VECTOR if s.vector else N.DRX_WALK_MATRIX
        w["out_p"] = s.out_p
        w["out_q"] = s.out_q
        for q in s.axes:
            axis.extend(int(v) for v in q)
        m = s.program.m
        if s.init == "pcf" and INT64_MIN <= int(s.a0) <= INT64_MAX:
            w["init_kind"] = N.DRX_INIT_PCF
            w["a0"] = int(s.a0)
        elif s.init == "identity":
            w["init_kind"] = N.DRX_INIT_IDENTITY
        else:
            if s.init == "pcf":  # a0 outside int64: give the residues
                pr = [int(p) for p in ctx.primes[: s.K]]
                r = np.zeros((4, s.K), dtype=np.uint32)
                r[0] = 1
                r[1] = [int(s.a0) % p for p in pr]
                r[3] = 1
                f = np.array([1.0, float(int(s.a0)), 0.0, 1.0])
            elif s.init == "given":
                rows = m if s.vector else m * m
                r = np.ascontiguousarray(s.init_res, dtype=np.uint32).reshape(rows, s.K)
                f = np.ascontiguousarray(s.init_f, dtype=np.float64).reshape(rows)
            else:
                raise ValueError(f"unknown init kind {s.init!r}")
            w["init_kind"] = N.DRX_INIT_GIVEN
            w["init_off"] = res_off
            w["init_f_off"] = len(init_f)
            init_res.append(r.reshape(-1))
            init_f.extend(float(v) for v in f)
            res_off += r.size
    for v in axis:
        if not INT64_MIN <= v <= INT64_MAX:
            raise OverflowError("shift/trajectory value does not fit in int64")
    axis_a = np.array(axis if axis else [0], dtype=np.int64)
    init_res_a = np.concatenate(init_res) if init_res else np.zeros(1, np.uint32)
    init_f_a = np.array(init_f if init_f else [0.0], dtype=np.float64)
    return walks, axis_a, init_res_a, init_f_a


def batch_signature(specs: Sequence[WalkSpec]) -> tuple:
    """The shape mix of a batch, for the autotuner's cache: counts of
    (matrix order, K, forward-difference or not, depth octave), relative to
    the batch size in 1/16ths."""
    from collections import Counter

    c = Counter()
    for s in specs:
        meta = N.pack_program(s.program)[2]
        c[(s.program.m, int(s.K), int(meta["fd_ok"]), max(0, int(s.depth)).bit_length())] += 1
    n = len(specs)
    return tuple(sorted((k, round(16 * v / n)) for k, v in c.items())) + (("n", n.bit_length()),)


_TUNED: Dict[tuple, Dict[str, float]] = {}


def autotune(specs: Sequence[WalkSpec], ctx: Optional[N.Context] = None, max_walks: int = 16384,
             budget_s: float = 20.0, interp_only: bool = False) -> Dict[str, float]:
    """Launch shape for this batch -- workgroup size, chunk length, walks per
    workgroup and the prime stripe of wide walks (drx_autotune: whole
    pipeline runs on a sample of at most ``max_walks`` walks, shrunk to fit
    ``budget_s``) -- cached per batch signature."""
    ctx = ctx or N.context()
    key = (batch_signature(specs), bool(interp_only))
    if key in _TUNED:
        return _TUNED[key]
    walks, axis, init_res, init_f = _pack_walks(specs, ctx)
    b = N.Batch()
    b.walks = walks.ctypes.data
    b.n = len(specs)
    b.axis_pool = axis.ctypes.data
    b.init_res_pool = init_res.ctypes.data
    b.init_f_pool = init_f.ctypes.data
    base = N.LaunchCfg(N.MODE_BATCHED, 0, 0, 0, 1 if interp_only else 0)
    best = N.LaunchCfg()
    ms = C.c_double(0.0)
    ctx._check(ctx.lib.drx_autotune(ctx.h, C.byref(b), C.byref(base), int(max_walks), float(budget_s),
                                    C.byref(best), C.byref(ms)))
    out = {"threads": int(best.threads), "chunk_steps": int(best.chunk_steps),
           "walks_per_wg": max(1, int(best.walks_per_wg)), "slice": int(best.slice) or 256,
           "sample_walk_ms": ms.value}
    _TUNED[key] = out
    return out


class BatchPlan:
    """Packed, device-resident batch; see the module docstring."""

    MODES = {"batched": N.MODE_BATCHED, "plain": N.MODE_PLAIN, "persistent": N.MODE_PERSISTENT}

    def __init__(self, specs: Sequence[WalkSpec], ctx: Optional[N.Context] = None, mode: str = "batched",
                 threads: int = 0, chunk_steps: int = 0, use_graph: bool = False, crt: bool = False,
                 full: bool = False, top_k: int = 0, interp_only: bool = False, tune: bool = False,
                 chain_cus: int = 0, walks_per_wg: int = 0, slice: int = 0):
        self.ctx = ctx or N.context()
        n = len(specs)
        if n == 0:
            raise ValueError("empty batch")
        kmax = max(int(s.K) for s in specs)
        ctx = self.ctx
        walks, self._axis, self._init_res, self._init_f = _pack_walks(specs, ctx)
        self._walks = walks
        stride = kmax
        if tune and mode != "plain" and not (threads or chunk_steps):
            t = autotune(specs, ctx=ctx, interp_only=interp_only)
            threads, chunk_steps = t["threads"], t["chunk_steps"]
            walks_per_wg, slice = t["walks_per_wg"], t["slice"]
        self.threads, self.chunk_steps = int(threads), int(chunk_steps)
        self.use_graph = bool(use_graph)
        self.walks_per_wg, self.slice = int(walks_per_wg), int(slice)
        # outputs live in page-locked memory: the copies back run at link speed
        z = N.pinned_zeros
        self.result = BatchResult(
            residues=z((n, 2, stride), np.int64),
            out=z((n,), N.OUT_DTYPE),
            primes=ctx.primes[:kmax].copy(),
            Ks=walks["K"].copy(),
            limbs=z((n, 2, stride), np.uint32) if crt else None,
            neg=z((n, 2), np.uint8) if crt else None,
            full=z((n, 16, stride), np.int64) if full else None,
            full_f=z((n, 16), np.float64) if full else None,
            topk=z((top_k,), np.int32) if top_k else None,
        )
        r = self.result
        b = N.Batch()
        b.walks = walks.ctypes.data
        b.n = n
        b.axis_pool = self._axis.ctypes.data
        b.init_res_pool = self._init_res.ctypes.data
        b.init_f_pool = self._init_f.ctypes.data
        b.out = r.out.ctypes.data
        b.residues = r.residues.ctypes.data
        b.res_stride = stride
        b.full = r.full.ctypes.data if full else None
        b.full_f = r.full_f.ctypes.data if full else None
        b.crt_limbs = r.limbs.ctypes.data if crt else None
        b.crt_neg = r.neg.ctypes.data if crt else None
        b.top_k = int(top_k)
        b.topk_idx = r.topk.ctypes.data if top_k else None
        cfg = N.LaunchCfg(self.MODES[mode], int(threads), int(chunk_steps), 1 if use_graph else 0,
                          1 if interp_only else 0, int(chain_cus), int(walks_per_wg), int(slice))
        self._b, self._cfg = b, cfg
        h = C.c_void_p()
        ctx._check(ctx.lib.drx_plan_create(ctx.h, C.byref(b), C.byref(cfg), C.byref(h)))
        self.h = h
        self.n = n
        self.mode = mode

    def run(self, fetch: bool = True) -> BatchResult:
        """Run the batch. With ``fetch=False`` the outputs stay in device
        memory (``self.result`` keeps the previous contents until ``fetch``)."""
        if fetch:
            self.ctx._check(self.ctx.lib.drx_plan_run(self.h))
        else:
            self.ctx._check(self.ctx.lib.drx_plan_run_resident(self.h))
        return self.result

    def fetch(self) -> BatchResult:
        """Copy the last run's outputs into ``self.result``."""
        self.ctx._check(self.ctx.lib.drx_plan_fetch(self.h))
        return self.result

    def stats(self) -> Dict[str, float]:
        s = np.zeros(4, dtype=np.float64)
        self.ctx._check(self.ctx.lib.drx_plan_stats(self.h, s.ctypes.data))
        return {"residue_steps": s[0], "modmul": s[1], "modadd": s[2], "tasks": s[3]}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx.lib.drx_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_specs(specs: Sequence[WalkSpec], **kw) -> BatchResult:
    """Plan, run once, release."""
    with BatchPlan(specs, **kw) as plan:
        return plan.run()

"""pFq CMF sweeps on the GPU: specs x trajectories x shifts
(python/a100_sweep.py:284-303 loaders, 336-528 run_cmf_sweep_mode).

The reference compiles one program per (spec, trajectory) and walks the
512 shift offsets of that pair as one batch, one batch at a time
(a100_sweep.py:400-418). Here the walks of many (spec, trajectory) pairs
share one device batch: the batch holds up to ``max_walks`` walks of any
number of programs, so the device stays full however few shifts there are.
The integer offsets ignore the shifts' denominators, so distinct shifts
often start at the same offsets (the 512 shifts of dim5_shifts.json give 129
offset vectors): each distinct start is walked once and its result is
reported for every shift that maps to it.

Every walk starts from the identity; axis i starts at the integer offset
max(1, 1 + num_i) of its shift (cmf_walk.py:157-169; offsets padded with 1
to the spec's dim, a100_sweep.py:391-398) and moves by trajectory[i] per
step. Results are the reference's per-walk dicts (gpu_walk.py:478-487), each
with its own FP64 shadow.
"""

from __future__ import annotations

import csv
import json
import os
import time
from dataclasses import dataclass, field
from typing import Any, Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np

from .cmf_walk import compile_cmf_spec, shift_to_axis_offsets
from .engine import BatchPlan, BatchResult, cmf_spec

__all__ = ["load_cmf_specs", "load_trajectories", "load_shifts", "sweep_offsets", "SweepBatch", "CmfSweep",
           "run_cmf_sweep", "save_hits"]


def load_cmf_specs(path: str) -> List[Dict]:
    """One pFq spec per non-empty JSONL line (a100_sweep.py:284-291)."""
    with open(path) as f:
        return [json.loads(line) for line in (ln.strip() for ln in f) if line]


def load_trajectories(path: str) -> List[List[int]]:
    """The "trajectories" list of a trajectory file (a100_sweep.py:294-297)."""
    with open(path) as f:
        return json.load(f)["trajectories"]


def load_shifts(path: str) -> List[Dict]:
    """The "shifts" list of a shift file (a100_sweep.py:300-303)."""
    with open(path) as f:
        return json.load(f)["shifts"]


def sweep_offsets(shifts: Sequence[Dict], p: int, dim: int) -> List[List[int]]:
    """Integer axis starts of every shift for a spec of ``dim`` axes: padded
    with 1, cut to dim (a100_sweep.py:391-398)."""
    out = []
    for sh in shifts:
        so = shift_to_axis_offsets(sh, p)
        so = so + [1] * (dim - len(so))
        out.append(so[:dim])
    return out


@dataclass
class SweepBatch:
    """One device batch of a sweep. Row i is the walk of spec ``spec_idx[i]``
    along trajectory ``traj_idx[i]`` from the axis offsets ``offsets[i]``;
    ``shift_rows[i]`` lists the shift indices with those offsets (shifts that
    differ only in their denominators start at the same offsets, so their
    walks are one walk)."""

    spec_idx: np.ndarray
    traj_idx: np.ndarray
    offsets: List[List[int]]
    shift_rows: List[List[int]]
    result: BatchResult
    crt: bool = False

    def __len__(self) -> int:
        return len(self.spec_idx)

    @property
    def n_results(self) -> int:
        return sum(len(r) for r in self.shift_rows)

    def walk_dict(self, i: int, shift_idx: Optional[int] = None) -> Dict[str, Any]:
        """run_cmf_walk_batch_gpu's dict for row i (gpu_walk.py:478-487) plus
        the sweep indices; with crt, the centred CRT integers p_crt / q_crt."""
        d = self.result.walk_dict(i)
        d["shift_vals"] = list(self.offsets[i])
        d["spec_idx"] = int(self.spec_idx[i])
        d["traj_idx"] = int(self.traj_idx[i])
        d["shift_idx"] = int(self.shift_rows[i][0] if shift_idx is None else shift_idx)
        if self.crt:
            d["p_crt"], d["q_crt"] = self.result.crt_values(i)
        return d

    def records(self) -> Iterator[Dict[str, Any]]:
        """One dict per (spec, trajectory, shift), in sweep order per row."""
        for i in range(len(self)):
            base = self.walk_dict(i)
            for h in self.shift_rows[i]:
                d = dict(base)
                d["shift_idx"] = int(h)
                yield d


@dataclass
class CmfSweep:
    """Iterate the sweep batch by batch.

    ``specs`` are pFq spec dicts, ``trajectories`` integer direction vectors
    of length dim, ``shifts`` shift records ({"nums", "dens"}). Pairs whose
    program does not compile (None, or an exception, a100_sweep.py:403-410)
    are listed in ``skipped`` as (spec index, trajectory index, reason).
    With ``dedupe`` (default) shifts of equal axis offsets are walked once.
    """

    specs: Sequence[Dict]
    trajectories: Sequence[Sequence[int]]
    shifts: Sequence[Dict]
    depth: int = 2000
    K: int = 1000
    max_walks: int = 8192
    crt: bool = False
    dedupe: bool = True
    plan_kw: Dict[str, Any] = field(default_factory=dict)
    skipped: List[Tuple[int, int, str]] = field(default_factory=list)

    @property
    def total_walks(self) -> int:
        """Walks of the reference's sweep: specs x trajectories x shifts."""
        return len(self.specs) * len(self.trajectories) * len(self.shifts)

    def _rows(self) -> Iterator[Tuple[int, int, Any, List[int], List[int]]]:
        for si, spec in enumerate(self.specs):
            dim = int(spec.get("dim", int(spec["p"]) + int(spec["q"])))
            offs = sweep_offsets(self.shifts, int(spec["p"]), dim)
            groups: Dict[Tuple[int, ...], List[int]] = {}
            if self.dedupe:
                for hi, so in enumerate(offs):
                    groups.setdefault(tuple(so), []).append(hi)
                uniq = [(list(k), v) for k, v in groups.items()]
            else:
                uniq = [(so, [hi]) for hi, so in enumerate(offs)]
            for ti, traj in enumerate(self.trajectories):
                try:
                    prog = compile_cmf_spec(spec, trajectory=list(traj))
                except Exception as e:  # a100_sweep.py skips the pair and goes on
                    self.skipped.append((si, ti, f"{type(e).__name__}: {e}"))
                    continue
                if prog is None:
                    self.skipped.append((si, ti, "not real"))
                    continue
                for so, hs in uniq:
                    yield si, ti, prog, so, hs

    def _run(self, rows) -> SweepBatch:
        specs = [cmf_spec(prog, self.depth, self.K, so) for (_, _, prog, so, _) in rows]
        with BatchPlan(specs, crt=self.crt, **self.plan_kw) as plan:
            res = plan.run()
        return SweepBatch(spec_idx=np.array([r[0] for r in rows], dtype=np.int32),
                          traj_idx=np.array([r[1] for r in rows], dtype=np.int32),
                          offsets=[r[3] for r in rows], shift_rows=[r[4] for r in rows], result=res, crt=self.crt)

    def __iter__(self) -> Iterator[SweepBatch]:
        self.skipped = []
        rows: List[Tuple[int, int, Any, List[int], List[int]]] = []
        for r in self._rows():
            rows.append(r)
            if len(rows) >= self.max_walks:
                yield self._run(rows)
                rows = []
        if rows:
            yield self._run(rows)


def run_cmf_sweep(specs: Sequence[Dict], trajectories: Sequence[Sequence[int]], shifts: Sequence[Dict],
                  constants: Sequence[Dict[str, Any]], depth: int = 2000, K: int = 1000, proximity: float = 1e-3,
                  dps: int = 200, max_walks: int = 8192, **plan_kw) -> Tuple[List[Dict[str, Any]], Dict[str, Any]]:
    """The CMF sweep with its hit selection (a100_sweep.py:336-528): every
    spec x trajectory x shift walked on the GPU (GPU CRT), then
    ``postwalk.sweep_hits``. Returns (hits, stats) with stats total_runs
    (specs x trajectories x shifts, skipped pairs included, as the reference
    counts them), total_hits, skipped, elapsed_s, device_s (batches) and
    post_s (hit selection)."""
    from .postwalk import sweep_hits

    t0 = time.perf_counter()
    sw = CmfSweep(specs, trajectories, shifts, depth=depth, K=K, max_walks=max_walks, crt=True, plan_kw=plan_kw)
    hits: List[Dict[str, Any]] = []
    dev = post = 0.0
    it = iter(sw)
    while True:
        t = time.perf_counter()
        try:
            batch = next(it)
        except StopIteration:
            break
        dev += time.perf_counter() - t
        t = time.perf_counter()
        hits += sweep_hits(batch, specs, trajectories, constants, proximity, dps)
        post += time.perf_counter() - t
    return hits, {"total_runs": sw.total_walks, "total_hits": len(hits), "skipped": list(sw.skipped),
                  "elapsed_s": time.perf_counter() - t0, "device_s": dev, "post_s": post}


def save_hits(hits: Sequence[Dict[str, Any]], output_dir: str, prefix: str = "hits") -> str:
    """Hits to <prefix>.jsonl (every field) and, when there are any,
    <prefix>.csv (columns of the first hit, lists written as their str)
    (a100_sweep.py:533-556). Returns the JSONL path."""
    os.makedirs(output_dir, exist_ok=True)
    path = os.path.join(output_dir, f"{prefix}.jsonl")
    with open(path, "w") as f:
        for h in hits:
            f.write(json.dumps(h, default=str) + "\n")
    if hits:
        keys = list(hits[0].keys())
        with open(os.path.join(output_dir, f"{prefix}.csv"), "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=keys)
            w.writeheader()
            for h in hits:
                w.writerow({k: str(v) if isinstance(v, list) else v for k, v in h.items()})
    return path

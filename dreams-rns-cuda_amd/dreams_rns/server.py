"""Continuous-batching walk server (Python side of drx_server_*).

Requests (:class:`~dreams_rns.engine.WalkSpec`) are queued; a native thread
admits them into device-resident slots as slots and memory free up and
advances every resident walk by ``round_steps`` steps per round (several such
step ranges side by side while few walks are resident), so requests that
arrive while others run join at the next round and share its launches. Each
finished request returns the reference's result dict, the centred CRT
integers of p and q (computed on the GPU in the request's last round, as
the reference's drivers follow each walk with crt_reconstruct + centered,
python/euler2ai_full_verify.py:331-342) and its timing.
"""

from __future__ import annotations

import ctypes as C
import time
from typing import Any, Dict, List, Optional

import numpy as np

from . import _native as N
from .engine import WalkSpec


class WalkServer:
    def __init__(self, ctx: Optional[N.Context] = None, max_slots: int = 0, round_steps: int = 1000,
                 threads: int = 0, mem_budget: int = 0, max_K: int = 128, max_walk_tasks: int = 0):
        self.ctx = ctx or N.context()
        self.ctx.ensure_primes(max_K)
        cfg = N.ServerCfg(int(max_slots), int(round_steps), int(threads), int(max_walk_tasks), int(mem_budget))
        h = C.c_void_p()
        self.ctx._check(self.ctx.lib.drx_server_create(self.ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self.stride = min(len(self.ctx.primes), 256)
        self.submitted = 0
        self.completed = 0
        self._pending: Dict[int, WalkSpec] = {}
        self._pids: Dict[int, tuple] = {}  # id(program) -> (program, program id, K covered)
        self._cap = 0

    def _program_id(self, program, K: int) -> int:
        hit = self._pids.get(id(program))
        if hit is not None and hit[0] is program and hit[2] >= K:
            return hit[1]
        pid = self.ctx.program_id(program, N.DRX_WALK_MATRIX, int(K))
        self._pids[id(program)] = (program, pid, int(K))
        return pid

    def submit(self, spec: WalkSpec) -> int:
        if spec.init not in ("pcf", "identity"):
            raise ValueError("the server takes PCF or identity initial matrices")
        w = N.Walk()
        w.prog = self._program_id(spec.program, spec.K)
        w.K = spec.K
        w.depth = spec.depth
        w.init_kind = N.DRX_INIT_PCF if spec.init == "pcf" else N.DRX_INIT_IDENTITY
        w.a0 = int(spec.a0)
        w.target = spec.target
        axis = np.array([int(v) for q in spec.axes for v in q] or [0], dtype=np.int64)
        t = C.c_uint64()
        self.ctx._check(self.ctx.lib.drx_server_submit(self.h, C.byref(w), axis.ctypes.data, C.byref(t)))
        self.submitted += 1
        self._pending[t.value] = spec
        return t.value

    def _buffers(self, n: int):
        if n > self._cap:
            self._cap = max(n, 2 * self._cap, 64)
            self._res = (N.ServerResult * self._cap)()
            self._resid = np.zeros((self._cap, 2, self.stride), dtype=np.int64)
            self._limbs = np.zeros((self._cap, 2, self.stride), dtype=np.uint32)
        return self._res, self._resid, self._limbs

    def poll(self, max_results: int = 4096) -> List[Dict[str, Any]]:
        """Finished requests: the reference's walk dict plus ``p_big`` /
        ``q_big`` (centred CRT integers), ``delta_float``, ``status`` and
        timestamps (``latency_s`` = submission to CRT integers in host
        memory)."""
        want = max(1, min(int(max_results), self.submitted - self.completed))
        res, resid, limbs = self._buffers(want)
        n = C.c_int(0)
        self.ctx._check(self.ctx.lib.drx_server_poll(self.h, res, resid.ctypes.data, limbs.ctypes.data, self.stride,
                                                     want, C.byref(n)))
        out = []
        for i in range(n.value):
            r = res[i]
            spec = self._pending.pop(r.ticket, None)
            K = r.K
            d = {
                "ticket": r.ticket, "status": r.status,
                "p_residues": resid[i, 0, :K].copy(), "q_residues": resid[i, 1, :K].copy(),
                "p_float": np.float64(r.out.p_float), "q_float": np.float64(r.out.q_float),
                "log_scale": float(r.out.log_scale), "delta_float": float(r.out.delta_float),
                "primes": self.ctx.primes[:K].copy(),
                "t_submit": r.t_submit, "t_admit": r.t_admit, "t_done": r.t_done,
                "latency_s": r.t_done - r.t_submit, "spec": spec,
            }
            if r.status == 0:
                for v, key in ((0, "p_big"), (1, "q_big")):
                    mag = int.from_bytes(limbs[i, v, :K].tobytes(), "little")
                    d[key] = -mag if r.neg[v] else mag
            out.append(d)
        self.completed += n.value
        return out

    def pending(self):
        q, r = C.c_int(), C.c_int()
        self.ctx._check(self.ctx.lib.drx_server_pending(self.h, C.byref(q), C.byref(r)))
        return q.value, r.value

    def capacity(self) -> int:
        n = C.c_int()
        self.ctx._check(self.ctx.lib.drx_server_capacity(self.h, C.byref(n)))
        return n.value

    def drain(self, timeout: float = 600.0, sleep: float = 0.0005) -> List[Dict[str, Any]]:
        """Wait until every submitted request has finished; returns them."""
        out = []
        t0 = time.perf_counter()
        while self.completed < self.submitted:
            got = self.poll()
            out.extend(got)
            if not got:
                if time.perf_counter() - t0 > timeout:
                    raise TimeoutError(f"{self.submitted - self.completed} requests still running")
                time.sleep(sleep)
        return out

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.drx_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

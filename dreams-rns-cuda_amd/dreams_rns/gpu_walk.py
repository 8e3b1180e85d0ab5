"""Batched GPU walk API with the reference's names (python/dreams_rns/gpu_walk.py).

Unlike the reference's CuPy path, every walk of a batch gets its own FP64
shadow (the reference copies the first walk's shadow into every result,
gpu_walk.py:345-377), INV never leaves the device, and there is no silent
CPU fallback: without the HIP library these functions raise.
"""

from __future__ import annotations

from typing import Any, Dict, List, Optional

from . import _native as N
from .compiler import CmfProgram
from .engine import BatchPlan, cmf_spec, pcf_spec

__all__ = ["gpu_available", "run_pcf_walk_gpu", "run_pcf_walk_batch_gpu", "run_cmf_walk_gpu",
           "run_cmf_walk_batch_gpu"]


def gpu_available() -> bool:
    """True when libdreams_hip loads and a HIP device is visible (gpu_walk.py:49-51)."""
    try:
        return N.device_count() > 0
    except N.NativeUnavailable:
        return False


def run_pcf_walk_gpu(program: CmfProgram, a0: int, depth: int, K: int = 32, shift_val: int = 1) -> Dict[str, Any]:
    """Single PCF walk; same interface as run_pcf_walk (gpu_walk.py:270-274)."""
    return run_pcf_walk_batch_gpu(program, a0, depth, K, [shift_val])[0]


def run_pcf_walk_batch_gpu(program: CmfProgram, a0: int, depth: int, K: int = 32,
                           shift_vals: Optional[List[int]] = None, batch_size: int = 4096) -> List[Dict[str, Any]]:
    """One walk per shift value, all in one device batch per ``batch_size``
    walks (gpu_walk.py:279-379). Each dict also carries ``shift_val``."""
    assert program.m == 2, "GPU walk only supports 2x2 companion matrices"
    shift_vals = [1] if shift_vals is None else list(shift_vals)
    out: List[Dict[str, Any]] = []
    for b0 in range(0, len(shift_vals), max(1, batch_size)):
        chunk = shift_vals[b0:b0 + batch_size]
        specs = [pcf_spec(program, a0, depth, K, sv) for sv in chunk]
        with BatchPlan(specs) as plan:
            res = plan.run()
        for i, sv in enumerate(chunk):
            d = res.walk_dict(i)
            d["shift_val"] = sv
            out.append(d)
    return out


def run_cmf_walk_gpu(program: CmfProgram, depth: int, K: int, shift_vals) -> Dict[str, Any]:
    """Single CMF walk (gpu_walk.py:384-392)."""
    return run_cmf_walk_batch_gpu(program, depth, K, [shift_vals])[0]


def run_cmf_walk_batch_gpu(program: CmfProgram, depth: int, K: int = 32, shift_val_list=None,
                           batch_size: int = 4096) -> List[Dict[str, Any]]:
    """One CMF walk per shift vector (gpu_walk.py:395-489). Each dict also
    carries ``shift_vals``."""
    shift_val_list = [[1] * program.dim] if shift_val_list is None else list(shift_val_list)
    out: List[Dict[str, Any]] = []
    for b0 in range(0, len(shift_val_list), max(1, batch_size)):
        chunk = shift_val_list[b0:b0 + batch_size]
        specs = [cmf_spec(program, depth, K, sv) for sv in chunk]
        with BatchPlan(specs) as plan:
            res = plan.run()
        for i, sv in enumerate(chunk):
            d = res.walk_dict(i)
            d["shift_vals"] = list(sv)
            out.append(d)
    return out

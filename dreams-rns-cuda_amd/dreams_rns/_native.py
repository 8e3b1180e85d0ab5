"""ctypes binding of libdreams_hip.so (include/dreams_hip.h).

The library is required: nothing here falls back to a CPU implementation. A
missing library or a machine without a HIP device raises ``NativeUnavailable``.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
import weakref
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._build import LIB_PATH

ABI_VERSION = 5
PACK_META = 13  # values drx_program_pack reports
MAX_K = 8192  # primes one context holds (walks of more than 256 primes run as prime slices)
DRX_INIT_PCF, DRX_INIT_IDENTITY, DRX_INIT_GIVEN = 0, 1, 2
DRX_WALK_MATRIX, DRX_WALK_VECTOR = 0, 1
MODE_BATCHED, MODE_PLAIN, MODE_PERSISTENT = 0, 1, 2

_ERRORS = {-1: ValueError, -2: RuntimeError, -3: MemoryError, -4: OverflowError, -5: RuntimeError, -6: BlockingIOError}


class NativeUnavailable(RuntimeError):
    """libdreams_hip.so is not built or no HIP device is present."""


class Walk(C.Structure):
    _fields_ = [
        ("prog", C.c_int32), ("K", C.c_int32), ("depth", C.c_int64),
        ("init_kind", C.c_int32), ("mode", C.c_int32), ("a0", C.c_int64),
        ("init_off", C.c_int64), ("init_f_off", C.c_int64), ("axis_off", C.c_int64),
        ("target", C.c_double), ("out_p", C.c_int32), ("out_q", C.c_int32),
    ]


WALK_DTYPE = np.dtype([
    ("prog", "<i4"), ("K", "<i4"), ("depth", "<i8"), ("init_kind", "<i4"), ("mode", "<i4"),
    ("a0", "<i8"), ("init_off", "<i8"), ("init_f_off", "<i8"), ("axis_off", "<i8"), ("target", "<f8"),
    ("out_p", "<i4"), ("out_q", "<i4"),
], align=True)
assert WALK_DTYPE.itemsize == C.sizeof(Walk)

OUT_DTYPE = np.dtype([("p_float", "<f8"), ("q_float", "<f8"), ("log_scale", "<f8"),
                      ("delta_float", "<f8"), ("log_q", "<f8"), ("reserved", "<f8")])


class WalkOut(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("p_float", "q_float", "log_scale", "delta_float", "log_q", "reserved")]


class Batch(C.Structure):
    _fields_ = [
        ("walks", C.c_void_p), ("n", C.c_int), ("axis_pool", C.c_void_p),
        ("init_res_pool", C.c_void_p), ("init_f_pool", C.c_void_p), ("out", C.c_void_p),
        ("residues", C.c_void_p), ("res_stride", C.c_int), ("full", C.c_void_p), ("full_f", C.c_void_p),
        ("crt_limbs", C.c_void_p), ("crt_neg", C.c_void_p), ("top_k", C.c_int),
        ("topk_idx", C.c_void_p),
    ]


class LaunchCfg(C.Structure):
    _fields_ = [("mode", C.c_int), ("threads", C.c_int), ("chunk_steps", C.c_int), ("use_graph", C.c_int),
                ("interp_only", C.c_int), ("chain_cus", C.c_int), ("walks_per_wg", C.c_int), ("slice", C.c_int)]


class ServerCfg(C.Structure):
    _fields_ = [("max_slots", C.c_int), ("round_steps", C.c_int), ("threads", C.c_int),
                ("max_walk_tasks", C.c_int), ("mem_budget", C.c_int64)]


class ServerResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("out", WalkOut), ("K", C.c_int32), ("status", C.c_int32),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("neg", C.c_uint8 * 2), ("pad", C.c_uint8 * 6)]


# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "drx_abi_version": (C.c_int, []),
    "drx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "drx_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "drx_ctx_destroy": (None, [C.c_void_p]),
    "drx_last_error": (C.c_char_p, [C.c_void_p]),
    "drx_set_primes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "drx_program_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "drx_program_upload_mode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_int)]),
    "drx_program_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_uint32)]),
    "drx_program_release_all": (C.c_int, [C.c_void_p]),
    "drx_program_pack": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                   C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                   C.c_int, C.c_char_p, C.c_int]),
    "drx_walk_batch": (C.c_int, [C.c_void_p, C.POINTER(Batch), C.POINTER(LaunchCfg)]),
    "drx_plan_create": (C.c_int, [C.c_void_p, C.POINTER(Batch), C.POINTER(LaunchCfg), C.POINTER(C.c_void_p)]),
    "drx_plan_run": (C.c_int, [C.c_void_p]),
    "drx_plan_run_resident": (C.c_int, [C.c_void_p]),
    "drx_plan_fetch": (C.c_int, [C.c_void_p]),
    "drx_plan_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "drx_plan_destroy": (None, [C.c_void_p]),
    "drx_crt": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "drx_crt_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "drx_topk": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "drx_copy_calibrate": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p]),
    "drx_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]),
    "drx_host_alloc": (C.c_int, [C.c_uint64, C.c_void_p]),
    "drx_host_free": (None, [C.c_void_p]),
    "drx_last_timing": (C.c_int, [C.c_void_p, C.c_void_p]),
    "drx_valu_rate": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "drx_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerCfg), C.POINTER(C.c_void_p)]),
    "drx_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Walk), C.c_void_p, C.POINTER(C.c_uint64)]),
    "drx_server_poll": (C.c_int, [C.c_void_p, C.POINTER(ServerResult), C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                  C.POINTER(C.c_int)]),
    "drx_server_pending": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "drx_server_capacity": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "drx_server_destroy": (None, [C.c_void_p]),
}

_lib = None
_lib_lock = threading.Lock()


def load_library(path: Optional[str] = None):
    """Load and type the shared library (no device is touched)."""
    global _lib
    with _lib_lock:
        if _lib is not None and path is None:
            return _lib
        p = path or os.environ.get("DRX_LIBRARY", LIB_PATH)
        if not os.path.exists(p):
            raise NativeUnavailable(f"{p} is missing: build it with dreams_rns._build.build_library()")
        lib = C.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.drx_abi_version() != ABI_VERSION:
            raise NativeUnavailable("libdreams_hip ABI version mismatch")
        if path is None:
            _lib = lib
        return lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class Context:
    """One HIP device context of the engine plus its prime table and program cache."""

    def __init__(self, device: int = 0, n_primes: int = 256):
        self.lib = load_library()
        n = C.c_int(0)
        self.lib.drx_device_count(C.byref(n))
        if n.value <= 0:
            raise NativeUnavailable("no HIP device is visible")
        h = C.c_void_p()
        self._check(self.lib.drx_ctx_create(device, C.byref(h)), None)
        self.h = h
        self.device = device
        self.primes = np.zeros(0, dtype=np.int64)
        self._programs: Dict[tuple, Tuple[int, int]] = {}  # key -> (program id, primes its constants cover)
        self._lock = threading.RLock()
        self.ensure_primes(n_primes)

    # -- errors
    def _check(self, rc: int, handle=0):
        if rc == 0:
            return
        h = self.h if handle == 0 else handle
        msg = self.lib.drx_last_error(h).decode() if h else "libdreams_hip error"
        raise _ERRORS.get(rc, RuntimeError)(f"libdreams_hip: {msg} (code {rc})")

    def close(self):
        if getattr(self, "h", None):
            self.lib.drx_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    # -- primes
    def ensure_primes(self, K: int) -> None:
        from .primes import generate_rns_primes

        with self._lock:
            if K <= len(self.primes):
                return
            if K > MAX_K:
                raise ValueError(f"K={K} above the {MAX_K} primes of one context")
            want = min(max(K, 2 * len(self.primes), 64), MAX_K)
            pr = generate_rns_primes(want)
            if len(pr) < K:
                raise ValueError(f"only {len(pr)} primes in (2^30, 2^31) are available")
            # an extension keeps the loaded programs (each covers its own prefix of primes)
            arr = np.ascontiguousarray(pr, dtype=np.int64)
            self._check(self.lib.drx_set_primes(self.h, _ptr(arr), len(arr)))
            self.primes = arr

    # -- programs
    def program_id(self, program, mode: int = DRX_WALK_MATRIX, K: int = 0) -> int:
        """Upload ``program`` once per packing mode (DRX_WALK_MATRIX: the
        code computes M(s); DRX_WALK_VECTOR: it computes v' = M(s) v), with
        constant residues for at least the first ``K`` primes (a power of two
        of at least 256 primes, so that walks of a few more primes reuse it);
        a wider walk later uploads it again."""
        key = (program.m, program.dim, tuple((int(i.op), i.arg0, i.arg1, i.arg2) for i in program.opcodes),
               tuple(int(c) for c in program.constants), int(mode))
        with self._lock:
            hit = self._programs.get(key)
            if hit is not None and hit[1] >= K:
                return hit[0]
            kcap = 256
            while kcap < K:
                kcap *= 2
            kcap = min(kcap, len(self.primes))
            if kcap < K:
                raise ValueError(f"K={K} above the {len(self.primes)} primes set")
            ops = np.array(key[2], dtype=np.uint32).reshape(-1, 4) if key[2] else np.zeros((0, 4), np.uint32)
            ops = np.ascontiguousarray(ops)
            consts = [int(c) for c in program.constants]
            nc = len(consts)
            out = C.c_int(-1)
            if all(-(1 << 63) <= c < (1 << 63) for c in consts):
                carr = np.array(consts, dtype=np.int64) if nc else np.zeros(1, np.int64)
                rc = self.lib.drx_program_upload_mode(self.h, _ptr(ops), len(ops), program.m, program.dim,
                                                      _ptr(carr), None, None, nc, kcap, int(mode), C.byref(out))
            else:
                pr = [int(p) for p in self.primes[:kcap]]
                res = np.array([[c % p for p in pr] for c in consts], dtype=np.uint32)
                flt = np.array([float(c) for c in consts], dtype=np.float64)
                rc = self.lib.drx_program_upload_mode(self.h, _ptr(ops), len(ops), program.m, program.dim,
                                                      None, _ptr(res), _ptr(flt), nc, kcap, int(mode),
                                                      C.byref(out))
            self._check(rc)
            self._programs[key] = (out.value, kcap)
            return out.value

    def program_info(self, pid: int) -> Tuple[int, int, int]:
        n_ops, n_c, kinds = C.c_int(), C.c_int(), C.c_uint32()
        self._check(self.lib.drx_program_info(self.h, pid, C.byref(n_ops), C.byref(n_c), C.byref(kinds)))
        return n_ops.value, n_c.value, kinds.value

    def last_timing(self) -> np.ndarray:
        t = np.zeros(6, dtype=np.float64)
        self._check(self.lib.drx_last_timing(self.h, _ptr(t)))
        return t

    def release_programs(self) -> None:
        """Drop every uploaded program; plans built before are refused."""
        with self._lock:
            self._check(self.lib.drx_program_release_all(self.h))
            self._programs.clear()

    def topk(self, values, k: int) -> np.ndarray:
        """Indices of the ``k`` largest ``values`` on the device (NaN last,
        ties to the lower index), as the batched path ranks delta_float."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        idx = np.zeros(max(1, int(k)), dtype=np.int32)
        self._check(self.lib.drx_topk(self.h, _ptr(v), len(v), int(k), _ptr(idx)))
        return idx[: int(k)]

    def valu_rate(self, op: int):
        """(measured ops/s, peak lane-ops/s) of VALU op 0 FP64 FMA, 1 32-bit
        multiply-high, 2 32x32+64 multiply-add, 3 the residue modmul."""
        r, pk = C.c_double(0.0), C.c_double(0.0)
        self._check(self.lib.drx_valu_rate(self.h, int(op), C.byref(r), C.byref(pk)))
        return r.value, pk.value

    def copy_calibrate(self, nbytes: int, reps: int = 10) -> float:
        """Mean device ms of a plain copy of ``nbytes`` (reads and writes
        exactly ``nbytes`` each): byte-counter calibration and achievable
        copy bandwidth."""
        ms = C.c_double(0.0)
        self._check(self.lib.drx_copy_calibrate(self.h, int(nbytes), int(reps), C.byref(ms)))
        return ms.value


_ctx: Optional[Context] = None
_ctx_lock = threading.Lock()


def context(device: int = 0) -> Context:
    """Process-wide engine context on ``device`` (created on first use)."""
    global _ctx
    with _ctx_lock:
        if _ctx is None:
            _ctx = Context(device)
        return _ctx


def device_count() -> int:
    lib = load_library()
    n = C.c_int(0)
    lib.drx_device_count(C.byref(n))
    return n.value


def pinned_zeros(shape, dtype) -> np.ndarray:
    """Zero-filled array in page-locked host memory (freed with the array)."""
    lib = load_library()
    dt = np.dtype(dtype)
    count = int(np.prod(shape, dtype=np.int64)) if len(shape) else 1
    nbytes = max(1, count * dt.itemsize)
    ptr = C.c_void_p()
    if lib.drx_host_alloc(nbytes, C.byref(ptr)) != 0 or not ptr.value:
        raise MemoryError(f"page-locked allocation of {nbytes} bytes failed")
    buf = (C.c_char * nbytes).from_address(ptr.value)
    weakref.finalize(buf, lib.drx_host_free, ptr.value)
    arr = np.frombuffer(buf, dtype=np.uint8, count=count * dt.itemsize).view(dt).reshape(shape)
    arr[...] = 0
    return arr


def pack_program(program, mode: int = DRX_WALK_MATRIX):
    """Packed device code of ``program`` (no device needed): (words uint32,
    const_slots int list, meta dict). ``words`` holds the residue code (one
    word per instruction, ``meta["words"]``) and then, from the next even
    word, the FP64 code (two words per instruction, ``meta["fwords"]``);
    ``meta["derived"]`` lists the folded constants as (op, a, b) triples.
    Constants must fit in int64."""
    lib = load_library()
    ops = np.ascontiguousarray(
        np.array([(int(i.op), i.arg0, i.arg1, i.arg2) for i in program.opcodes] or [(12, 0, 0, 0)],
                 dtype=np.uint32))
    consts = np.array([int(c) for c in program.constants] or [0], dtype=np.int64)
    n = len(program.opcodes)
    cap = 16 * n + 256
    words = np.zeros(cap, dtype=np.uint32)
    slots = np.zeros(n + len(program.constants) + 1, dtype=np.int32)
    derived = np.zeros((n + 1, 3), dtype=np.int32)
    meta = np.zeros(PACK_META, dtype=np.int64)
    err = C.create_string_buffer(256)
    rc = lib.drx_program_pack(ops.ctypes.data, n, program.m, program.dim, consts.ctypes.data,
                              len(program.constants), int(mode), words.ctypes.data, cap, slots.ctypes.data, len(slots),
                              derived.ctypes.data, len(derived), meta.ctypes.data, PACK_META, err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    nw, nfw = int(meta[0]), int(meta[8])
    total = nw + (nw & 1) + nfw
    if total > cap:
        raise RuntimeError("packed code larger than the buffer")
    return words[:total].copy(), [int(s) for s in slots[: int(meta[5])]], {
        "words": nw, "kinds": int(meta[1]), "colred": int(meta[2]), "const_lds": int(meta[3]),
        "nslots": int(meta[4]), "nconst": int(meta[5]), "fd_ok": int(meta[6]), "fd_degree": int(meta[7]),
        "fwords": nfw, "fnslots": int(meta[9]), "inv_batched": int(meta[11]), "src_ops": int(meta[12]),
        "derived": [tuple(int(v) for v in derived[i]) for i in range(int(meta[10]))]}

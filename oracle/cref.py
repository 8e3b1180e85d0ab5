"""ctypes wrapper of oracle/cwalk.c (test infrastructure only).

``build()`` compiles the C restatement with gcc into oracle/_build/.
``walk(...)`` runs one walk with the reference semantics and returns the
final residue matrix, the final FP64 shadow matrix and log_scale.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cwalk.c")
LIB = os.path.join(HERE, "_build", "libcwalk.so")
_lib = None


def build() -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-builtin-pow",
           SRC, "-o", LIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.cw_walk.restype = C.c_int
        lib.cw_walk.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.POINTER(C.c_double)]
        _lib = lib
    return _lib


def walk(program, depth: int, primes, axes, init_res=None, init_f=None):
    """Walk ``program`` for ``depth`` steps over ``primes``.

    axes: (sn, sd, tn, td) per axis; init_res (m*m, K) canonical residues and
    init_f (m*m,) default to the identity. Returns (res (m, m, K) int64,
    P_f (m, m) float64, log_scale float). Constants must fit in int64.
    """
    lib = _load()
    m, dim = program.m, program.dim
    primes = np.ascontiguousarray(primes, dtype=np.int64)
    K = len(primes)
    ops = np.array([(int(i.op), i.arg0, i.arg1, i.arg2) for i in program.opcodes] or [(12, 0, 0, 0)],
                   dtype=np.uint32)
    consts = np.array([int(c) for c in program.constants] or [0], dtype=np.int64)
    ax = np.array(list(axes) or [(0, 1, 0, 1)], dtype=np.int64).reshape(-1, 4)
    if init_res is None:
        init_res = np.zeros((m * m, K), dtype=np.int64)
        init_f = np.zeros(m * m)
        for d in range(m):
            init_res[d * m + d] = 1
            init_f[d * m + d] = 1.0
    init_res = np.ascontiguousarray(init_res, dtype=np.int64).reshape(m * m, K)
    init_f = np.ascontiguousarray(init_f, dtype=np.float64).reshape(m * m)
    out_res = np.zeros((m * m, K), dtype=np.int64)
    out_f = np.zeros(m * m)
    ls = C.c_double(0.0)
    rc = lib.cw_walk(m, dim, ops.ctypes.data, len(program.opcodes), consts.ctypes.data,
                     len(program.constants), primes.ctypes.data, K, int(depth), init_res.ctypes.data,
                     init_f.ctypes.data, ax.ctypes.data, out_res.ctypes.data, out_f.ctypes.data, C.byref(ls))
    if rc != 0:
        raise RuntimeError("cw_walk failed")
    return out_res.reshape(m, m, K), out_f.reshape(m, m), ls.value


def pcf_walk(program, a0: int, depth: int, primes, shift_val: int = 1):
    """run_pcf_walk's walk (runner.py:338-413) through the C restatement."""
    primes = np.asarray(primes, dtype=np.int64)
    K = len(primes)
    init_res = np.zeros((4, K), dtype=np.int64)
    init_res[0] = 1
    init_res[1] = [int(a0) % int(p) for p in primes]
    init_res[3] = 1
    init_f = np.array([1.0, float(a0), 0.0, 1.0])
    axes = [(shift_val, 1, int(d), 1) for d in program.directions[: program.dim]]
    return walk(program, depth, primes, axes, init_res, init_f)

"""dreams_rns -- PCF/CMF walk engine for AMD MI355X (gfx950).

Drop-in for the walk path of VesterlundCoder/Dreams-RNS-CUDA: the same public
names and signatures (python/dreams_rns/__init__.py:11-48), with walks, CRT,
the FP64 shadow and top-k executed by hand-written HIP kernels in
libdreams_hip.so through a ctypes C ABI (include/dreams_hip.h).

Convention: M(n) = [[0, b(n)], [1, a(n)]], P(N) = A * M(1) * ... * M(N),
A = [[1, a(0)], [0, 1]], p = P[0, m-1], q = P[1, m-1],
delta = -(1 + log|p/q - L| / log|q|).
"""

from .compiler import CmfCompiler, CmfProgram, Instruction, Opcode, compile_cmf, compile_cmf_from_dict
from .runner import (
    WalkConfig,
    centered,
    compile_pcf,
    compute_dreams_delta_exact,
    compute_dreams_delta_float,
    crt_reconstruct,
    generate_rns_primes,
    pcf_initial_values,
    run_pcf_walk,
    verify_pcf,
)
from .cmf_walk import compile_cmf_spec, run_cmf_walk, run_cmf_walk_vec, run_cmf_walk_vec_batch, shift_to_axis_offsets
from .gpu_walk import gpu_available, run_cmf_walk_batch_gpu, run_cmf_walk_gpu, run_pcf_walk_batch_gpu, run_pcf_walk_gpu

__version__ = "0.2.0"
__all__ = [
    "compile_cmf",
    "compile_cmf_from_dict",
    "CmfCompiler",
    "CmfProgram",
    "Instruction",
    "Opcode",
    "WalkConfig",
    "generate_rns_primes",
    "crt_reconstruct",
    "centered",
    "compute_dreams_delta_float",
    "compute_dreams_delta_exact",
    "compile_pcf",
    "pcf_initial_values",
    "run_pcf_walk",
    "verify_pcf",
    "run_cmf_walk",
    "run_cmf_walk_vec",
    "run_cmf_walk_vec_batch",
    "compile_cmf_spec",
    "shift_to_axis_offsets",
    "gpu_available",
    "run_pcf_walk_gpu",
    "run_pcf_walk_batch_gpu",
    "run_cmf_walk_gpu",
    "run_cmf_walk_batch_gpu",
]

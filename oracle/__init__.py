"""CPU oracle for the PCF/CMF walk engine -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference algorithm of
VesterlundCoder/Dreams-RNS-CUDA for the walk path, and adds independent
checkers that pin it. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
or as the timed CPU baseline. The product package never imports it.

Modules
-------
``walk``   numpy restatement of python/dreams_rns/runner.py and
           python/dreams_rns/cmf_walk.py: per step a numpy-over-K bytecode
           interpretation with a fresh 512-register file, a Python-level
           Fermat INV, an int64 ``%`` matrix product, and the OpenBLAS
           ``P_f @ M_f`` FP64 shadow with the ``> 1e10`` rescale.  It is
           the "port" CPU baseline.
``exact``  independent pins: an exact big-integer / Fraction walk, a
           pure-Python modular bytecode interpreter (INV(0) = 0), a
           deterministic FMA-chain FP64 shadow, CRT by textbook Garner.
``cref``   ctypes binding to ``oracle/cwalk.c`` (gcc), a C restatement of
           the residue walk and of the FMA-chain shadow, fast enough to
           check thousands of walks.
``postwalk`` the CMF sweep's per-walk post-processing (CRT check, constant
           proximity, delta against the bank) restated walk by walk.

Parity status of each function is listed in the module docstrings; nothing
here is "parity unpinned" except where a docstring says so.
"""

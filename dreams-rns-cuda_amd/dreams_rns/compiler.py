"""Bytecode ISA and the SymPy -> bytecode compiler for PCF/CMF matrix entries.

The instruction set, its 16-byte packing, the program serialisation and the
order in which SymPy trees are folded into instructions are an interface: the
FP64 shadow walk evaluates the same bytecode in floating point, so two
compilers that fold ``a + b + c`` differently produce different shadow bits.
They therefore follow the reference exactly:

* ISA and packing ........ python/dreams_rns/compiler.py:14-41
* CmfProgram (de)serialise python/dreams_rns/compiler.py:44-89
* expression lowering .... python/dreams_rns/compiler.py:200-278
* entry / register reset . python/dreams_rns/compiler.py:280-306
* dict front end ......... python/dreams_rns/compiler.py:367-404
* ramanujantools front end python/dreams_rns/compiler.py:309-364

The lowering rules (one fresh register per produced value; registers reclaimed
after every STORE):

====================  =====================================================
SymPy node            emitted code
====================  =====================================================
Integer c             LOAD_C idx(c) (constants de-duplicated by value)
Symbol ``n``          LOAD_N
axis symbol           LOAD_X axis
Add(a0, a1, ...)      a0; a1; ADD; a2; ADD ...   (left fold, SymPy arg order)
Mul(-1, rest...)      Mul(rest...); NEG
Mul(a0, a1, ...)      left fold with MUL
Pow(b, 2) / (b, 3)    b; POW2 / POW3
Pow(b, -1)            b; INV
Pow(b, k>=4)          b; (k-1) x MUL(acc, b)
Pow(b, k<=-2)         b; (|k|-1) x MUL(acc, b); INV
Rational p/q          LOAD_C p; LOAD_C q; INV; MUL
====================  =====================================================
"""

from __future__ import annotations

import struct
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Dict, List, Optional, Sequence, Tuple

__all__ = [
    "Opcode",
    "Instruction",
    "CmfProgram",
    "CmfCompiler",
    "compile_cmf",
    "compile_cmf_from_dict",
]


class Opcode(IntEnum):
    """Bytecode opcodes (python/dreams_rns/compiler.py:14-28)."""

    NOP = 0
    LOAD_X = 1   # arg0 = axis, arg1 = dst
    LOAD_C = 2   # arg0 = constant index, arg1 = dst
    LOAD_N = 3   # arg0 = dst
    ADD = 4      # arg0, arg1 -> arg2
    SUB = 5
    MUL = 6
    NEG = 7      # arg0 -> arg1
    POW2 = 8
    POW3 = 9
    INV = 10
    STORE = 11   # arg0 = src, arg1 = row, arg2 = col
    END = 12


_WORD = struct.Struct("<IIII")


@dataclass
class Instruction:
    """One 16-byte instruction: opcode and three unsigned 32-bit arguments."""

    op: Opcode
    arg0: int = 0
    arg1: int = 0
    arg2: int = 0

    def pack(self) -> bytes:
        return _WORD.pack(int(self.op), self.arg0, self.arg1, self.arg2)


@dataclass
class CmfProgram:
    """A compiled m x m matrix field over ``dim`` axes.

    ``constants`` holds Python ints of any size; ``directions`` is the
    default per-axis step used by walks that do not pass a trajectory.
    """

    m: int
    dim: int
    opcodes: List[Instruction] = field(default_factory=list)
    constants: List[int] = field(default_factory=list)
    directions: List[int] = field(default_factory=list)

    def serialize(self) -> bytes:
        """Header ``IIII`` (m, dim, n_ops, n_const), 16 B per instruction,
        int64 constants, int32 directions (compiler.py:53-62)."""
        parts = [_WORD.pack(self.m, self.dim, len(self.opcodes), len(self.constants))]
        parts.extend(ins.pack() for ins in self.opcodes)
        parts.extend(struct.pack("q", c) for c in self.constants)
        parts.extend(struct.pack("i", d) for d in self.directions)
        return b"".join(parts)

    @classmethod
    def deserialize(cls, data: bytes) -> "CmfProgram":
        """Inverse of :meth:`serialize` (compiler.py:64-89)."""
        m, dim, n_ops, n_const = _WORD.unpack_from(data, 0)
        pos = _WORD.size
        ops = []
        for _ in range(n_ops):
            op, a0, a1, a2 = _WORD.unpack_from(data, pos)
            ops.append(Instruction(Opcode(op), a0, a1, a2))
            pos += _WORD.size
        consts = list(struct.unpack_from(f"{n_const}q", data, pos)) if n_const else []
        pos += 8 * n_const
        dirs = list(struct.unpack_from(f"{dim}i", data, pos)) if dim else []
        return cls(m=m, dim=dim, opcodes=ops, constants=consts, directions=dirs)


class CmfCompiler:
    """Incremental compiler: one ``compile_matrix_entry`` call per entry,
    then :meth:`build` (compiler.py:92-306)."""

    MAX_REGS = 512

    def __init__(self, m: int, dim: int, directions: Optional[List[int]] = None):
        self.m = m
        self.dim = dim
        self.directions = directions or [1] * dim
        self.opcodes: List[Instruction] = []
        self.constants: Dict[int, int] = {}
        self.next_reg = 0
        self.entry_regs: Dict[Tuple[int, int], int] = {}

    # -- register / constant bookkeeping ---------------------------------
    def _alloc_reg(self) -> int:
        reg = self.next_reg
        self.next_reg = reg + 1
        if self.next_reg >= self.MAX_REGS:
            raise RuntimeError("Out of registers")
        return reg

    def _add_const(self, value: int) -> int:
        return self.constants.setdefault(value, len(self.constants))

    def _emit(self, op: Opcode, arg0: int = 0, arg1: int = 0, arg2: int = 0) -> None:
        self.opcodes.append(Instruction(op, arg0, arg1, arg2))

    def _unary(self, op: Opcode, src: int) -> int:
        dst = self._alloc_reg()
        self._emit(op, src, dst)
        return dst

    def _binary(self, op: Opcode, a: int, b: int) -> int:
        dst = self._alloc_reg()
        self._emit(op, a, b, dst)
        return dst

    # -- public single-instruction helpers -------------------------------
    def load_axis(self, axis: int) -> int:
        dst = self._alloc_reg()
        self._emit(Opcode.LOAD_X, axis, dst)
        return dst

    def load_const(self, value: int) -> int:
        idx = self._add_const(value)
        dst = self._alloc_reg()
        self._emit(Opcode.LOAD_C, idx, dst)
        return dst

    def load_n(self) -> int:
        dst = self._alloc_reg()
        self._emit(Opcode.LOAD_N, dst)
        return dst

    def add(self, r1: int, r2: int) -> int:
        return self._binary(Opcode.ADD, r1, r2)

    def sub(self, r1: int, r2: int) -> int:
        return self._binary(Opcode.SUB, r1, r2)

    def mul(self, r1: int, r2: int) -> int:
        return self._binary(Opcode.MUL, r1, r2)

    def neg(self, r: int) -> int:
        return self._unary(Opcode.NEG, r)

    def pow2(self, r: int) -> int:
        return self._unary(Opcode.POW2, r)

    def pow3(self, r: int) -> int:
        return self._unary(Opcode.POW3, r)

    def inv(self, r: int) -> int:
        return self._unary(Opcode.INV, r)

    def store(self, r: int, row: int, col: int) -> None:
        self._emit(Opcode.STORE, r, row, col)
        self.entry_regs[(row, col)] = r

    # -- lowering ----------------------------------------------------------
    def compile_sympy_expr(self, expr, axis_symbols: Dict[str, int]) -> int:
        """Lower one SymPy expression; returns the register holding it."""
        import sympy as sp

        if isinstance(expr, (int, sp.Integer)):
            return self.load_const(int(expr))
        if isinstance(expr, sp.Symbol):
            return self._lower_symbol(str(expr), axis_symbols)
        if isinstance(expr, sp.Add):
            return self._fold(expr.args, self.add, axis_symbols)
        if isinstance(expr, sp.Mul):
            args = expr.args
            if args[0] == -1:
                inner = self.compile_sympy_expr(sp.Mul(*args[1:]), axis_symbols)
                return self.neg(inner)
            return self._fold(args, self.mul, axis_symbols)
        if isinstance(expr, sp.Pow):
            return self._lower_pow(expr, axis_symbols)
        if isinstance(expr, sp.Rational):
            num = self.load_const(int(expr.p))
            den = self.load_const(int(expr.q))
            return self.mul(num, self.inv(den))
        raise ValueError(f"Unsupported expression type: {type(expr)}")

    def _lower_symbol(self, name: str, axis_symbols: Dict[str, int]) -> int:
        if name == "n":
            return self.load_n()
        if name in axis_symbols:
            return self.load_axis(axis_symbols[name])
        raise ValueError(f"Unknown symbol: {name}")

    def _fold(self, args: Sequence, combine, axis_symbols: Dict[str, int]) -> int:
        acc = self.compile_sympy_expr(args[0], axis_symbols)
        for term in args[1:]:
            acc = combine(acc, self.compile_sympy_expr(term, axis_symbols))
        return acc

    def _lower_pow(self, expr, axis_symbols: Dict[str, int]) -> int:
        import sympy as sp

        base, exp = expr.args
        if exp == 2:
            return self.pow2(self.compile_sympy_expr(base, axis_symbols))
        if exp == 3:
            return self.pow3(self.compile_sympy_expr(base, axis_symbols))
        if exp == -1:
            return self.inv(self.compile_sympy_expr(base, axis_symbols))
        k = int(exp) if isinstance(exp, (int, sp.Integer)) else 0
        if k == 0:
            raise ValueError(f"Unsupported power: {expr}")
        b = self.compile_sympy_expr(base, axis_symbols)
        acc = b
        for _ in range(abs(k) - 1):
            acc = self.mul(acc, b)
        return acc if k > 0 else self.inv(acc)

    def compile_matrix_entry(self, row: int, col: int, expr, axis_symbols: Dict[str, int]) -> None:
        reg = self.compile_sympy_expr(expr, axis_symbols)
        self.store(reg, row, col)
        self.next_reg = 0

    def build(self) -> CmfProgram:
        self._emit(Opcode.END)
        consts = [0] * len(self.constants)
        for value, idx in self.constants.items():
            consts[idx] = value
        return CmfProgram(
            m=self.m,
            dim=self.dim,
            opcodes=list(self.opcodes),
            constants=consts,
            directions=list(self.directions),
        )


def compile_cmf(cmf, axis_names: Optional[List[str]] = None) -> CmfProgram:
    """Compile an object exposing a SymPy matrix as ``.M`` or ``.matrix``
    (a ramanujantools CMF, compiler.py:309-364). Zero entries are skipped."""
    if hasattr(cmf, "M"):
        mat = cmf.M
    elif hasattr(cmf, "matrix"):
        mat = cmf.matrix
    else:
        raise ValueError("CMF object must have 'M' or 'matrix' attribute")
    m = mat.rows
    if axis_names is None:
        if hasattr(cmf, "symbols"):
            axis_names = [str(s) for s in cmf.symbols]
        else:
            axis_names = sorted(str(s) for s in mat.free_symbols if str(s) != "n")
    dim = len(axis_names)
    axis_symbols = {name: i for i, name in enumerate(axis_names)}
    directions = list(cmf.directions) if hasattr(cmf, "directions") else [1] * dim
    comp = CmfCompiler(m=m, dim=dim, directions=directions)
    for i in range(m):
        for j in range(m):
            entry = mat[i, j]
            if entry != 0:
                comp.compile_matrix_entry(i, j, entry, axis_symbols)
    return comp.build()


def compile_cmf_from_dict(
    matrix_dict: Dict[Tuple[int, int], str],
    m: int,
    dim: int,
    axis_names: Optional[List[str]] = None,
    directions: Optional[List[int]] = None,
) -> CmfProgram:
    """Compile ``{(row, col): "expr"}`` in dict order (compiler.py:367-404).

    Zero entries are compiled (LOAD_C 0; STORE) rather than skipped, exactly
    like the reference front end.
    """
    import sympy as sp

    if axis_names is None:
        axis_names = [f"x{i}" for i in range(dim)]
    axis_symbols = {name: i for i, name in enumerate(axis_names)}
    local_ns = {name: sp.Symbol(name) for name in axis_names}
    local_ns["n"] = sp.Symbol("n")
    comp = CmfCompiler(m=m, dim=dim, directions=directions or [1] * dim)
    for (row, col), text in matrix_dict.items():
        comp.compile_matrix_entry(row, col, sp.sympify(text, locals=local_ns), axis_symbols)
    return comp.build()

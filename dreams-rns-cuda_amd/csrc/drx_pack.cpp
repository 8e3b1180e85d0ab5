// Host-side program packer.
//
// The reference bytecode (python/dreams_rns/compiler.py:14-41) is SSA-like:
// every instruction writes a fresh register of a 512-entry file that is
// zeroed at the start of every step (runner.py:217), and STORE copies a
// register into a matrix entry. Packing:
//   1. rename registers to values (reads of never-written registers read 0);
//   2. keep only instructions reachable from the final STORE of each entry;
//   3. classify entries as zero / one / general (static per program);
//   4. insert RED where a sum could exceed 2^40 (the exactness bound of the
//      FP64 residue multiply), tracking magnitude bounds per value;
//   5. route operands: axes / step / constants read their banks directly,
//      a value used by the very next instruction is read from the
//      accumulator, everything else gets a register by linear scan (16 VGPR
//      slots, then LDS slots).
// Instruction order is preserved, so the FP64 shadow evaluates exactly the
// reference's operations in the reference's order.
#include "drx_pack.h"

#include <algorithm>
#include <cstdio>

#include "drx_device.h"
#include "drx_types.h"

namespace drx {

namespace {

enum RefOp : uint32_t {
  R_NOP = 0, R_LOAD_X = 1, R_LOAD_C = 2, R_LOAD_N = 3, R_ADD = 4, R_SUB = 5, R_MUL = 6,
  R_NEG = 7, R_POW2 = 8, R_POW3 = 9, R_INV = 10, R_STORE = 11, R_END = 12
};

enum ValKind { V_ZERO, V_X, V_N, V_C, V_OP };

struct Val {
  ValKind kind;
  int idx = 0;      // axis or constant index
  uint32_t op = 0;  // device op for V_OP
  int a = 0, b = 0; // operand values (V_OP)
};

struct Ins {            // device instruction before register assignment
  uint32_t op;
  int a, b;             // operand value ids (-1: none)
  int def;              // value id defined (-1: none)
  int entry = -1;       // >= 0: result goes to entry bank E[entry] (trailing MOVs)
};

constexpr double kBoundLimit = 1099511627776.0;  // 2^40
constexpr double kMulBound = 2147483648.0;       // 2^31

}  // namespace

std::string pack_program(const uint32_t* ops, int n_ops, int m, int dim, int n_const,
                         const std::vector<char>& const_is_one, const std::vector<char>& const_is_zero,
                         PackedProgram* out) {
  char msg[160];
  if (m < 2 || m > kMaxM) {
    std::snprintf(msg, sizeof msg, "matrix order m=%d outside [2, %d]", m, kMaxM);
    return msg;
  }
  if (dim < 0 || dim > kMaxDim) {
    std::snprintf(msg, sizeof msg, "dim=%d outside [0, %d]", dim, kMaxDim);
    return msg;
  }
  std::vector<Val> vals;
  vals.push_back(Val{V_ZERO});
  std::vector<int> reg(512, 0);
  std::vector<int> entry(m * m, 0);
  int n_src = 0;
  PackedProgram pp;
  pp.m = m;
  pp.dim = dim;
  auto chk_reg = [&](uint32_t r) -> bool { return r < 512u; };
  for (int i = 0; i < n_ops; ++i) {
    const uint32_t op = ops[4 * i], a0 = ops[4 * i + 1], a1 = ops[4 * i + 2], a2 = ops[4 * i + 3];
    if (op == R_END) break;
    ++n_src;
    switch (op) {
      case R_NOP:
        break;
      case R_LOAD_X:
        if ((int)a0 >= dim || !chk_reg(a1)) goto bad_operand;
        vals.push_back(Val{V_X, (int)a0});
        reg[a1] = (int)vals.size() - 1;
        break;
      case R_LOAD_C:
        if ((int)a0 >= n_const || !chk_reg(a1)) goto bad_operand;
        vals.push_back(Val{V_C, (int)a0});
        reg[a1] = (int)vals.size() - 1;
        break;
      case R_LOAD_N:
        if (!chk_reg(a0)) goto bad_operand;
        vals.push_back(Val{V_N});
        reg[a0] = (int)vals.size() - 1;
        break;
      case R_ADD:
      case R_SUB:
      case R_MUL: {
        if (!chk_reg(a0) || !chk_reg(a1) || !chk_reg(a2)) goto bad_operand;
        Val v{V_OP};
        v.op = op == R_ADD ? D_ADD : (op == R_SUB ? D_SUB : D_MUL);
        v.a = reg[a0];
        v.b = reg[a1];
        vals.push_back(v);
        reg[a2] = (int)vals.size() - 1;
        if (op == R_ADD) ++pp.n_add;
        else if (op == R_SUB) ++pp.n_sub;
        else ++pp.n_mul;
        break;
      }
      case R_NEG:
      case R_POW2:
      case R_POW3:
      case R_INV: {
        if (!chk_reg(a0) || !chk_reg(a1)) goto bad_operand;
        Val v{V_OP};
        v.op = op == R_NEG ? D_NEG : (op == R_POW2 ? D_SQR : (op == R_POW3 ? D_CUB : D_INV));
        v.a = reg[a0];
        v.b = 0;
        vals.push_back(v);
        reg[a1] = (int)vals.size() - 1;
        if (op == R_NEG) ++pp.n_neg;
        else if (op == R_POW2) ++pp.n_sqr;
        else if (op == R_POW3) ++pp.n_cub;
        else ++pp.n_inv;
        break;
      }
      case R_STORE:
        if (!chk_reg(a0) || (int)a1 >= m || (int)a2 >= m) goto bad_operand;
        entry[a1 * m + a2] = reg[a0];
        break;
      default:
        std::snprintf(msg, sizeof msg, "instruction %d: unknown opcode %u", i, op);
        return msg;
    }
    continue;
  bad_operand:
    std::snprintf(msg, sizeof msg, "instruction %d (op %u): operand out of range", i, op);
    return msg;
  }
  pp.n_src_ops = n_src;

  // ---- entry kinds and liveness
  const int nv = (int)vals.size();
  std::vector<char> needed(nv, 0);
  std::vector<int> stack;
  auto is_const_val = [&](int v, const std::vector<char>& tab) {
    return vals[v].kind == V_C && tab[vals[v].idx];
  };
  for (int e = 0; e < m * m; ++e) {
    const int v = entry[e];
    uint32_t kind = E_GEN;
    if (vals[v].kind == V_ZERO || is_const_val(v, const_is_zero)) kind = E_ZERO;
    else if (is_const_val(v, const_is_one)) kind = E_ONE;
    pp.kinds |= kind << (2 * e);
    if (kind == E_GEN) stack.push_back(v);
  }
  while (!stack.empty()) {
    const int v = stack.back();
    stack.pop_back();
    if (needed[v]) continue;
    needed[v] = 1;
    if (vals[v].kind == V_OP) {
      stack.push_back(vals[v].a);
      stack.push_back(vals[v].b);
    }
  }
  // polynomial degree of every needed value in the step (-1: not a polynomial)
  {
    std::vector<int> deg(nv, 0);
    for (int v = 0; v < nv; ++v) {
      const Val& x = vals[v];
      switch (x.kind) {
        case V_ZERO: deg[v] = 0; break;
        case V_C: deg[v] = 0; break;
        case V_X: deg[v] = 1; break;
        case V_N: deg[v] = 1; break;
        default: {
          const int da = deg[x.a], db = deg[x.b];
          if (x.op == D_INV || da < 0 || (db < 0 && (x.op == D_ADD || x.op == D_SUB || x.op == D_MUL))) {
            deg[v] = -1;
          } else if (x.op == D_ADD || x.op == D_SUB) {
            deg[v] = std::max(da, db);
          } else if (x.op == D_MUL) {
            deg[v] = da + db;
          } else if (x.op == D_SQR) {
            deg[v] = 2 * da;
          } else if (x.op == D_CUB) {
            deg[v] = 3 * da;
          } else {
            deg[v] = da;  // NEG
          }
          if (deg[v] > 15) deg[v] = -1;  // beyond any table the kernels keep
        }
      }
    }
    int ok = 1, dmax = 0;
    uint64_t packed = 0;
    for (int e = 0; e < m * m; ++e) {
      if (((pp.kinds >> (2 * e)) & 3u) != E_GEN) continue;
      const int d = deg[entry[e]];
      if (d < 0) {
        ok = 0;
        break;
      }
      dmax = std::max(dmax, d);
      packed |= (uint64_t)d << (4 * e);
    }
    pp.fd_ok = ok;
    pp.fd_degree = ok ? dmax : 0;
    pp.fd_deg = ok ? packed : 0;
  }
  for (int j = 0; j < m; ++j) {
    int nz = 0, ones = 0;
    for (int k = 0; k < m; ++k) {
      const uint32_t kd = (pp.kinds >> (2 * (k * m + j))) & 3u;
      nz += kd != E_ZERO;
      ones += kd == E_ONE;
    }
    if (ones > 0 && nz >= 2) pp.colred |= 1u << j;
  }

  // ---- emit (with RED for magnitude control), values renumbered by Ins
  std::vector<double> bound(nv, 0.0);
  std::vector<Ins> code;
  auto src_bound = [&](int v) -> double {
    switch (vals[v].kind) {
      case V_ZERO: return 0.0;
      case V_X: return kMulBound;
      case V_N: return kMulBound;
      case V_C: return kMulBound / 2;
      default: return bound[v];
    }
  };
  for (int v = 0; v < nv; ++v) {
    if (!needed[v] || vals[v].kind != V_OP) continue;
    const Val x = vals[v];
    double bd;
    switch (x.op) {
      case D_ADD:
      case D_SUB: bd = src_bound(x.a) + src_bound(x.b); break;
      case D_NEG: bd = src_bound(x.a); break;
      default: bd = kMulBound; break;
    }
    const bool unary = !(x.op == D_ADD || x.op == D_SUB || x.op == D_MUL);
    if (bd > kBoundLimit) {
      // the sum goes to the accumulator only; RED(A) produces value v
      const int tmp = (int)vals.size();
      vals.push_back(x);
      bound.push_back(bd);
      needed.push_back(1);
      code.push_back(Ins{x.op, x.a, unary ? -1 : x.b, tmp});
      Val r{V_OP};
      r.op = D_RED;
      r.a = tmp;
      vals[v] = r;
      bound[v] = kMulBound;
      code.push_back(Ins{D_RED, tmp, -1, v});
    } else {
      bound[v] = bd;
      code.push_back(Ins{x.op, x.a, unary ? -1 : x.b, v});
    }
  }
  // trailing moves for entries whose value cannot be written into E by its
  // producer: inputs, values shared by several entries or used later
  const int nvals = (int)vals.size();
  std::vector<int> producer(nvals, -1);
  for (int i = 0; i < (int)code.size(); ++i) producer[code[i].def] = i;
  std::vector<int> n_entry_of(nvals, 0);
  for (int e = 0; e < m * m; ++e)
    if (((pp.kinds >> (2 * e)) & 3u) == E_GEN) ++n_entry_of[entry[e]];
  std::vector<int> uses(nvals, 0);
  for (const Ins& in : code) {
    if (in.a >= 0) ++uses[in.a];
    if (in.b >= 0) ++uses[in.b];
  }
  std::vector<int> entry_direct(nvals, -1);  // value -> entry its producer writes
  for (int e = 0; e < m * m; ++e) {
    if (((pp.kinds >> (2 * e)) & 3u) != E_GEN) continue;
    const int v = entry[e];
    if (vals[v].kind == V_OP && n_entry_of[v] == 1 && uses[v] == 0) {
      entry_direct[v] = e;
    } else {
      code.push_back(Ins{D_MOV, v, -1, -1, e});
      ++uses[v];
    }
  }

  // ---- operand routing and register allocation
  const int nc = (int)code.size();
  std::vector<int> def_pos(nvals, -1), last_use(nvals, -1);
  std::vector<char> only_next(nvals, 1);
  for (int i = 0; i < nc; ++i) {
    if (code[i].def >= 0) def_pos[code[i].def] = i;
  }
  for (int i = 0; i < nc; ++i) {
    for (int u : {code[i].a, code[i].b}) {
      if (u < 0 || vals[u].kind != V_OP) continue;
      last_use[u] = std::max(last_use[u], i);
      if (i != def_pos[u] + 1) only_next[u] = 0;
    }
  }
  // constant slots in order of first use
  std::vector<int> cslot(n_const, -1);
  for (int i = 0; i < nc; ++i) {
    for (int u : {code[i].a, code[i].b}) {
      if (u >= 0 && vals[u].kind == V_C && cslot[vals[u].idx] < 0) {
        cslot[vals[u].idx] = (int)pp.const_slots.size();
        pp.const_slots.push_back(vals[u].idx);
      }
    }
  }
  pp.const_lds = pp.const_slots.size() > (size_t)kBank ? 1 : 0;
  if (pp.const_slots.size() > 8191) return "too many constants";

  // linear scan: slot 0..15 = VGPR bank R, 16.. = LDS bank S
  std::vector<int> slot(nvals, -1);
  std::vector<int> free_r, free_s;
  for (int s = kBank - 1; s >= 0; --s) free_r.push_back(s);
  int n_s = 0;
  std::vector<std::vector<int>> dying(nc);
  for (int v = 0; v < nvals; ++v)
    if (last_use[v] >= 0 && def_pos[v] >= 0 && !(only_next[v] && entry_direct[v] < 0))
      dying[last_use[v]].push_back(v);
  std::vector<uint32_t> words;
  words.reserve(2 * nc);
  for (int i = 0; i < nc; ++i) {
    const Ins& in = code[i];
    uint32_t sb[2] = {SB_Z, SB_Z}, si[2] = {0, 0};
    const int opnd[2] = {in.a, in.b};
    for (int q = 0; q < 2; ++q) {
      const int u = opnd[q];
      if (u < 0) continue;
      const Val& x = vals[u];
      if (x.kind == V_ZERO) {
        sb[q] = SB_Z;
      } else if (x.kind == V_X) {
        sb[q] = SB_X;
        si[q] = (uint32_t)x.idx;
      } else if (x.kind == V_N) {
        sb[q] = SB_X;
        si[q] = kStepSlot;
      } else if (x.kind == V_C) {
        sb[q] = pp.const_lds ? SB_L : SB_C;
        si[q] = (uint32_t)cslot[x.idx];
      } else if (def_pos[u] == i - 1) {
        sb[q] = SB_A;
      } else {
        const int s = slot[u];
        if (s < 0) return "internal: operand without a register";
        sb[q] = s < kBank ? SB_R : SB_S;
        si[q] = s < kBank ? (uint32_t)s : (uint32_t)(s - kBank);
      }
    }
    // free registers of values dying here (before assigning the destination)
    for (int v : dying[i]) {
      const int s = slot[v];
      if (s < 0) continue;
      if (s < kBank) free_r.push_back(s);
      else free_s.push_back(s);
    }
    uint32_t db = DB_NONE, di = 0;
    if (in.entry >= 0) {
      db = DB_E;
      di = (uint32_t)in.entry;
    } else if (in.def >= 0) {
      const int v = in.def;
      if (entry_direct[v] >= 0) {
        db = DB_E;
        di = (uint32_t)entry_direct[v];
      } else if (last_use[v] >= 0 && !only_next[v]) {
        int s;
        if (!free_r.empty()) {
          s = free_r.back();
          free_r.pop_back();
        } else if (!free_s.empty()) {
          s = free_s.back();
          free_s.pop_back();
        } else {
          s = kBank + n_s++;
        }
        slot[v] = s;
        db = s < kBank ? DB_R : DB_S;
        di = s < kBank ? (uint32_t)s : (uint32_t)(s - kBank);
      }
    }
    words.push_back(enc_w0(in.op, db, di));
    words.push_back(enc_w1(sb[0], si[0], sb[1], si[1]));
  }
  pp.nslots = n_s;
  pp.code.swap(words);
  *out = std::move(pp);
  return std::string();
}

}  // namespace drx

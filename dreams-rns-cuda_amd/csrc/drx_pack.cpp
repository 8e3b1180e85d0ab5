// Host-side program packer.
//
// The reference bytecode (python/dreams_rns/compiler.py:14-41) is SSA-like:
// every instruction writes a fresh register of a 512-entry file that is
// zeroed at the start of every step (runner.py:217), and STORE copies a
// register into a matrix entry. Packing:
//   1. rename registers to values (reads of never-written registers read 0),
//      one value per distinct operation on the same operands;
//   2. fold operations on constants into derived constants (their residues
//      and FP64 values are computed at upload with the same operations);
//   3. keep only values reachable from a general entry; classify entries as
//      zero / one / general; find the entries' degrees in the step;
//   4. FP64 code: the remaining operations, one reciprocal per INV;
//   5. residue code: step-dependent INVs whose arguments do not depend on
//      another INV share one inversion (Montgomery's trick: prefix products,
//      one INV, back-substitution; a zero argument counts as 1 in the
//      products and its result is masked to 0, so INV(0) = 0 as in the
//      reference); RED is inserted where a sum could exceed 2^40 (the
//      exactness bound of the FP64 residue multiply), tracking magnitude
//      bounds per value;
//   6. residue code: lowering to the R / C register banks (x operand in R,
//      y in R or C, results in R, entries also stored to LDS rows) with
//      furthest-next-use register allocation (spills to LDS rows); FP64
//      code: operand routing over its banks and linear-scan allocation. The
//      evaluation order is the best of a few candidates for each code.
// Reordering independent operations changes no result: every value is the
// same operation on the same operand values, in FP64 as in residues.
#include "drx_pack.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <map>
#include <queue>

#include "drx_device.h"
#include "drx_types.h"

namespace drx {

namespace {

enum RefOp : uint32_t {
  R_NOP = 0, R_LOAD_X = 1, R_LOAD_C = 2, R_LOAD_N = 3, R_ADD = 4, R_SUB = 5, R_MUL = 6,
  R_NEG = 7, R_POW2 = 8, R_POW3 = 9, R_INV = 10, R_STORE = 11, R_END = 12
};

enum ValKind { V_ZERO, V_X, V_N, V_C, V_OP, V_S };  // V_S: state entry idx (vector walks)

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

bool binary_op(uint32_t op) {
  return op == D_ADD || op == D_SUB || op == D_MUL || op == D_ZSEL || op == D_MULNZ;
}

// A code list over a value graph, with its entry routing.
struct CodeList {
  std::vector<Ins> code;
  std::vector<int> entry_direct;  // value -> entry its producer writes (-1: none)
};

// Instructions for the values in `order` (RED inserted when `with_red`;
// with `reduce_out` also before an output that could pass 2^31), then
// trailing moves for outputs whose value cannot be written by its producer:
// inputs, values shared by several outputs or used later. Output o (an entry
// of M, or a state entry of a vector walk) is `out[o]`, emitted when
// `out_gen[o]`.
CodeList build_code(std::vector<Val>& vals, const std::vector<int>& order, std::vector<int> out,
                    const std::vector<char>& out_gen, bool with_red, bool reduce_out) {
  CodeList cl;
  std::vector<double> bound(vals.size(), 0.0);
  auto src_bound = [&](int v) -> double {
    switch (vals[v].kind) {
      case V_ZERO: return 0.0;
      case V_X: return kMulBound;
      case V_N: return kMulBound;
      case V_S: return kMulBound;
      case V_C: return kMulBound / 2;
      default: return bound[v];
    }
  };
  for (int v : order) {
    const Val x = vals[v];
    const bool unary = !binary_op(x.op);
    double bd;
    switch (x.op) {
      case D_ADD:
      case D_SUB: bd = src_bound(x.a) + src_bound(x.b); break;
      case D_NEG:
      case D_NZ1:
      case D_MOV: bd = src_bound(x.a); break;
      case D_ZSEL: bd = src_bound(x.a); break;
      default: bd = kMulBound; break;
    }
    if (with_red && bd > kBoundLimit) {
      // the sum goes to the accumulator only; RED(A) produces value v
      const int tmp = (int)vals.size();
      vals.push_back(x);
      bound.push_back(bd);
      cl.code.push_back(Ins{x.op, x.a, unary ? -1 : x.b, tmp});
      Val r{V_OP};
      r.op = D_RED;
      r.a = tmp;
      vals[v] = r;
      bound[v] = kMulBound;
      cl.code.push_back(Ins{D_RED, tmp, -1, v});
    } else {
      bound[v] = bd;
      cl.code.push_back(Ins{x.op, x.a, unary ? -1 : x.b, v});
    }
  }
  const int nout = (int)out.size();
  if (with_red && reduce_out) {
    // a vector walk's state must come back reduced: it is next step's input
    for (int o = 0; o < nout; ++o) {
      const int v = out[o];
      if (!out_gen[o] || vals[v].kind != V_OP || bound[v] <= kMulBound) continue;
      Val r{V_OP};
      r.op = D_RED;
      r.a = v;
      vals.push_back(r);
      bound.push_back(kMulBound);
      const int nv2 = (int)vals.size() - 1;
      cl.code.push_back(Ins{D_RED, v, -1, nv2});
      for (int o2 = o; o2 < nout; ++o2)
        if (out[o2] == v) out[o2] = nv2;
    }
  }
  const int nvals = (int)vals.size();
  std::vector<int> n_entry_of(nvals, 0), uses(nvals, 0);
  for (int e = 0; e < nout; ++e)
    if (out_gen[e]) ++n_entry_of[out[e]];
  for (const Ins& in : cl.code) {
    if (in.a >= 0) ++uses[in.a];
    if (in.b >= 0) ++uses[in.b];
  }
  cl.entry_direct.assign(nvals, -1);
  for (int e = 0; e < nout; ++e) {
    if (!out_gen[e]) continue;
    const int v = out[e];
    if (vals[v].kind == V_OP && n_entry_of[v] == 1 && uses[v] == 0 && cl.entry_direct[v] < 0) {
      cl.entry_direct[v] = e;
    } else {
      cl.code.push_back(Ins{D_MOV, v, -1, -1, e});
      ++uses[v];
    }
  }
  return cl;
}

// Constant slots: constants of the residue code by use count (ties: first
// use), then those only the FP64 code reads.
void assign_const_slots(const CodeList& rcl, const std::vector<Val>& rvals, const CodeList& fcl,
                        const std::vector<Val>& fvals, int n_const_all, std::vector<int>& cslot,
                        std::vector<int>& slots) {
  std::vector<int> uses(n_const_all, 0), first(n_const_all, -1);
  int pos = 0;
  for (const Ins& in : rcl.code) {
    for (int u : {in.a, in.b}) {
      if (u < 0 || rvals[u].kind != V_C) continue;
      const int ci = rvals[u].idx;
      ++uses[ci];
      if (first[ci] < 0) first[ci] = pos++;
    }
  }
  std::vector<int> order;
  for (int ci = 0; ci < n_const_all; ++ci)
    if (first[ci] >= 0) order.push_back(ci);
  std::sort(order.begin(), order.end(),
            [&](int x, int y) { return uses[x] != uses[y] ? uses[x] > uses[y] : first[x] < first[y]; });
  for (const Ins& in : fcl.code) {
    for (int u : {in.a, in.b}) {
      if (u < 0 || fvals[u].kind != V_C) continue;
      const int ci = fvals[u].idx;
      if (first[ci] < 0) {
        first[ci] = pos++;
        order.push_back(ci);
      }
    }
  }
  cslot.assign(n_const_all, -1);
  slots = order;
  for (size_t i = 0; i < order.size(); ++i) cslot[order[i]] = (int)i;
}

// Residue graph: the values, with every needed INV whose argument does not
// depend on another INV replaced by its share of one inversion (Montgomery's
// trick, zero arguments counted as 1): c_0 = NZ1(a_0), c_j = MULNZ(c_{j-1},
// a_j), t = INV(c_{k-1}); going down, w_j = t c_{j-1} and t <- MULNZ(t, a_j);
// INV(a_j) = ZSEL(w_j, a_j). prio: the position in the reference's order
// (new values: just after their last argument).
struct ResGraph {
  std::vector<Val> vals;
  std::vector<double> prio;
  int n_batched = 0;
};

ResGraph residue_graph(const std::vector<Val>& vals, const std::vector<char>& needed) {
  ResGraph g;
  g.vals = vals;
  const int nv = (int)vals.size();
  g.prio.resize(nv);
  for (int v = 0; v < nv; ++v) g.prio[v] = v;
  std::vector<char> dep_inv(nv, 0);  // value depends on an INV result
  for (int v = 0; v < nv; ++v) {
    const Val& x = vals[v];
    if (x.kind != V_OP) continue;
    dep_inv[v] = x.op == D_INV || dep_inv[x.a] || (binary_op(x.op) && dep_inv[x.b]);
  }
  std::vector<int> inv_vals;
  for (int v = 0; v < nv; ++v)
    if (needed[v] && vals[v].kind == V_OP && vals[v].op == D_INV && !dep_inv[vals[v].a]) inv_vals.push_back(v);
  const int k = (int)inv_vals.size();
  if (k < 2) return g;
  auto add = [&](uint32_t op, int a, int b, double pr) {
    Val x{V_OP};
    x.op = op;
    x.a = a;
    x.b = b;
    g.vals.push_back(x);
    g.prio.push_back(pr);
    return (int)g.vals.size() - 1;
  };
  std::vector<int> arg(k), c(k);
  double last = 0;
  for (int j = 0; j < k; ++j) {
    arg[j] = vals[inv_vals[j]].a;
    const double pr = g.prio[arg[j]] + 0.5;
    last = std::max(last, pr);
    c[j] = j == 0 ? add(D_NZ1, arg[0], 0, pr) : add(D_MULNZ, c[j - 1], arg[j], pr);
  }
  int t = add(D_INV, c[k - 1], 0, last);
  std::vector<int> w(k);
  for (int j = k - 1; j >= 1; --j) {
    w[j] = add(D_MUL, t, c[j - 1], last);
    t = add(D_MULNZ, t, arg[j], last);
  }
  w[0] = t;
  for (int j = 0; j < k; ++j) {
    Val r{V_OP};
    r.op = D_ZSEL;
    r.a = w[j];
    r.b = arg[j];
    g.vals[inv_vals[j]] = r;
  }
  g.n_batched = k;
  return g;
}

constexpr int kPolicies = 4;

// Evaluation order of the values reachable from `roots` (operands first):
// 0: depth-first from the entries in order; 1: depth-first, the operand
// needing more registers first (Sethi-Ullman labels); 2: depth-first from the
// entries in reverse order; 3: list scheduling by prio.
std::vector<int> schedule(const std::vector<Val>& vals, const std::vector<double>& prio,
                          const std::vector<int>& roots, int policy) {
  const int n = (int)vals.size();
  auto dfs = [&](const std::vector<int>& rts, const std::vector<int>* label) {
    std::vector<int> order;
    std::vector<char> state(n, 0);  // 0 new, 1 open, 2 done
    std::vector<std::pair<int, int>> st;
    for (int r : rts) {
      st.push_back({r, 0});
      while (!st.empty()) {
        auto& top = st.back();
        const int v = top.first;
        if (vals[v].kind != V_OP || state[v] == 2) {
          st.pop_back();
          continue;
        }
        state[v] = 1;
        int first = vals[v].a, second = binary_op(vals[v].op) ? vals[v].b : -1;
        if (label && second >= 0 && (*label)[second] > (*label)[first]) std::swap(first, second);
        if (top.second == 0) {
          top.second = 1;
          st.push_back({first, 0});
        } else if (top.second == 1) {
          top.second = 2;
          if (second >= 0) st.push_back({second, 0});
        } else {
          state[v] = 2;
          order.push_back(v);
          st.pop_back();
        }
      }
    }
    return order;
  };
  std::vector<int> base = dfs(roots, nullptr);
  if (policy == 0) return base;
  if (policy == 2) {
    std::vector<int> rev(roots.rbegin(), roots.rend());
    return dfs(rev, nullptr);
  }
  if (policy == 1) {
    std::vector<int> label(n, 0);
    for (int v : base) {
      const int la = label[vals[v].a], lb = binary_op(vals[v].op) ? label[vals[v].b] : 0;
      label[v] = la == lb ? la + 1 : std::max(la, lb);
    }
    return dfs(roots, &label);
  }
  // list scheduling over the reachable values
  std::vector<int> pending(n, 0);
  std::vector<std::vector<int>> users(n);
  for (int v : base) {
    for (int u : {vals[v].a, binary_op(vals[v].op) ? vals[v].b : -1}) {
      if (u < 0 || vals[u].kind != V_OP) continue;
      ++pending[v];
      users[u].push_back(v);
    }
  }
  auto cmp = [&](int x, int y) { return prio[x] != prio[y] ? prio[x] > prio[y] : x > y; };
  std::priority_queue<int, std::vector<int>, decltype(cmp)> ready(cmp);
  for (int v : base)
    if (pending[v] == 0) ready.push(v);
  std::vector<int> order;
  while (!ready.empty()) {
    const int v = ready.top();
    ready.pop();
    order.push_back(v);
    for (int u : users[v])
      if (--pending[u] == 0) ready.push(u);
  }
  return order;
}

// Residue code (and, with `fp`, the FP64 code of the M_f pass): lowering to
// the bank instruction set with register allocation over the 16 temps of
// bank B. Axes, the step and the first
// 15 - dim constant slots sit in B's input slots and are read in place;
// further constants are loaded from LDS (LDC), a vector walk's state entries
// too (LDV), and zero is made by LIN. A value brought into a temp stays there
// while it has uses (an LDS input is simply dropped when its temp is
// needed). Every result goes to a temp; a matrix entry (or new state entry)
// then leaves through STE. When all temps are taken, the value used furthest
// in the future leaves; a computed one is spilled to an LDS row (STM) first
// and reloaded (LDM) before its next use.
std::string encode_res(const CodeList& cl, const std::vector<Val>& vals, const std::vector<int>& cslot, int dim,
                       std::vector<uint32_t>& words, int& nrows, bool fp) {
  const std::vector<Ins>& code = cl.code;
  const int nc = (int)code.size(), nvals = (int)vals.size();
  const int n_cbank = kResStep - dim;  // constant slots held in the bank
  std::vector<std::vector<int>> uses(nvals);
  for (int i = 0; i < nc; ++i)
    for (int u : {code[i].a, code[i].b})
      if (u >= 0) uses[u].push_back(i);
  std::vector<size_t> next(nvals, 0);
  auto next_use = [&](int v) { return next[v] < uses[v].size() ? uses[v][next[v]] : INT32_MAX; };
  constexpr int kTemp = -2;
  int held[kResTemps];
  for (int& h : held) h = -1;
  bool pinned[kResTemps];
  for (bool& p : pinned) p = false;
  std::vector<int> reg(nvals, -1), row(nvals, -1);
  int n_rows = 0;
  std::vector<int> free_rows;  // rows of spilled values that are dead
  words.clear();
  const uint32_t any = kResCBase + kResStep;  // an operand that is always finite (the step)
  auto free_reg = [&]() -> int {
    for (int r = 0; r < kResTemps; ++r)
      if (held[r] == -1) return r;
    int best = -1, far = -1;
    for (int r = 0; r < kResTemps; ++r) {
      if (pinned[r] || held[r] < 0) continue;
      const int nu = next_use(held[r]);
      if (nu > far) {
        far = nu;
        best = r;
      }
    }
    if (best < 0) return -1;
    const int v = held[best];
    if (vals[v].kind == V_OP && row[v] < 0) {
      if (!free_rows.empty()) {
        row[v] = free_rows.back();
        free_rows.pop_back();
      } else {
        row[v] = n_rows++;
      }
      words.push_back(enc_res(RO_STM, (uint32_t)best, any, (uint32_t)best, (uint32_t)row[v]));
    }
    reg[v] = -1;
    held[best] = -1;
    return best;
  };
  // input held in the bank's input slots: its bank index, else -1
  auto in_bank = [&](int u) -> int {
    const Val& x = vals[u];
    if (x.kind == V_X) return kResCBase + kResStep - 1 - x.idx;
    if (x.kind == V_N) return kResCBase + kResStep;
    if (x.kind == V_C && cslot[x.idx] < n_cbank) return kResCBase + cslot[x.idx];
    return -1;
  };
  // operand u: the bank index holding it, after bringing it into a temp
  // (reload, LDS input or zero) when it is not in the bank; the temp is pinned
  auto operand = [&](int u) -> int {
    if (reg[u] >= 0) {
      pinned[reg[u]] = true;
      return reg[u];
    }
    const int b = in_bank(u);
    if (b >= 0) return b;
    const Val& x = vals[u];
    if (x.kind == V_OP && row[u] < 0) return -1;  // not computed yet: the order is not topological
    const int r = free_reg();
    if (r < 0) return -1;
    if (x.kind == V_OP) words.push_back(enc_res(RO_LDM, any, any, (uint32_t)r, (uint32_t)row[u]));
    else if (x.kind == V_S) words.push_back(enc_res(RO_LDV, any, any, (uint32_t)r, (uint32_t)x.idx));
    else if (x.kind == V_C) words.push_back(enc_res(RO_LDC, any, any, (uint32_t)r, (uint32_t)(cslot[x.idx] - n_cbank)));
    else if (fp) words.push_back(enc_res(FO_ZERO, any, any, (uint32_t)r, 0));
    else words.push_back(enc_res(RO_LIN, any, any, (uint32_t)r, lin_signs(0, 0)));  // zero
    reg[u] = r;
    held[r] = u;
    pinned[r] = true;
    return r;
  };
  for (int i = 0; i < nc; ++i) {
    const Ins& in = code[i];
    uint32_t op = RO_MUL, aux = 0;
    int x = 0, y = 0;
    const uint32_t kind = in.op;
    int e = -1;  // entry row the result leaves through
    if (in.entry >= 0) e = in.entry;
    else if (in.def >= 0 && cl.entry_direct[in.def] >= 0) e = cl.entry_direct[in.def];
    if (kind == D_MOV) {
      // an output that is an input or a value also used elsewhere: stored from where it is
      x = operand(in.a);
      if (x < 0) return "internal: no free register";
      if (e < 0) return "internal: a move that is no output";
      op = RO_STE;
      aux = (uint32_t)e;
      y = (int)any;
    } else if (kind == D_MUL || kind == D_ADD || kind == D_SUB) {
      x = operand(in.a);
      if (x < 0) return "internal: no free register";
      y = operand(in.b);
      if (y < 0) return "internal: no free register";
      if (kind == D_ADD) op = RO_ADD;
      else if (kind == D_SUB) {
        op = RO_LIN;
        aux = lin_signs(1, -1);
      }
    } else if (fp && (kind == D_CUB || kind == D_INV || kind == D_NEG || kind == D_RED)) {
      // single FP64 operations: the correctly rounded cube, 1/x with 0 for
      // |x| <= 1e-300, the sign flip, and RED, which is no operation in FP64
      x = operand(in.a);
      if (x < 0) return "internal: no free register";
      y = x;
      op = kind == D_CUB ? FO_CUB : (kind == D_INV ? FO_INV : (kind == D_NEG ? FO_NEG : FO_MOV));
    } else if (kind == D_SQR || kind == D_CUB) {
      x = operand(in.a);
      if (x < 0) return "internal: no free register";
      y = x;
      if (kind == D_CUB) {
        const int t = free_reg();
        if (t < 0) return "internal: no free register";
        words.push_back(enc_res(RO_MUL, (uint32_t)x, (uint32_t)x, (uint32_t)t, 0));  // t = x^2
        held[t] = kTemp;
        pinned[t] = true;
        x = t;  // then x^2 * x
      }
    } else if (kind == D_NEG) {
      x = operand(in.a);
      if (x < 0) return "internal: no free register";
      y = x;
      op = RO_LIN;
      aux = lin_signs(-1, 0);
    } else if (kind == D_INV || kind == D_RED || kind == D_NZ1) {
      x = operand(in.a);
      if (x < 0) return "internal: no free register";
      y = (int)any;
      op = kind == D_INV ? RO_INV : (kind == D_RED ? RO_RED : RO_NZ1);
    } else if (fp) {
      return "internal: residue-only operation in the FP64 code";
    } else if (kind == D_ZSEL || kind == D_MULNZ) {
      x = operand(in.a);
      if (x < 0) return "internal: no free register";
      y = operand(in.b);
      if (y < 0) return "internal: no free register";
      op = kind == D_ZSEL ? RO_ZSEL : RO_MULNZ;
    } else {
      return "internal: unknown operation";
    }
    // operands read: advance next uses; free temps and values used for the last time
    for (int u : {in.a, in.b})
      if (u >= 0 && next[u] < uses[u].size() && uses[u][next[u]] == i) ++next[u];
    for (int u : {in.a, in.b}) {
      if (u < 0 || next_use(u) != INT32_MAX) continue;
      if (reg[u] >= 0) {
        held[reg[u]] = -1;
        reg[u] = -1;
      }
      if (row[u] >= 0) {
        free_rows.push_back(row[u]);
        row[u] = -1;
      }
    }
    for (int r = 0; r < kResTemps; ++r)
      if (held[r] == kTemp) held[r] = -1;
    for (bool& p : pinned) p = false;
    if (op == RO_STE) {
      words.push_back(enc_res(RO_STE, (uint32_t)x, (uint32_t)y, (uint32_t)x, aux));
      continue;
    }
    // destination temp; an entry also leaves through STE
    const int d = free_reg();
    if (d < 0) return "internal: no free register";
    if (in.def >= 0 && !uses[in.def].empty()) {
      held[d] = in.def;
      reg[in.def] = d;
    }
    words.push_back(enc_res(op, (uint32_t)x, (uint32_t)y, (uint32_t)d, aux));
    if (e >= 0) words.push_back(enc_res(RO_STE, (uint32_t)d, any, (uint32_t)d, (uint32_t)e));
  }
  if (n_rows > (int)kMaxRow) return "too many spilled values";
  nrows = n_rows;
  return std::string();
}

}  // namespace

std::string pack_program(const uint32_t* ops, int n_ops, int m, int dim, int n_const,
                         const std::vector<char>& const_is_one, const std::vector<char>& const_is_zero,
                         PackedProgram* out, int mode) {
  char msg[160];
  const bool vec = mode == kPackVector;
  const int mmax = vec ? kMaxVecM : kMaxM;
  if (mode != kPackMatrix && mode != kPackVector) return "unknown packing mode";
  if (m < 2 || m > mmax) {
    std::snprintf(msg, sizeof msg, "%s order m=%d outside [2, %d]", vec ? "state vector" : "matrix", m, mmax);
    return msg;
  }
  if (dim < 0 || dim > kMaxDim) {
    std::snprintf(msg, sizeof msg, "dim=%d outside [0, %d]", dim, kMaxDim);
    return msg;
  }
  std::vector<Val> vals;
  vals.push_back(Val{V_ZERO});
  // value numbering: an operation on the same operand values is the same
  // value (exact in residues and in FP64, where the operation is
  // deterministic and ADD / MUL are commutative bit for bit)
  std::map<std::array<int64_t, 4>, int> seen;
  auto intern = [&](Val x) -> int {
    if (x.kind == V_OP && (x.op == D_ADD || x.op == D_MUL) && x.a > x.b) std::swap(x.a, x.b);
    const std::array<int64_t, 4> key = {x.kind, x.kind == V_OP ? (int64_t)x.op : x.idx, x.a, x.b};
    auto it = seen.find(key);
    if (it != seen.end()) return it->second;
    vals.push_back(x);
    seen.emplace(key, (int)vals.size() - 1);
    return (int)vals.size() - 1;
  };
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
        reg[a1] = intern(Val{V_X, (int)a0});
        break;
      case R_LOAD_C:
        if ((int)a0 >= n_const || !chk_reg(a1)) goto bad_operand;
        reg[a1] = intern(Val{V_C, (int)a0});
        break;
      case R_LOAD_N:
        if (!chk_reg(a0)) goto bad_operand;
        reg[a0] = intern(Val{V_N});
        break;
      case R_ADD:
      case R_SUB:
      case R_MUL: {
        if (!chk_reg(a0) || !chk_reg(a1) || !chk_reg(a2)) goto bad_operand;
        Val v{V_OP};
        v.op = op == R_ADD ? D_ADD : (op == R_SUB ? D_SUB : D_MUL);
        v.a = reg[a0];
        v.b = reg[a1];
        reg[a2] = intern(v);
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
        reg[a1] = intern(v);
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

  // ---- constant folding (values are in topological order)
  int n_const_all = n_const;
  for (size_t v = 1; v < vals.size(); ++v) {
    Val& x = vals[v];
    if (x.kind != V_OP) continue;
    const bool unary = !binary_op(x.op);
    const bool foldable_op = x.op == D_ADD || x.op == D_SUB || x.op == D_MUL || x.op == D_NEG || x.op == D_SQR ||
                             x.op == D_INV;
    if (!foldable_op || vals[x.a].kind != V_C || (!unary && vals[x.b].kind != V_C)) continue;
    pp.derived.push_back(DerivedConst{x.op, vals[x.a].idx, unary ? -1 : vals[x.b].idx});
    const Val c{V_C, n_const_all++};
    x = c;
  }
  if (n_const_all > kMaxConst) return "too many constants after folding";
  std::vector<char> is_one(const_is_one), is_zero(const_is_zero);
  is_one.resize(n_const_all, 0);
  is_zero.resize(n_const_all, 0);

  // ---- entry kinds
  auto is_const_val = [&](int v, const std::vector<char>& tab) { return vals[v].kind == V_C && tab[vals[v].idx]; };
  std::vector<uint32_t> ekind(m * m);
  for (int e = 0; e < m * m; ++e) {
    const int v = entry[e];
    ekind[e] = E_GEN;
    if (vals[v].kind == V_ZERO || is_const_val(v, is_zero)) ekind[e] = E_ZERO;
    else if (is_const_val(v, is_one)) ekind[e] = E_ONE;
    if (!vec) pp.kinds |= ekind[e] << (2 * e);
  }
  // ---- outputs: the general entries of M, or for a vector walk the new
  // state v'_i = sum_j M_ij v_j over the nonzero entries in column order
  std::vector<int> outv;
  std::vector<char> out_gen;
  if (!vec) {
    outv = entry;
    for (int e = 0; e < m * m; ++e) out_gen.push_back(ekind[e] == E_GEN);
  } else {
    std::vector<int> state(m);
    for (int j = 0; j < m; ++j) {
      vals.push_back(Val{V_S, j});
      state[j] = (int)vals.size() - 1;
    }
    for (int i = 0; i < m; ++i) {
      int acc = 0;  // V_ZERO: a row without nonzero entries
      for (int j = 0; j < m; ++j) {
        const uint32_t kd = ekind[i * m + j];
        if (kd == E_ZERO) continue;
        int term = state[j];
        if (kd == E_GEN) {
          Val t{V_OP};
          t.op = D_MUL;
          t.a = entry[i * m + j];
          t.b = state[j];
          vals.push_back(t);
          term = (int)vals.size() - 1;
        }
        if (acc == 0) {
          acc = term;
        } else {
          Val t{V_OP};
          t.op = D_ADD;
          t.a = acc;
          t.b = term;
          vals.push_back(t);
          acc = (int)vals.size() - 1;
        }
      }
      outv.push_back(acc);
      out_gen.push_back(1);
    }
  }
  const int nout = (int)outv.size();

  // ---- liveness
  const int nv = (int)vals.size();
  std::vector<char> needed(nv, 0);
  std::vector<int> stack;
  for (int o = 0; o < nout; ++o)
    if (out_gen[o]) stack.push_back(outv[o]);
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
  // polynomial degree of every needed value in the step (-1: not a
  // polynomial); forward differences serve matrix walks only
  if (!vec) {
    std::vector<int> deg(nv, 0);
    for (int v = 0; v < nv; ++v) {
      const Val& x = vals[v];
      switch (x.kind) {
        case V_ZERO: deg[v] = 0; break;
        case V_C: deg[v] = 0; break;
        case V_X: deg[v] = 1; break;
        case V_N: deg[v] = 1; break;
        case V_S: deg[v] = -1; break;
        default: {
          const int da = deg[x.a], db = deg[x.b];
          if (x.op == D_INV || da < 0 || (db < 0 && binary_op(x.op))) {
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
      if (ekind[e] != E_GEN) continue;
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
    for (int j = 0; j < m; ++j) {
      int nz = 0, ones = 0;
      for (int k = 0; k < m; ++k) {
        nz += ekind[k * m + j] != E_ZERO;
        ones += ekind[k * m + j] == E_ONE;
      }
      if (ones > 0 && nz >= 2) pp.colred |= 1u << j;
    }
  }

  std::vector<int> roots;
  for (int o = 0; o < nout; ++o)
    if (out_gen[o]) roots.push_back(outv[o]);
  std::vector<int> dummy_cslot(n_const_all, 0);

  // ---- FP64 code: the same operations as the reference (one reciprocal per
  // INV), in the candidate order giving the shortest code, then the fewest
  // LDS spill rows
  CodeList fcl;
  std::vector<Val> fvals;
  {
    std::vector<double> prio(nv);
    for (int v = 0; v < nv; ++v) prio[v] = v;
    int best = -1;
    size_t best_flen = 0;
    for (int policy = 0; policy < kPolicies; ++policy) {
      std::vector<Val> cv = vals;
      CodeList cl = build_code(cv, schedule(cv, prio, roots, policy), outv, out_gen, false, false);
      std::vector<uint32_t> scratch_words;
      int ns = 0;
      const std::string err = encode_res(cl, cv, dummy_cslot, dim, scratch_words, ns, true);
      if (!err.empty()) return err;
      if (best < 0 || scratch_words.size() < best_flen || (scratch_words.size() == best_flen && ns < best)) {
        best = ns;
        best_flen = scratch_words.size();
        fcl = std::move(cl);
        fvals = std::move(cv);
      }
    }
  }

  // ---- residue code: shared inversion, then the evaluation order (of a few
  // candidates) giving the shortest code, then the fewest LDS spill rows
  CodeList rcl;
  std::vector<Val> rvals;
  int best_slots = -1;
  size_t best_len = 0;
  {
    const ResGraph g = residue_graph(vals, needed);
    for (int policy = 0; policy < kPolicies; ++policy) {
      std::vector<Val> cv = g.vals;
      const std::vector<int> order = schedule(cv, g.prio, roots, policy);
      CodeList cl = build_code(cv, order, outv, out_gen, true, vec);
      std::vector<uint32_t> scratch_words;
      int ns = 0;
      const std::string err = encode_res(cl, cv, dummy_cslot, dim, scratch_words, ns, false);
      if (!err.empty()) return err;
      if (best_slots < 0 || scratch_words.size() < best_len || (scratch_words.size() == best_len && ns < best_slots)) {
        best_slots = ns;
        best_len = scratch_words.size();
        rcl = std::move(cl);
        rvals = std::move(cv);
        pp.n_inv_batched = g.n_batched;
      }
    }
  }

  // ---- constant slots (the residue code's most used first: slots 0..15 sit
  // in VGPRs, the rest in LDS), then routing and allocation of both codes
  std::vector<int> cslot;
  assign_const_slots(rcl, rvals, fcl, fvals, n_const_all, cslot, pp.const_slots);
  pp.const_lds = pp.const_slots.size() > (size_t)(kResStep - dim) ? 1 : 0;
  if (pp.const_slots.size() > kMaxRow + (size_t)(kResStep - dim)) return "too many constants";
  std::string err = encode_res(rcl, rvals, cslot, dim, pp.code, pp.nslots, false);
  if (!err.empty()) return err;
  while (pp.code.size() % kResGroup) pp.code.push_back(kResNop);
  err = encode_res(fcl, fvals, cslot, dim, pp.fcode, pp.fnslots, true);
  if (!err.empty()) return err;
  while (pp.fcode.size() % kResGroup) pp.fcode.push_back(kResNop);
  *out = std::move(pp);
  return std::string();
}

}  // namespace drx

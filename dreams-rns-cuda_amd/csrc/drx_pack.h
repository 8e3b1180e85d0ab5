// Host-side program packer: reference bytecode -> banked device code.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace drx {

// A constant computed from other constants at upload time (constant
// subexpression folding): value = op(a, b) with a, b constant indices
// (b = -1 for unary ops); its index is n_const + its position in `derived`.
struct DerivedConst {
  uint32_t op;  // D_ADD, D_SUB, D_MUL, D_NEG, D_SQR or D_INV
  int a, b;
};

struct PackedProgram {
  int m = 0, dim = 0;
  // Residue code: two words per device instruction. Step-dependent INVs
  // whose arguments do not depend on another INV share one inversion
  // (Montgomery's trick with zero arguments masked, so INV(0) stays 0).
  std::vector<uint32_t> code;
  // FP64 code (the M_f pass): the reference's operations, one reciprocal per
  // INV, so the shadow's bits follow the reference.
  std::vector<uint32_t> fcode;
  std::vector<int> const_slots;      // device constant slot -> constant index (>= n_const: derived)
  std::vector<DerivedConst> derived;
  uint32_t kinds = 0;                // EntryKind per entry (2 bits)
  uint32_t colred = 0;               // columns reduced after P*M
  int const_lds = 0;                 // constant slots beyond the 16 of the VGPR bank live in LDS
  int nslots = 0;                    // LDS temp slots of the residue code
  int fnslots = 0;                   // LDS temp slots of the FP64 code
  int n_src_ops = 0;                 // reference instructions before END
  int n_inv_batched = 0;             // INVs served by the shared inversion
  // Degree in the step of every general entry when the program has no INV
  // on any path to an entry (then every entry is a polynomial in the step):
  // 4 bits per entry; fd_degree = max over entries; fd_ok = 0 otherwise.
  int fd_ok = 0;
  int fd_degree = 0;
  uint64_t fd_deg = 0;
  // per-instruction statistics of the reference program (for the roofline)
  int n_mul = 0, n_sqr = 0, n_cub = 0, n_inv = 0, n_add = 0, n_sub = 0, n_neg = 0;
};

constexpr int kPackMatrix = 0;  // outputs: the general entries of M(s)
constexpr int kPackVector = 1;  // outputs: the new state v' = M(s) v of a vector walk
constexpr int kMaxVecM = 16;    // state entries of a vector walk

// `const_is_one[i]` / `const_is_zero[i]`: constant i equals 1 / 0 exactly.
// Returns an empty string on success, else an error message.
std::string pack_program(const uint32_t* ops, int n_ops, int m, int dim, int n_const,
                         const std::vector<char>& const_is_one, const std::vector<char>& const_is_zero,
                         PackedProgram* out, int mode = kPackMatrix);

}  // namespace drx

// Host-side program packer: reference bytecode -> banked device code.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace drx {

struct PackedProgram {
  int m = 0, dim = 0;
  std::vector<uint32_t> code;        // two words per device instruction
  std::vector<int> const_slots;      // device constant slot -> program constant index
  uint32_t kinds = 0;                // EntryKind per entry (2 bits)
  uint32_t colred = 0;               // columns reduced after P*M
  int const_lds = 0;                 // constants in LDS instead of the VGPR bank
  int nslots = 0;                    // LDS temp slots
  int n_src_ops = 0;                 // reference instructions before END
  // Degree in the step of every general entry when the program has no INV
  // on any path to an entry (then every entry is a polynomial in the step):
  // 4 bits per entry; fd_degree = max over entries; fd_ok = 0 otherwise.
  int fd_ok = 0;
  int fd_degree = 0;
  uint64_t fd_deg = 0;
  // per-instruction statistics of the reference program (for the roofline)
  int n_mul = 0, n_sqr = 0, n_cub = 0, n_inv = 0, n_add = 0, n_sub = 0, n_neg = 0;
};

// `const_is_one[i]` / `const_is_zero[i]`: constant i equals 1 / 0 exactly.
// Returns an empty string on success, else an error message.
std::string pack_program(const uint32_t* ops, int n_ops, int m, int dim, int n_const,
                         const std::vector<char>& const_is_one, const std::vector<char>& const_is_zero,
                         PackedProgram* out);

}  // namespace drx

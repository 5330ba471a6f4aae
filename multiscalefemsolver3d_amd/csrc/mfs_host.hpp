// Host-side setup shared by the C-ABI (C++17, no device code).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mfs {

// Sparse pattern in CSSD row order (down ++ diag ++ up, sorted): a sorted CSR
// whose every row holds its diagonal.
struct Pattern {
  int64_t n = 0;
  std::vector<int32_t> rowptr;  // n+1
  std::vector<int32_t> col;     // nnz
  int64_t nnz() const { return rowptr.empty() ? 0 : rowptr.back(); }
};

// Build the node-adjacency pattern of a tet mesh (nodes already in DOF order).
void build_pattern(int64_t n_rows, int64_t n_tets, const int32_t* tets, Pattern& P);

// Gather lists for deterministic assembly: for every pattern slot, the
// contributions (tet << 4 | i << 2 | j) in ascending tet order, where local
// node i of the tet maps to the slot's row and local node j to its column.
struct GatherLists {
  std::vector<int64_t> ptr;       // nnz+1
  std::vector<uint32_t> entry;    // 16 per tet in total
};
void build_gather_lists(const Pattern& P, int64_t n_tets, const int32_t* tets, GatherLists& G);
// node -> tets containing it (ascending tet order per node)
void node_tets(int64_t nn, int64_t nt, const int32_t* tets, std::vector<int64_t>& ptr, std::vector<int32_t>& list);
// per slice of C consecutive nodes, the distinct tets touching them (ascending)
void build_slice_tets(int64_t nn, int64_t nt, const int32_t* tets, int C, std::vector<int64_t>& sptr,
                      std::vector<int32_t>& stet);

// Mechanical renumbering (Dirichlet first, then first appearance by material).
int64_t renumber_mech(int64_t nn, int64_t nt, const int32_t* tets, const int32_t* mat, int64_t nfix,
                      const int32_t* fixed, int32_t* map);

// SELL-C layout of a (block) pattern: C rows per slice, every slice padded to
// its longest row; entry j of row r lives at slice_base[r/C] + j*C + r%C.
struct SellLayout {
  int C = 64;
  int64_t n_slices = 0;
  std::vector<int64_t> slice_base;  // n_slices+1, in entries (not values)
  std::vector<int32_t> slice_width; // n_slices
  std::vector<int32_t> col;         // padded columns (own row for padding)
  std::vector<int64_t> pos;         // CSR slot -> SELL entry index
};
void build_sell(const Pattern& P, int C, SellLayout& S);

void set_error(const std::string& msg);

}  // namespace mfs

// Internal device-side definitions shared by the HIP sources of libmfs.so.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mfs.h"
#include "mfs_host.hpp"

#define MFS_WAVE 64

// HIP call check: records the error text and returns MFS_ERR_HIP (or OOM).
#define MFS_HIP(call)                                                                         \
  do {                                                                                        \
    hipError_t _e = (call);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      mfs::set_error(std::string(#call) + ": " + hipGetErrorString(_e));                     \
      return _e == hipErrorOutOfMemory ? MFS_ERR_OOM : MFS_ERR_HIP;                          \
    }                                                                                         \
  } while (0)

#define MFS_CHECK_ARG(cond, msg)  \
  do {                            \
    if (!(cond)) {                \
      mfs::set_error(msg);        \
      return MFS_ERR_ARG;         \
    }                             \
  } while (0)

#define MFS_TRY(expr)          \
  do {                         \
    int _s = (expr);           \
    if (_s != MFS_OK) return _s; \
  } while (0)

// Contexts and matrices are reference counted: every object created on a
// context holds a reference to it, and a solver holds its matrix, so the
// destroy calls may come in any order (as they do from garbage collectors).
struct mfs_ctx_s {
  std::atomic<int> refs{1};
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  char name[256] = {0};
  // small pinned host buffer for status read-backs
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  // device address of this device's copy of the library's fixed scratch
  // array (mfs_solve_noalloc), resolved once per device at context creation
  void* noalloc_scratch = nullptr;
  size_t noalloc_scratch_bytes = 0;
};

struct mfs_vec_s {
  mfs_ctx ctx = nullptr;
  int64_t n = 0;
  int prec = MFS_F64;
  void* d = nullptr;
};

// Sparse matrix: pattern in CSSD order (host copy for export), device values
// in a SELL-64 layout (thread per block row, column-major inside a slice).
struct mfs_mat_s {
  std::atomic<int> refs{1};
  mfs_ctx ctx = nullptr;
  int bs = 1;            // 1 = CSR, 3 = BSR3
  int64_t nb = 0;        // block rows
  int64_t nnzb = 0;      // block nonzeros
  mfs::Pattern pat;      // block pattern
  mfs::SellLayout sell;  // host copy of the layout
  int64_t sell_entries = 0;
  // device
  int64_t* d_slice_base = nullptr;
  int32_t* d_slice_width = nullptr;
  int32_t* d_col = nullptr;     // sell_entries
  int16_t* d_col16 = nullptr;   // sell_entries: column - row, when all fit (else null)
  int64_t* d_pos = nullptr;     // nnzb: CSR slot -> SELL entry
  int64_t* d_diag_pos = nullptr;// nb: SELL entry of the diagonal block
  double* d_val64 = nullptr;    // bs*bs*sell_entries
  float* d_val32 = nullptr;
  double* d_dinv64 = nullptr;   // bs*nb inverse diagonal (Jacobi)
  float* d_dinv32 = nullptr;
  bool val32_valid = false;
  int xcd_map = 1;  // SELL kernels: contiguous slice ranges per XCD (xcd_block)
  int64_t n() const { return bs * nb; }
  int64_t nnz() const { return (int64_t)bs * bs * nnzb; }
};

namespace mfs {

// reference counting (see mfs_ctx_s)
void ctx_retain(mfs_ctx c);
void ctx_release(mfs_ctx c);
void mat_retain(mfs_mat A);
void mat_release(mfs_mat A);

// allocate + track
int dmalloc(void** p, size_t bytes);
void dfree(void* p);

// matrix internals used by assembly and MsFEM
int mat_create_from_pattern(mfs_ctx ctx, int bs, Pattern&& P, mfs_mat* out);
int mat_refresh_derived(mfs_mat A);  // FP32 copy + inverse diagonals from d_val64
int spmv_launch(mfs_mat A, int prec, const void* x, void* y, hipStream_t s);

// Algorithmic bytes of one SpMV (reference formula): nnz values, block
// indices, row pointers, x and y.
double spmv_bytes(mfs_mat A, int prec);

// Grid-wide deterministic reduction scratch: per-block partials + counter.
struct ReduceScratch {
  double* partials = nullptr;  // capacity doubles
  unsigned* counter = nullptr;
  int64_t capacity = 0;
};
int reduce_scratch_alloc(ReduceScratch& r, int64_t capacity);
void reduce_scratch_free(ReduceScratch& r);

}  // namespace mfs

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = MFS_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, MFS_WAVE);
  return v;
}

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of the
// MI355X (workgroup b runs on XCD b % 8), and each XCD has its own L2.
// xcd_block(1) renumbers them so that XCD x takes one contiguous range of
// the logical block order: in a SELL kernel, neighbouring slices, which
// gather the same operand lines, then share one L2.  A bijection of
// [0, gridDim.x) for any grid size.  xcd_block(0) keeps the dispatch order.
__device__ __forceinline__ unsigned xcd_block(int on) {
  if (!on) return blockIdx.x;
  const unsigned G = gridDim.x, b = blockIdx.x, x = b & 7u, per = G >> 3, rem = G & 7u;
  return x * per + min(x, rem) + (b >> 3);
}

// Element range [e0, e1) of this block of a vector-update kernel that
// feeds a SELL kernel of grid g1 (256 block rows of bs elements per block).
// With the XCD slice order (xcd = 1) the update blocks running on XCD x
// split the rows that the SELL kernel's blocks on XCD x own, so the values
// written are gathered next from the same L2; otherwise equal contiguous
// chunks.  Chunks are multiples of gran elements.
__device__ __forceinline__ void update_range(int64_t n, int xcd, unsigned g1, int bs, int gran, int64_t& e0,
                                             int64_t& e1) {
  const unsigned G = gridDim.x, b = blockIdx.x;
  if (!xcd || G < 8) {
    const int64_t chunk = ((n + G - 1) / G + gran - 1) / gran * gran;
    e0 = min(n, (int64_t)b * chunk);
    e1 = min(n, e0 + chunk);
    return;
  }
  const unsigned x = b & 7u, k = b >> 3, cnt = (G - x + 7) / 8;
  const unsigned per = g1 >> 3, rem = g1 & 7u;
  const int64_t s1 = (int64_t)x * per + min(x, rem), c1 = per + (x < rem ? 1 : 0);
  const int64_t r0 = min(n, s1 * 256 * bs), r1 = min(n, (s1 + c1) * 256 * bs);
  const int64_t chunk = ((r1 - r0 + cnt - 1) / cnt + gran - 1) / gran * gran;
  e0 = min(r1, r0 + (int64_t)k * chunk);
  e1 = min(r1, e0 + chunk);
}

// SELL-64 row product: thread `lane` of slice `slice` accumulates its block
// row of A X into acc[BS], where X(i) is supplied by the functor xl (x[i],
// or z[i] + beta p[i] for the fused CG kernel).  Values of entry j,
// component e live at val[(slice_base + j*64)*BS*BS + e*64 + lane]; matrix
// streams are read with non-temporal loads (used once per product), X
// through the caches.
//
// The row loop has a runtime trip count, so the compiler would issue one
// index load, wait, gather, wait per entry (two memory latencies each).
// Here the indices and values of U entries are loaded first and the U
// gathers are then issued together: two latencies per U entries.  Entries
// past the slice width re-read the last entry (a valid index) and use value
// 0, so every load is unconditional.
// Column indices are int32 block columns, or (CI = int16_t) block-column
// offsets from the row, which halve the index traffic when every offset of
// the matrix fits in 16 bits.
template <typename T, int BS, typename XL, typename CI>
__device__ __forceinline__ void sell_row_x(const int64_t* __restrict__ slice_base,
                                           const int32_t* __restrict__ slice_width, const CI* __restrict__ col,
                                           const T* __restrict__ val, int64_t slice, int lane, int64_t nb, XL xl,
                                           T (&acc)[BS]) {
  constexpr int C = 64;
  constexpr int U = BS == 1 ? 8 : 4;
  constexpr int B2 = BS * BS;
#pragma unroll
  for (int a = 0; a < BS; a++) acc[a] = 0;
  (void)nb;
  const int64_t sb = slice_base[slice];
  const int w = slice_width[slice];
  const CI* cp = col + sb + lane;
  const int32_t rbase = sizeof(CI) == 2 ? (int32_t)(slice * C + lane) : 0;
  const T* vp = val + sb * B2 + lane;
  for (int j0 = 0; j0 < w; j0 += U) {
    int32_t c[U];
    T v[U][B2];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = min(j0 + u, w - 1);
      const bool in = j0 + u < w;
      c[u] = (int32_t)__builtin_nontemporal_load(cp + (int64_t)j * C) + rbase;
#pragma unroll
      for (int e = 0; e < B2; e++) {
        const T t = __builtin_nontemporal_load(vp + ((int64_t)j * B2 + e) * C);
        v[u][e] = in ? t : (T)0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      T xv[BS];
#pragma unroll
      for (int b = 0; b < BS; b++) xv[b] = xl((int64_t)BS * c[u] + b);
#pragma unroll
      for (int a = 0; a < BS; a++)
#pragma unroll
        for (int b = 0; b < BS; b++) acc[a] += v[u][BS * a + b] * xv[b];
    }
  }
}

// Row products with two operands in one pass over the matrix: a1 = A_row·x1,
// a2 = A_row·x2 (same batching as sell_row_x).  Nothing here depends on a
// scalar computed by the caller, so a kernel can issue these loads before it
// has finished reducing the previous kernel's partial sums.
//
// The caller passes the slice's start and width (sb, w; w = 0 for a slice
// past the end), so that it can load them as early as it likes.
template <typename T, int BS, typename CI>
__device__ __forceinline__ void sell_row_pair(int64_t sb, int w, const CI* __restrict__ col,
                                              const T* __restrict__ val, const T* __restrict__ x1,
                                              const T* __restrict__ x2, int64_t slice, int lane, T (&a1)[BS],
                                              T (&a2)[BS]) {
  constexpr int C = 64;
  constexpr int U = BS == 1 ? 8 : 4;
  constexpr int B2 = BS * BS;
#pragma unroll
  for (int a = 0; a < BS; a++) a1[a] = a2[a] = 0;
  const CI* cp = col + sb + lane;
  const int32_t rbase = sizeof(CI) == 2 ? (int32_t)(slice * C + lane) : 0;
  const T* vp = val + sb * B2 + lane;
  for (int j0 = 0; j0 < w; j0 += U) {
    int32_t c[U];
    T v[U][B2];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = min(j0 + u, w - 1);
      const bool in = j0 + u < w;
      c[u] = (int32_t)__builtin_nontemporal_load(cp + (int64_t)j * C) + rbase;
#pragma unroll
      for (int e = 0; e < B2; e++) {
        const T t = __builtin_nontemporal_load(vp + ((int64_t)j * B2 + e) * C);
        v[u][e] = in ? t : (T)0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      T y1[BS], y2[BS];
#pragma unroll
      for (int b = 0; b < BS; b++) {
        y1[b] = x1[(int64_t)BS * c[u] + b];
        y2[b] = x2[(int64_t)BS * c[u] + b];
      }
#pragma unroll
      for (int a = 0; a < BS; a++)
#pragma unroll
        for (int b = 0; b < BS; b++) {
          a1[a] += v[u][BS * a + b] * y1[b];
          a2[a] += v[u][BS * a + b] * y2[b];
        }
    }
  }
}

template <typename T, int BS, typename CI>
__device__ __forceinline__ void sell_row(const int64_t* __restrict__ slice_base, const int32_t* __restrict__ slice_width,
                                         const CI* __restrict__ col, const T* __restrict__ val,
                                         const T* __restrict__ x, int64_t slice, int lane, int64_t nb, T (&acc)[BS]) {
  sell_row_x<T, BS>(slice_base, slice_width, col, val, slice, lane, nb, [&](int64_t i) { return x[i]; }, acc);
}

namespace mfs {
int spmv_dot_launch(mfs_mat A, int prec, const void* x, void* y, double* partials, unsigned* counter, double* out,
                    hipStream_t s);
int spmv_residual_launch(mfs_mat A, int prec, const void* x, void* r, const void* b, double* partials,
                         unsigned* counter, double* out, hipStream_t s);
int64_t spmv_grid(mfs_mat A);
}  // namespace mfs

// Fence-free two-stage reductions: a kernel writes one partial per block and
// value (part[k*G + b]); k_finish_sums, one block per value, adds them in
// block order.
template <typename T, int NV>
__device__ __forceinline__ void write_block_partials(T (&v)[NV], double* __restrict__ part) {
  __shared__ double red_bp[NV][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum((double)v[k]);
    if (lane == 0) red_bp[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0;
    for (int w = 0; w < nw; w++) s += red_bp[threadIdx.x][w];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void k_finish_sums(const double* __restrict__ part, unsigned G, double* __restrict__ out);

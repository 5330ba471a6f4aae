// GPU assembly of the FEM system on gfx950.
//
//  1. k_tet_geometry: one thread per tet computes the P1 gradients and the
//     volume-scaled material factors (V*sigma, or V*lambda and V*mu).
//  2. k_gather: one thread per SELL-64 entry sums that entry's element
//     contributions in ascending tet order (the order in which the
//     reference's serial SummPartOfMatrix visits elements, FEM.h:2146-2155)
//     and writes the entry's values.  The contribution lists are stored
//     interleaved per group of 64 entries (entry k of the 64 lists side by
//     side), so a wave reads one list position of its 64 entries with one
//     coalesced load and writes each value of its 64 blocks with one
//     coalesced store.  No atomics, so results are deterministic.
//  3. k_dirichlet_*: elimination of fixed DOFs (FEM.h:505-619): fixed rows
//     become identity rows, fixed columns move to the right-hand side.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "mfs_device.hpp"

using namespace mfs;

struct mfs_asm_s {
  mfs_ctx ctx = nullptr;
  int problem = MFS_PROBLEM_SCALAR;
  int convention = MFS_ELEM_REFERENCE;
  int bs = 1;
  int64_t nn = 0, nt = 0;
  std::vector<int32_t> map;  // original node -> DOF node
  mfs_mat A = nullptr;
  mfs_vec b = nullptr, x0 = nullptr;
  // device
  double* d_xyz = nullptr;      // 3 per DOF node
  int32_t* d_tets = nullptr;    // 4 per tet, DOF node ids
  double* d_coef = nullptr;     // per tet
  double* d_tet = nullptr;      // 16 doubles per tet: g[4][3], f0, f1, V, pad
  int64_t* d_goff = nullptr;    // per group of 64 SELL entries: start of its interleaved lists (+1)
  uint32_t* d_gil = nullptr;    // interleaved contribution lists (kNoContribution pads)
  int32_t* d_browptr = nullptr; // nb+1
  uint8_t* d_fixed = nullptr;   // per scalar DOF
  double* d_g = nullptr;        // per scalar DOF (fixed values)
  double* d_fbase = nullptr;    // Neumann loads per scalar DOF
  // fused per-chunk path (when every chunk touches at most kMaxSliceTets tets);
  // a chunk is chunk_rows consecutive block rows of one SELL slice
  bool fused = false;
  int chunk_rows = 64;
  size_t fused_lds = 0;
  int64_t* d_sptr = nullptr;    // n_chunks+1
  int32_t* d_stet = nullptr;    // per chunk, its tets (ascending)
  int4* d_snode = nullptr;      // per chunk, its tets' node ids (same order)
  uint16_t* d_gil16 = nullptr;  // interleaved lists of local codes (chunk-local tet << 4 | i << 2 | j)
  uint8_t* d_slice_fix = nullptr;  // per chunk: a row or an entry column is fixed
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
// Gradients of the 4 barycentric functions: grad phi_1..3 = rows of J^{-T}
// with J = [p1-p0, p2-p0, p3-p0]; grad phi_0 = -(sum).  V = |det J| / 6.
// o[0..11] = g[4][3], o[12] = V sigma (scalar) or V lambda, o[13] = 0 or
// V mu, o[14] = V.
__device__ __forceinline__ void tet_coords(const double* __restrict__ xyz, int4 t, double (&p)[4][3]) {
  const int32_t id[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
  for (int v = 0; v < 4; v++)
#pragma unroll
    for (int a = 0; a < 3; a++) p[v][a] = xyz[3 * (int64_t)id[v] + a];
}

__device__ __forceinline__ void tet_record_from(const double (&p)[4][3], double c, int problem, double nu,
                                                double (&o)[15]) {
  // no fused multiply-adds: every assembly path (and the oracle, built with
  // -ffp-contract=off) rounds the same operations the same way
#pragma clang fp contract(off)
  double J[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int a = 0; a < 3; a++) J[r][a] = p[r + 1][a] - p[0][a];
  // cofactors: inverse(J)^T rows give grad phi_{1..3}
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double c10 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  const double c11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  const double c12 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  const double c20 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  const double c21 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  const double c22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double inv = 1.0 / det;
  double g[4][3];
  // grad phi_k (k=1..3) = column k-1 of J^{-1} = (cofactor row k-1) / det
  g[1][0] = c00 * inv; g[1][1] = c01 * inv; g[1][2] = c02 * inv;
  g[2][0] = c10 * inv; g[2][1] = c11 * inv; g[2][2] = c12 * inv;
  g[3][0] = c20 * inv; g[3][1] = c21 * inv; g[3][2] = c22 * inv;
#pragma unroll
  for (int a = 0; a < 3; a++) g[0][a] = -(g[1][a] + g[2][a] + g[3][a]);
  const double V = fabs(det) / 6.0;
#pragma unroll
  for (int v = 0; v < 4; v++)
#pragma unroll
    for (int a = 0; a < 3; a++) o[3 * v + a] = g[v][a];
  if (problem == MFS_PROBLEM_SCALAR) {
    o[12] = V * c;
    o[13] = 0.0;
  } else {
    const double lam = c * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = c / (2.0 * (1.0 + nu));
    o[12] = V * lam;
    o[13] = V * mu;
  }
  o[14] = V;
}

__device__ __forceinline__ void tet_record(const double* __restrict__ xyz, const int32_t* __restrict__ tets,
                                           const double* __restrict__ coef, int64_t e, int problem, double nu,
                                           double (&o)[15]) {
  double p[4][3];
  tet_coords(xyz, *reinterpret_cast<const int4*>(tets + 4 * e), p);
  tet_record_from(p, coef[e], problem, nu, o);
}

__global__ __launch_bounds__(256) void k_tet_geometry(const double* __restrict__ xyz, const int32_t* __restrict__ tets,
                                                      const double* __restrict__ coef, int64_t nt, int problem,
                                                      double nu, double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nt; e += (int64_t)gridDim.x * blockDim.x) {
    double o[15];
    tet_record(xyz, tets, coef, e, problem, nu, o);
    double* d = out + 16 * e;
#pragma unroll
    for (int k = 0; k < 15; k++) d[k] = o[k];
    d[15] = 0.0;
  }
}

constexpr uint32_t kNoContribution = 0xffffffffu;

// Element block (I=i, J=j) of tet record o (BS = 1: sigma V gI.gJ; BS = 3:
//   textbook  K[a][b] = V(lam dI_a dJ_b + mu dI_b dJ_a + mu delta_ab dI.dJ)
//   reference K[a][b] = textbook[b][a] (FEM_Grid.h:2159-2191 puts B_J^T on
//   the left of D)), added to acc.
template <int BS, int CONV>
__device__ __forceinline__ void add_block(const double (&g)[8], double (&acc)[BS * BS]) {
#pragma clang fp contract(off)
  const double* dI = g;
  const double* dJ = g + 3;
  const double dd = dI[0] * dJ[0] + dI[1] * dJ[1] + dI[2] * dJ[2];
  if constexpr (BS == 1) {
    acc[0] += g[6] * dd;
  } else {
    const double vl = g[6], vm = g[7];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        double v;
        if (CONV == MFS_ELEM_TEXTBOOK)
          v = vl * dI[a] * dJ[b] + vm * dI[b] * dJ[a];
        else
          v = vl * dI[b] * dJ[a] + vm * dI[a] * dJ[b];
        if (a == b) v += vm * dd;
        acc[3 * a + b] += v;
      }
  }
}

// One thread per SELL entry p (group p / 64, lane p % 64): the entry's
// contributions are gil[goff[g] + k * 64 + lane], k ascending = tet order.
template <int BS, int CONV>
__global__ __launch_bounds__(256) void k_gather(const int64_t* __restrict__ goff, const uint32_t* __restrict__ gil,
                                                const double* __restrict__ tet, int64_t n_entries,
                                                double* __restrict__ sell) {
  constexpr int C = 64;
  constexpr int U = 4;  // list positions whose loads are in flight together
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n_entries;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t grp = p / C;
    const int lane = (int)(p % C);
    const int64_t k0 = goff[grp];
    const int K = (int)((goff[grp + 1] - k0) / C);
    double acc[BS * BS];
#pragma unroll
    for (int e = 0; e < BS * BS; e++) acc[e] = 0.0;
    for (int kb = 0; kb < K; kb += U) {
      // list entries first, then the records they point at, then the sums in
      // ascending order; positions past this lane's list read tet 0 and are
      // skipped
      uint32_t c[U];
#pragma unroll
      for (int u = 0; u < U; u++) c[u] = kb + u < K ? gil[k0 + (int64_t)(kb + u) * C + lane] : kNoContribution;
      double g[U][8];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t cu = c[u] == kNoContribution ? 0u : c[u];
        const double* o = tet + 16 * (int64_t)(cu >> 4);
        const int i = (cu >> 2) & 3, j = cu & 3;
        g[u][0] = o[3 * i];
        g[u][1] = o[3 * i + 1];
        g[u][2] = o[3 * i + 2];
        g[u][3] = o[3 * j];
        g[u][4] = o[3 * j + 1];
        g[u][5] = o[3 * j + 2];
        g[u][6] = o[12];
        g[u][7] = o[13];
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (c[u] != kNoContribution) add_block<BS, CONV>(g[u], acc);
    }
    if constexpr (BS == 1) {
      sell[p] = acc[0];
    } else {
      const int64_t base = grp * C * 9 + lane;
#pragma unroll
      for (int e = 0; e < 9; e++) sell[base + (int64_t)e * C] = acc[e];
    }
  }
}

// Dirichlet elimination, one thread per block row.  fixed/g indexed by
// scalar DOF.  Row-fixed: zero the row, diagonal 1, rhs = g.  Otherwise every
// fixed column c: rhs -= a_rc g_c, a_rc = 0.
template <int BS>
__global__ __launch_bounds__(256) void k_dirichlet(const int32_t* __restrict__ browptr, const int32_t* __restrict__ col,
                                                   const int64_t* __restrict__ slice_base, double* __restrict__ sell,
                                                   const uint8_t* __restrict__ fixed, const double* __restrict__ g,
                                                   const double* __restrict__ fbase, double* __restrict__ rhs,
                                                   double* __restrict__ x0, int64_t nb) {
#pragma clang fp contract(off)
  constexpr int C = 64;
  constexpr int U = 4;  // entries whose index / flag loads are in flight together
  for (int64_t R = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; R < nb; R += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = R / C, lane = R % C;
    const int64_t sb = slice_base[q];
    const int len = browptr[R + 1] - browptr[R];
    double f[BS];
    bool rf[BS];
#pragma unroll
    for (int a = 0; a < BS; a++) {
      f[a] = fbase[BS * R + a];
      rf[a] = fixed[BS * R + a] != 0;
    }
    for (int j0 = 0; j0 < len; j0 += U) {
      // column indices first, then the fixed flags and values they select;
      // entries past the row end re-read the last one and are skipped
      int32_t c[U];
#pragma unroll
      for (int u = 0; u < U; u++) c[u] = col[sb + (int64_t)min(j0 + u, len - 1) * C + lane];
      uint8_t fx[U][BS];
      double gv[U][BS];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int b = 0; b < BS; b++) {
          const int64_t cj = (int64_t)BS * c[u] + b;
          fx[u][b] = fixed[cj];
          gv[u][b] = g[cj];
        }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (j0 + u >= len) continue;
        const int64_t vbase = (sb + (int64_t)(j0 + u) * C) * (BS * BS) + lane;
#pragma unroll
        for (int a = 0; a < BS; a++)
#pragma unroll
          for (int b = 0; b < BS; b++) {
            const int64_t vi = vbase + (int64_t)(BS * a + b) * C;
            const int64_t cj = (int64_t)BS * c[u] + b;
            if (rf[a]) {
              sell[vi] = (cj == BS * R + a) ? 1.0 : 0.0;
            } else if (fx[u][b]) {
              f[a] -= sell[vi] * gv[u][b];
              sell[vi] = 0.0;
            }
          }
      }
    }
#pragma unroll
    for (int a = 0; a < BS; a++) {
      const int64_t i = BS * R + a;
      rhs[i] = rf[a] ? g[i] : f[a];
      x0[i] = rf[a] ? g[i] : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused assembly, one workgroup per chunk of CL consecutive block rows of a
// SELL slice (CL = 64, 32 or 16; a slice's 64 rows split into 64 / CL
// chunks, each with its own tet list, so a smaller chunk needs less LDS and
// lets more workgroups share a CU; along a row of nodes the chunks' tet
// lists hardly overlap, so the split recomputes few records):
//  1. the records of the tets touching the chunk's rows (host list, in
//     ascending tet order) are computed into LDS, 14 doubles each
//     (grad phi_0..3 and the two material factors, as tet_record forms
//     them);
//  2. one thread per SELL entry of the chunk sums its contributions (16-bit
//     chunk-local codes, interleaved per group of 64 entries, ascending tet
//     order) from LDS,
//     turns fixed rows into identity rows, and writes the FP64 value, its
//     FP32 copy and, at the diagonal entry, the inverse diagonal;
//  3. the chunk's rows then get their right-hand side; where the host
//     flagged a fixed column in the chunk, the row moves those columns to
//     the right-hand side in slot order (as k_dirichlet does) and
//     refreshes its inverse diagonal.
// The arithmetic and its order are those of k_tet_geometry + k_gather +
// k_dirichlet + the FP32 / diagonal derivation, so both paths give the
// same bits; this one reads each tet record from LDS instead of HBM and
// writes each value once.
constexpr int kRec = 14;               // doubles per tet record in LDS
constexpr int kMaxSliceTets = 1400;    // 1400 * 112 B = 153.1 KB of LDS
constexpr uint16_t kNoContribution16 = 0xffff;
constexpr int kDefaultChunkRows = 16;
// threads per chunk workgroup: 16 per chunk row, so that one pass over the
// chunk's entries (about 15 per row for the elastic stencil) needs one round
constexpr int kChunkThreads(int cl) { return 16 * cl; }

template <int BS, int CONV, int CL, int NTH>
__global__ __launch_bounds__(NTH) void k_assemble_slice(
    const double* __restrict__ xyz, const int32_t* __restrict__ tets, const double* __restrict__ coef, int problem,
    double nu, const int64_t* __restrict__ sptr, const int32_t* __restrict__ stet, const int4* __restrict__ snode,
    const int64_t* __restrict__ goff,
    const uint16_t* __restrict__ gil, const int64_t* __restrict__ slice_base, const int32_t* __restrict__ slice_width,
    const int32_t* __restrict__ browptr, const int32_t* __restrict__ col, const uint8_t* __restrict__ slice_fix,
    const uint8_t* __restrict__ fixed, const double* __restrict__ g, const double* __restrict__ fbase,
    double* __restrict__ val64, float* __restrict__ val32, double* __restrict__ dinv64, float* __restrict__ dinv32,
    double* __restrict__ rhs, double* __restrict__ x0, int64_t nb) {
#pragma clang fp contract(off)
  constexpr int C = 64, SUB = C / CL;
  static_assert(C % CL == 0, "a chunk is a whole fraction of a slice");
  extern __shared__ double rec[];
  const int64_t chunk = blockIdx.x;
  const int64_t q = chunk / SUB;           // SELL slice
  const int h0 = (int)(chunk % SUB) * CL;  // first lane of the chunk
  const int tid = threadIdx.x, nth = NTH;
  const int64_t t0 = sptr[chunk];
  const int T = (int)(sptr[chunk + 1] - t0);
  const int w = slice_width[q];
  const int64_t sb = slice_base[q];
  constexpr int KB = 16;  // list positions whose codes are requested together
  // the list offsets and first KB codes of this thread's first entry are
  // requested before the records, so they arrive while phase 1 runs
  int64_t pk0 = 0;
  int pK = 0;
  uint16_t pc[KB];
  // and the entry's row facts for the write-out: row length, column, fixed DOFs
  int plen = 0;
  int32_t pcol = 0;
  uint8_t pfx[BS];
  double pfb[BS];  // Neumann loads of the row pass's row (threads < CL)
#pragma unroll
  for (int a = 0; a < BS; a++) {
    pfx[a] = 0;
    pfb[a] = 0.0;
  }
  const bool chunk_fix = slice_fix[chunk] != 0;
  if (tid < w * CL) {
    const int j = tid / CL, lane = h0 + tid % CL;
    const int64_t grp = sb / C + j;
    pk0 = goff[grp];
    pK = (int)((goff[grp + 1] - pk0) / C);
    const int64_t R = q * C + lane;
    if (R < nb) {
      plen = browptr[R + 1] - browptr[R];
      pcol = col[sb + (int64_t)j * C + lane];
#pragma unroll
      for (int a = 0; a < BS; a++) pfx[a] = fixed[BS * R + a];
      if (tid < CL)  // the row pass below handles this thread's row
#pragma unroll
        for (int a = 0; a < BS; a++) pfb[a] = fbase[BS * R + a];
    }
  }
  // phase 1: the chunk list carries each tet's node ids, so the coordinate
  // loads wait for one list load (and the modulus for the tet id); a thread
  // with two records requests both before computing either
  auto put = [&](int t, const double (&o)[15]) {
    double* r = rec + (int64_t)t * kRec;
#pragma unroll
    for (int k = 0; k < 12; k++) r[k] = o[k];
    r[12] = o[12];
    r[13] = o[13];
  };
  for (int t = tid; t < T; t += 2 * nth) {
    const bool two = t + nth < T;
    const int4 nd = snode[t0 + t];
    const int32_t e = stet[t0 + t];
    const int4 nd2 = two ? snode[t0 + t + nth] : nd;
    const int32_t e2 = two ? stet[t0 + t + nth] : e;
    if (t == tid && tid < w * CL) {
      const int lane = h0 + tid % CL;
#pragma unroll
      for (int u = 0; u < KB; u++) pc[u] = u < pK ? gil[pk0 + (int64_t)u * C + lane] : kNoContribution16;
    }
    double p[4][3], p2[4][3];
    tet_coords(xyz, nd, p);
    if (two) tet_coords(xyz, nd2, p2);
    const double ce = coef[e], ce2 = two ? coef[e2] : 0.0;
    double o[15];
    tet_record_from(p, ce, problem, nu, o);
    put(t, o);
    if (two) {
      tet_record_from(p2, ce2, problem, nu, o);
      put(t + nth, o);
    }
  }
  if (tid >= T && tid < w * CL) {  // a chunk with fewer tets than threads
    const int lane = h0 + tid % CL;
#pragma unroll
    for (int u = 0; u < KB; u++) pc[u] = u < pK ? gil[pk0 + (int64_t)u * C + lane] : kNoContribution16;
  }
  __syncthreads();
  for (int p = tid; p < w * CL; p += nth) {
    const int j = p / CL, lane = h0 + p % CL;
    const bool first = p == tid;  // the prefetched entry
    int64_t k0;
    int K;
    if (first) {
      k0 = pk0;
      K = pK;
    } else {
      const int64_t grp = sb / C + j;
      k0 = goff[grp];
      K = (int)((goff[grp + 1] - k0) / C);
    }
    double acc[BS * BS];
#pragma unroll
    for (int e = 0; e < BS * BS; e++) acc[e] = 0.0;
    for (int kb = 0; kb < K; kb += KB) {
      // all codes of the batch in flight at once, then the sums from LDS
      uint16_t c[KB];
      if (first && kb == 0) {
#pragma unroll
        for (int u = 0; u < KB; u++) c[u] = pc[u];
      } else {
#pragma unroll
        for (int u = 0; u < KB; u++) c[u] = kb + u < K ? gil[k0 + (int64_t)(kb + u) * C + lane] : kNoContribution16;
      }
#pragma unroll
      for (int u = 0; u < KB; u++) {
        if (c[u] == kNoContribution16) continue;
        const double* r = rec + (int64_t)(c[u] >> 4) * kRec;
        const int i = (c[u] >> 2) & 3, jj = c[u] & 3;
        double gr[8];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          gr[a] = r[3 * i + a];
          gr[3 + a] = r[3 * jj + a];
        }
        gr[6] = r[12];
        gr[7] = r[13];
        add_block<BS, CONV>(gr, acc);
      }
    }
    const int64_t R = q * C + lane;
    const int64_t ent = sb + (int64_t)j * C + lane;
    const int64_t vbase = (sb + (int64_t)j * C) * (BS * BS) + lane;
    // padding entries stay 0
    const bool real = R < nb && j < (first ? plen : browptr[R + 1] - browptr[R]);
    int32_t cc = 0;
    if (real) {
      cc = first ? pcol : col[ent];
#pragma unroll
      for (int a = 0; a < BS; a++)
        if (first ? pfx[a] : fixed[BS * R + a])
#pragma unroll
          for (int b = 0; b < BS; b++) acc[BS * a + b] = ((int64_t)BS * cc + b == BS * R + a) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int e = 0; e < BS * BS; e++) {
      val64[vbase + (int64_t)e * C] = acc[e];
      val32[vbase + (int64_t)e * C] = (float)acc[e];
    }
    if (real && cc == R) {
#pragma unroll
      for (int a = 0; a < BS; a++) {
        const double d = acc[BS * a + a];
        const double inv = d != 0.0 ? 1.0 / d : 1.0;
        dinv64[BS * R + a] = inv;
        dinv32[BS * R + a] = (float)inv;
      }
    }
  }
  __syncthreads();  // the chunk's values are visible to the row pass below
  if (tid < CL) {
    const int lane = h0 + tid;
    const int64_t R = q * C + lane;
    if (R < nb) {
      double f[BS];
      bool rf[BS];
#pragma unroll
      for (int a = 0; a < BS; a++) {
        f[a] = pfb[a];
        rf[a] = pfx[a] != 0;
      }
      if (chunk_fix) {
        const int len = browptr[R + 1] - browptr[R];
        int jd = -1;
        constexpr int U3 = 8;  // entries whose column and flag loads are in flight together
        for (int j0 = 0; j0 < len; j0 += U3) {
          int32_t cj0[U3];
          uint8_t fx[U3][BS];
#pragma unroll
          for (int u = 0; u < U3; u++) cj0[u] = col[sb + (int64_t)min(j0 + u, len - 1) * C + lane];
#pragma unroll
          for (int u = 0; u < U3; u++)
#pragma unroll
            for (int b = 0; b < BS; b++) fx[u][b] = fixed[(int64_t)BS * cj0[u] + b];
#pragma unroll
          for (int u = 0; u < U3; u++) {
            if (j0 + u >= len) continue;
            if (cj0[u] == R) jd = j0 + u;
            const int64_t vb = (sb + (int64_t)(j0 + u) * C) * (BS * BS) + lane;
#pragma unroll
            for (int a = 0; a < BS; a++)
#pragma unroll
              for (int b = 0; b < BS; b++) {
                const int64_t cj = (int64_t)BS * cj0[u] + b;
                if (!rf[a] && fx[u][b]) {
                  const int64_t vi = vb + (int64_t)(BS * a + b) * C;
                  f[a] -= val64[vi] * g[cj];
                  val64[vi] = 0.0;
                  val32[vi] = 0.0f;
                }
              }
          }
        }
        if (jd >= 0) {
          const int64_t vb = (sb + (int64_t)jd * C) * (BS * BS) + lane;
#pragma unroll
          for (int a = 0; a < BS; a++) {
            const double d = val64[vb + (int64_t)(BS * a + a) * C];
            const double inv = d != 0.0 ? 1.0 / d : 1.0;
            dinv64[BS * R + a] = inv;
            dinv32[BS * R + a] = (float)inv;
          }
        }
      }
#pragma unroll
      for (int a = 0; a < BS; a++) {
        const int64_t i = BS * R + a;
        rhs[i] = rf[a] ? g[i] : f[a];
        x0[i] = rf[a] ? g[i] : 0.0;
      }
    }
  }
}

static inline unsigned grid_cap(int64_t n, int64_t cap = 65536) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap));
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int mfs_assemble_destroy(mfs_asm S) {
  if (!S) return MFS_OK;
  if (S->ctx) (void)hipStreamSynchronize(S->ctx->stream);
  mfs_mat_destroy(S->A);
  mfs_vec_destroy(S->b);
  mfs_vec_destroy(S->x0);
  dfree(S->d_xyz);
  dfree(S->d_tets);
  dfree(S->d_coef);
  dfree(S->d_tet);
  dfree(S->d_goff);
  dfree(S->d_gil);
  dfree(S->d_browptr);
  dfree(S->d_fixed);
  dfree(S->d_g);
  dfree(S->d_fbase);
  dfree(S->d_sptr);
  dfree(S->d_stet);
  dfree(S->d_snode);
  dfree(S->d_gil16);
  dfree(S->d_slice_fix);
  mfs_ctx c = S->ctx;
  delete S;
  ctx_release(c);
  return MFS_OK;
}

extern "C" int mfs_assemble_create(mfs_ctx ctx, const mfs_fem_desc* d, mfs_asm* out) {
  MFS_CHECK_ARG(ctx && d && out, "mfs_assemble_create: bad argument");
  MFS_CHECK_ARG(d->problem == MFS_PROBLEM_SCALAR || d->problem == MFS_PROBLEM_ELASTIC,
                "mfs_assemble_create: unknown problem");
  MFS_CHECK_ARG(d->n_nodes > 0 && d->n_tets > 0 && d->xyz && d->tets, "mfs_assemble_create: empty mesh");
  MFS_CHECK_ARG(d->n_tets < ((int64_t)1 << 28), "mfs_assemble_create: more than 2^28 tets");
  MFS_CHECK_ARG(d->n_fixed == 0 || (d->fixed_nodes && d->fixed_values), "mfs_assemble_create: fixed arrays missing");
  MFS_CHECK_ARG(d->n_tris == 0 || (d->tris && d->traction), "mfs_assemble_create: Neumann arrays missing");
  MFS_CHECK_ARG(d->convention == MFS_ELEM_REFERENCE || d->convention == MFS_ELEM_TEXTBOOK,
                "mfs_assemble_create: unknown element convention");
  const int64_t nn = d->n_nodes, nt = d->n_tets;
  for (int64_t k = 0; k < 4 * nt; k++)
    MFS_CHECK_ARG(d->tets[k] >= 0 && d->tets[k] < nn, "mfs_assemble_create: tet node out of range");
  for (int64_t k = 0; k < d->n_fixed; k++)
    MFS_CHECK_ARG(d->fixed_nodes[k] >= 0 && d->fixed_nodes[k] < nn, "mfs_assemble_create: fixed node out of range");
  for (int64_t k = 0; k < 3 * d->n_tris; k++)
    MFS_CHECK_ARG(d->tris[k] >= 0 && d->tris[k] < nn, "mfs_assemble_create: triangle node out of range");
  MFS_HIP(hipSetDevice(ctx->device));
  mfs_asm S = new mfs_asm_s();
  S->ctx = ctx;
  ctx_retain(ctx);
  S->problem = d->problem;
  S->convention = d->convention;
  S->bs = d->problem == MFS_PROBLEM_ELASTIC ? 3 : 1;
  S->nn = nn;
  S->nt = nt;
  S->map.resize(nn);
  if (d->renumber)
    renumber_mech(nn, nt, d->tets, d->material, d->n_fixed, d->fixed_nodes, S->map.data());
  else
    std::iota(S->map.begin(), S->map.end(), 0);
  const std::vector<int32_t>& map = S->map;
  std::vector<int32_t> rt(4 * nt);
  for (int64_t k = 0; k < 4 * nt; k++) rt[k] = map[d->tets[k]];
  std::vector<double> rxyz(3 * nn);
  for (int64_t v = 0; v < nn; v++)
    for (int a = 0; a < 3; a++) rxyz[3 * (int64_t)map[v] + a] = d->xyz[3 * v + a];
  Pattern P;
  build_pattern(nn, nt, rt.data(), P);
  GatherLists G;
  build_gather_lists(P, nt, rt.data(), G);
  std::vector<int32_t> browptr(P.rowptr);
  const int bs = S->bs;
  const int64_t ndof = bs * nn;
  std::vector<uint8_t> fixed(ndof, 0);
  std::vector<double> g(ndof, 0.0), fbase(ndof, 0.0);
  for (int64_t k = 0; k < d->n_fixed; k++) {
    int64_t node = map[d->fixed_nodes[k]];
    for (int a = 0; a < bs; a++) {
      fixed[bs * node + a] = 1;
      g[bs * node + a] = d->fixed_values[bs * k + a];
    }
  }
  // Neumann loads: f_I += t * area / 3 for each triangle node (exact P1
  // integral of the reference's 4-point triangle rule, FEM_Grid.h:2622-2661)
  if (bs == 3)
    for (int64_t t = 0; t < d->n_tris; t++) {
      const double* p0 = d->xyz + 3 * (int64_t)d->tris[3 * t];
      const double* p1 = d->xyz + 3 * (int64_t)d->tris[3 * t + 1];
      const double* p2 = d->xyz + 3 * (int64_t)d->tris[3 * t + 2];
      const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      const double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
      const double cx = u[1] * w[2] - u[2] * w[1], cy = u[2] * w[0] - u[0] * w[2], cz = u[0] * w[1] - u[1] * w[0];
      const double third = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz) / 3.0;
      for (int q = 0; q < 3; q++) {
        const int64_t node = map[d->tris[3 * t + q]];
        for (int a = 0; a < 3; a++) fbase[3 * node + a] += d->traction[3 * t + a] * third;
      }
    }
  int st = mat_create_from_pattern(ctx, bs, std::move(P), &S->A);
  // contribution lists interleaved per group of 64 SELL entries (k-th
  // contribution of the group's 64 entries side by side, padded to the
  // group's longest list).  Fused path: codes local to the slice's tet
  // list (16 bits); otherwise global codes (32 bits).
  std::vector<int64_t> goff, sptr;
  std::vector<uint32_t> gil;
  std::vector<uint16_t> gil16;
  std::vector<int32_t> stet;
  std::vector<int4> snode;
  std::vector<uint8_t> slice_fix;
  if (st == MFS_OK) {
    const SellLayout& L = S->A->sell;
    const Pattern& Pt = S->A->pat;
    const int C = L.C;
    const int64_t ngroups = L.slice_base[L.n_slices] / C;
    // MFS_ASSEMBLY_CHUNK=64|32|16 picks the chunk height (A/B runs and tests)
    const char* evc = std::getenv("MFS_ASSEMBLY_CHUNK");
    const int want = evc ? std::atoi(evc) : kDefaultChunkRows;
    S->chunk_rows = want == 16 || want == 32 || want == 64 ? want : kDefaultChunkRows;
    const int CL = S->chunk_rows, SUB = C / CL;
    build_slice_tets(nn, nt, rt.data(), CL, sptr, stet);
    const int64_t n_chunks = L.n_slices * SUB;
    if ((int64_t)sptr.size() < n_chunks + 1) sptr.resize(n_chunks + 1, sptr.back());  // padded rows of the last slice
    int64_t tmax = 0;
    for (int64_t c = 0; c < n_chunks; c++) tmax = std::max(tmax, sptr[c + 1] - sptr[c]);
    // MFS_ASSEMBLY_SLICES=0 keeps the general path (A/B runs and tests)
    const char* ev = std::getenv("MFS_ASSEMBLY_SLICES");
    S->fused = tmax <= kMaxSliceTets && !(ev && ev[0] == '0');
    S->fused_lds = (size_t)std::max<int64_t>(1, tmax) * kRec * sizeof(double);
    if (S->fused) {
      snode.resize(stet.size());
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)stet.size(); i++) {
        const int32_t* t = rt.data() + 4 * (int64_t)stet[i];
        snode[i] = make_int4(t[0], t[1], t[2], t[3]);
      }
    }
    goff.assign(ngroups + 1, 0);
    auto slot_of = [&](int64_t q, int j, int lane) -> int64_t {
      const int64_t r = q * C + lane;
      if (r >= Pt.n || j >= Pt.rowptr[r + 1] - Pt.rowptr[r]) return -1;
      return Pt.rowptr[r] + j;
    };
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < L.n_slices; q++)
      for (int j = 0; j < L.slice_width[q]; j++) {
        int64_t K = 0;
        for (int lane = 0; lane < C; lane++) {
          const int64_t sl = slot_of(q, j, lane);
          if (sl >= 0) K = std::max<int64_t>(K, G.ptr[sl + 1] - G.ptr[sl]);
        }
        goff[L.slice_base[q] / C + j + 1] = K * C;
      }
    for (int64_t g = 0; g < ngroups; g++) goff[g + 1] += goff[g];
    if (S->fused) {
      gil16.assign(goff[ngroups], kNoContribution16);
      slice_fix.assign(n_chunks, 0);
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t ck = 0; ck < n_chunks; ck++) {
        const int64_t q = ck / SUB;
        const int h0 = (int)(ck % SUB) * CL;
        const int32_t* t0 = stet.data() + sptr[ck];
        const int32_t* t1 = stet.data() + sptr[ck + 1];
        bool fx = false;
        for (int lane = h0; lane < h0 + CL && !fx; lane++) {
          const int64_t r = q * C + lane;
          if (r >= Pt.n) break;
          for (int a = 0; a < bs; a++) fx = fx || fixed[bs * r + a];
          for (int64_t k = Pt.rowptr[r]; k < Pt.rowptr[r + 1] && !fx; k++)
            for (int a = 0; a < bs; a++) fx = fx || fixed[(int64_t)bs * Pt.col[k] + a];
        }
        slice_fix[ck] = fx;
        for (int j = 0; j < L.slice_width[q]; j++) {
          const int64_t base = goff[L.slice_base[q] / C + j];
          for (int lane = h0; lane < h0 + CL; lane++) {
            const int64_t sl = slot_of(q, j, lane);
            if (sl < 0) continue;
            for (int64_t k = G.ptr[sl]; k < G.ptr[sl + 1]; k++) {
              const uint32_t code = G.entry[k];
              const int32_t e = (int32_t)(code >> 4);
              const int64_t loc = std::lower_bound(t0, t1, e) - t0;
              gil16[base + (k - G.ptr[sl]) * C + lane] = (uint16_t)((loc << 4) | (code & 15));
            }
          }
        }
      }
    } else {
      gil.assign(goff[ngroups], kNoContribution);
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < L.n_slices; q++)
        for (int j = 0; j < L.slice_width[q]; j++) {
          const int64_t base = goff[L.slice_base[q] / C + j];
          for (int lane = 0; lane < C; lane++) {
            const int64_t sl = slot_of(q, j, lane);
            if (sl < 0) continue;
            for (int64_t k = G.ptr[sl]; k < G.ptr[sl + 1]; k++) gil[base + (k - G.ptr[sl]) * C + lane] = G.entry[k];
          }
        }
    }
  }
  if (st == MFS_OK) st = mfs_vec_create(ctx, ndof, MFS_F64, &S->b);
  if (st == MFS_OK) st = mfs_vec_create(ctx, ndof, MFS_F64, &S->x0);
  auto alloc = [&](void** p, size_t bytes) {
    if (st == MFS_OK) st = dmalloc(p, bytes);
  };
  alloc((void**)&S->d_xyz, 8 * 3 * nn);
  alloc((void**)&S->d_tets, 4 * 4 * nt);
  alloc((void**)&S->d_coef, 8 * nt);
  if (!S->fused) alloc((void**)&S->d_tet, 8 * 16 * nt);
  alloc((void**)&S->d_goff, 8 * goff.size());
  if (S->fused) {
    alloc((void**)&S->d_gil16, 2 * std::max<size_t>(1, gil16.size()));
    alloc((void**)&S->d_sptr, 8 * sptr.size());
    alloc((void**)&S->d_stet, 4 * std::max<size_t>(1, stet.size()));
    alloc((void**)&S->d_snode, 16 * std::max<size_t>(1, stet.size()));
    alloc((void**)&S->d_slice_fix, std::max<size_t>(1, slice_fix.size()));
  } else {
    alloc((void**)&S->d_gil, 4 * std::max<size_t>(1, gil.size()));
  }
  alloc((void**)&S->d_browptr, 4 * browptr.size());
  alloc((void**)&S->d_fixed, ndof);
  alloc((void**)&S->d_g, 8 * ndof);
  alloc((void**)&S->d_fbase, 8 * ndof);
  if (st != MFS_OK) {
    mfs_assemble_destroy(S);
    return st;
  }
  hipStream_t s = ctx->stream;
  MFS_HIP(hipMemcpyAsync(S->d_xyz, rxyz.data(), 8 * 3 * nn, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(S->d_tets, rt.data(), 4 * 4 * nt, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(S->d_goff, goff.data(), 8 * goff.size(), hipMemcpyHostToDevice, s));
  if (!gil.empty()) MFS_HIP(hipMemcpyAsync(S->d_gil, gil.data(), 4 * gil.size(), hipMemcpyHostToDevice, s));
  if (S->fused) {
    if (!gil16.empty())
      MFS_HIP(hipMemcpyAsync(S->d_gil16, gil16.data(), 2 * gil16.size(), hipMemcpyHostToDevice, s));
    MFS_HIP(hipMemcpyAsync(S->d_sptr, sptr.data(), 8 * sptr.size(), hipMemcpyHostToDevice, s));
    if (!stet.empty()) MFS_HIP(hipMemcpyAsync(S->d_stet, stet.data(), 4 * stet.size(), hipMemcpyHostToDevice, s));
    if (!snode.empty())
      MFS_HIP(hipMemcpyAsync(S->d_snode, snode.data(), 16 * snode.size(), hipMemcpyHostToDevice, s));
    MFS_HIP(hipMemcpyAsync(S->d_slice_fix, slice_fix.data(), slice_fix.size(), hipMemcpyHostToDevice, s));
  }
  MFS_HIP(hipMemcpyAsync(S->d_browptr, browptr.data(), 4 * browptr.size(), hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(S->d_fixed, fixed.data(), ndof, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(S->d_g, g.data(), 8 * ndof, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(S->d_fbase, fbase.data(), 8 * ndof, hipMemcpyHostToDevice, s));
  MFS_HIP(hipStreamSynchronize(s));
  *out = S;
  return MFS_OK;
}

static int assemble_fused(mfs_asm S, double nu) {
  hipStream_t s = S->ctx->stream;
  mfs_mat A = S->A;
  const int CL = S->chunk_rows;
  const unsigned grid = (unsigned)(A->sell.n_slices * (64 / CL));
  const size_t lds = S->fused_lds;
#define MFS_ASM_CHUNK(BS, CONV, CLV)                                                                                \
  do {                                                                                                              \
    constexpr int NTH = kChunkThreads(CLV);                                                                         \
    auto kern = k_assemble_slice<BS, CONV, CLV, NTH>;                                                               \
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);             \
    kern<<<grid, NTH, lds, s>>>(                                                                                    \
        S->d_xyz, S->d_tets, S->d_coef, S->problem, nu, S->d_sptr, S->d_stet, S->d_snode, S->d_goff, S->d_gil16,   \
        A->d_slice_base, A->d_slice_width, S->d_browptr, A->d_col, S->d_slice_fix, S->d_fixed, S->d_g, S->d_fbase,  \
        A->d_val64, A->d_val32, A->d_dinv64, A->d_dinv32, (double*)S->b->d, (double*)S->x0->d, A->nb);             \
  } while (0)
#define MFS_ASM_SLICE(BS, CONV)      \
  do {                               \
    if (CL == 16)                    \
      MFS_ASM_CHUNK(BS, CONV, 16);   \
    else if (CL == 32)               \
      MFS_ASM_CHUNK(BS, CONV, 32);   \
    else                             \
      MFS_ASM_CHUNK(BS, CONV, 64);   \
  } while (0)
  if (S->bs == 1)
    MFS_ASM_SLICE(1, MFS_ELEM_REFERENCE);
  else if (S->convention == MFS_ELEM_TEXTBOOK)
    MFS_ASM_SLICE(3, MFS_ELEM_TEXTBOOK);
  else
    MFS_ASM_SLICE(3, MFS_ELEM_REFERENCE);
#undef MFS_ASM_SLICE
#undef MFS_ASM_CHUNK
  MFS_HIP(hipGetLastError());
  A->val32_valid = true;
  return MFS_OK;
}

static int assemble_kernels(mfs_asm S, double nu) {
  if (S->fused) return assemble_fused(S, nu);
  hipStream_t s = S->ctx->stream;
  mfs_mat A = S->A;
  k_tet_geometry<<<grid_cap(S->nt), 256, 0, s>>>(S->d_xyz, S->d_tets, S->d_coef, S->nt, S->problem, nu, S->d_tet);
  const int64_t ne = A->sell_entries;
  if (S->bs == 1)
    k_gather<1, MFS_ELEM_REFERENCE><<<grid_cap(ne), 256, 0, s>>>(S->d_goff, S->d_gil, S->d_tet, ne, A->d_val64);
  else if (S->convention == MFS_ELEM_TEXTBOOK)
    k_gather<3, MFS_ELEM_TEXTBOOK><<<grid_cap(ne), 256, 0, s>>>(S->d_goff, S->d_gil, S->d_tet, ne, A->d_val64);
  else
    k_gather<3, MFS_ELEM_REFERENCE><<<grid_cap(ne), 256, 0, s>>>(S->d_goff, S->d_gil, S->d_tet, ne, A->d_val64);
  if (S->bs == 1)
    k_dirichlet<1><<<grid_cap(A->nb), 256, 0, s>>>(S->d_browptr, A->d_col, A->d_slice_base, A->d_val64, S->d_fixed,
                                                   S->d_g, S->d_fbase, (double*)S->b->d, (double*)S->x0->d, A->nb);
  else
    k_dirichlet<3><<<grid_cap(A->nb), 256, 0, s>>>(S->d_browptr, A->d_col, A->d_slice_base, A->d_val64, S->d_fixed,
                                                   S->d_g, S->d_fbase, (double*)S->b->d, (double*)S->x0->d, A->nb);
  MFS_HIP(hipGetLastError());
  return mat_refresh_derived(A);
}

extern "C" int mfs_assemble_run(mfs_asm S, const double* coef, double nu) {
  MFS_CHECK_ARG(S && coef, "mfs_assemble_run: bad argument");
  MFS_CHECK_ARG(S->problem == MFS_PROBLEM_SCALAR || (nu > -1.0 && nu < 0.5), "mfs_assemble_run: nu out of range");
  hipStream_t s = S->ctx->stream;
  MFS_HIP(hipMemcpyAsync(S->d_coef, coef, 8 * S->nt, hipMemcpyHostToDevice, s));
  MFS_TRY(assemble_kernels(S, nu));
  MFS_HIP(hipStreamSynchronize(s));
  S->A->val32_valid = true;
  return MFS_OK;
}

extern "C" int mfs_assemble_time(mfs_asm S, int reps, double* avg_ms) {
  MFS_CHECK_ARG(S && reps > 0 && avg_ms, "mfs_assemble_time: bad argument");
  hipStream_t s = S->ctx->stream;
  // nu of the last run is not stored; timing uses 0.3 (cost is independent)
  MFS_TRY(assemble_kernels(S, 0.3));
  hipEvent_t e0, e1;
  MFS_HIP(hipEventCreate(&e0));
  MFS_HIP(hipEventCreate(&e1));
  MFS_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) MFS_TRY(assemble_kernels(S, 0.3));
  MFS_HIP(hipEventRecord(e1, s));
  MFS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = ms / reps;
  return MFS_OK;
}

extern "C" int mfs_assemble_matrix(mfs_asm S, mfs_mat* A) {
  MFS_CHECK_ARG(S && A, "mfs_assemble_matrix: bad argument");
  *A = S->A;
  return MFS_OK;
}
extern "C" int mfs_assemble_rhs(mfs_asm S, mfs_vec* b) {
  MFS_CHECK_ARG(S && b, "mfs_assemble_rhs: bad argument");
  *b = S->b;
  return MFS_OK;
}
extern "C" int mfs_assemble_x0(mfs_asm S, mfs_vec* x0) {
  MFS_CHECK_ARG(S && x0, "mfs_assemble_x0: bad argument");
  *x0 = S->x0;
  return MFS_OK;
}
extern "C" int mfs_assemble_node_map(mfs_asm S, int32_t* map) {
  MFS_CHECK_ARG(S && map, "mfs_assemble_node_map: bad argument");
  std::memcpy(map, S->map.data(), 4 * S->nn);
  return MFS_OK;
}

// MsFEM local basis problems, batched on gfx950.
//
// Each coarse cell is a box of n^3 fine hexes (6 Freudenthal tets each) with a
// per-tet Young's modulus.  For each of its 8 coarse nodes c and 3 directions
// d the local problem is: K u = 0 inside the cell, u = hat_c * e_d on the
// cell boundary (hat_c = trilinear coarse hat).  This replaces the per-cell
// volume solves of Grid_forMech_Poly_Order1::CreateFunctionalSpaces
// (MsFEM/MsFEM_Grid.h:4795-5760 via FEM/FEM.h:2969-3795) in the
// three-direction formulation.
//
// Layout (device):
//  * cell matrix, interior rows only, in the structured 15-point stencil of
//    the Freudenthal mesh: value (s, e) of interior node v at
//    K[cell_base*135 + (s*9 + e)*nint + v] (structure-of-arrays, no column
//    indices).  Blocks that couple to boundary nodes are zero here; their
//    effect is in the right-hand side b = -K_IB g.
//  * b, x: FP64 [interior node][3][24] per cell.
//  * one task = (cell, group of 8 bases); a persistent kernel pulls tasks
//    from an atomic queue (largest cells first) and runs FP32 Jacobi-PCG with
//    FP64 refinement for its 8 right-hand sides, search directions in LDS.
//
// Reference-trace mode (MFS_TRACE_REFERENCE) replaces the boundary data
// hat_c * e_d by the reference's edge -> face hierarchy (MsFEM_Grid.h:
// 5026-5506): per coarse edge a 1-D elastic solve with the coarse-node
// value at one end and 0 at the other, per coarse face and face corner a
// 2-D elastic solve with the two edge solutions that are 1 at that corner,
// then the volume solve with the face solutions of the three faces meeting
// at the node.  Basis 3c+d is driven with e_d; the reference drives (1,1,1)
// and its basis of node c is the sum over d (each step is linear).
//  * edge profiles ET[cell.et0 + (edge*(n+1) + t)*2 + end]: the 1-D chain
//    solution is e_d times a scalar profile (axis-aligned segments give
//    diagonal 3x3 blocks, so the components decouple);
//  * face traces FT[cell.ft0 + ((face*(n+1)^2 + u + (n+1)v)*3 + a)*12 +
//    3*lc + d] for the 4 face corners lc and 3 drives d, interior face nodes
//    from a batched FP64 Jacobi-PCG (one workgroup per face, 12 right-hand
//    sides), face-boundary nodes copied from the edge profiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "mfs_device.hpp"

using namespace mfs;

namespace {

constexpr int NS = 15;   // stencil points
constexpr int RG = 8;    // right-hand sides per task
// Row stride (values) of p in LDS, and of d when it lives there: the 3 x RG
// values of a row plus padding.  A wave's lanes own consecutive rows, so its
// reads of one component address consecutive rows; with packed rows (96 B
// FP32, 192 B FP64) they had 2- / 4-way bank conflicts.  FP64 and the FP32
// register variants use 16-byte reads: 208 B and 112 B rows put any 8 (or
// 16) consecutive rows in distinct banks.  The FP32 slab variants use 104 B
// rows (8-byte reads, distinct banks for 16 consecutive rows), which keeps
// two workgroups per CU on cells of up to 729 rows.
template <typename T, bool REG>
constexpr int lds_row() { return sizeof(T) == 8 ? 3 * RG + 2 : REG ? 3 * RG + 4 : 3 * RG + 2; }
template <typename T, bool REG>
__device__ __forceinline__ int64_t lds_at(int v, int a, int r) { return (int64_t)v * lds_row<T, REG>() + a * RG + r; }
constexpr int NB = 24;   // bases per cell
constexpr int MAXN = 12; // largest supported cell subdivision (FP32 inner; FP64 inner: 10)
// Kernel variants per precision, picked by the autotuner: the number of
// rows a thread carries through the inner-iteration phases at once (more
// rows keep more loads in flight per wave, fewer rows leave registers for
// two waves per SIMD), and whether the per-row vectors r, d, q live in
// registers (variants 3 and 4; only for batches whose cells all have
// nint <= rows * workgroup size) or in a global scratch slab.  (Measured
// and dropped: K rings of 5 stencil points, part of each row's K kept in
// registers for the whole task, and forcing three or four waves per SIMD;
// each costs more in spills than it hides in latency.)
constexpr int kRptF32[4] = {2, 3, 1, 2}, kRptF64[4] = {1, 2, 1, 2};
constexpr bool kRegVar[4] = {false, false, true, true};
constexpr int kVariants[2] = {4, 4};  // per precision: FP32 inner, FP64 inner
// Register-resident variants keep d (the inner correction) in LDS behind p,
// except the one-row FP32 variant, which has the registers for it.
constexpr bool d_in_lds(bool fp64, int rpt) { return fp64 || rpt != 1; }
// Largest workgroup of a variant: one register-resident row per thread
// needs a thread per interior node, so that form takes up to 512 threads.
constexpr int max_threads(bool reg, int rpt) { return reg && rpt == 1 ? 512 : 256; }
// Waves per SIMD a variant is compiled for: two where its registers allow
// it (the compiler then spills cold code instead of dropping to one).
constexpr int min_waves(bool fp64, int rpt, bool reg) { return (reg && rpt == 1) || (!reg && rpt <= (fp64 ? 1 : 2)) ? 2 : 1; }

// Stencil offsets (di, dj, dk), sorted by linear offset (dk, dj, di lexicographic)
__constant__ int c_stencil[NS][3];
static const int h_stencil[NS][3] = {{-1, -1, -1}, {0, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {-1, -1, 0},
                                     {0, -1, 0},   {-1, 0, 0},  {0, 0, 0},   {1, 0, 0},   {0, 1, 0},
                                     {1, 1, 0},    {0, 0, 1},   {1, 0, 1},   {0, 1, 1},   {1, 1, 1}};
constexpr int S_CENTER = 7;
// the same table as a compile-time constant for the unrolled device loops
__device__ constexpr int kStencil[NS][3] = {{-1, -1, -1}, {0, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {-1, -1, 0},
                                            {0, -1, 0},   {-1, 0, 0},  {0, 0, 0},   {1, 0, 0},   {0, 1, 0},
                                            {1, 1, 0},    {0, 0, 1},   {1, 0, 1},   {0, 1, 1},   {1, 1, 1}};

struct Cell {
  int n;        // fine hexes per edge
  int nint;     // interior nodes (n-1)^3
  int64_t tet0; // first tet in E
  int64_t int0; // first interior node (global numbering over cells)
  int tmpl;     // template id
  double ox, oy, oz;
  int64_t et0;  // reference-trace mode: first edge-profile entry (12*(n+1)*2 per cell)
  int64_t ft0;  // reference-trace mode: first face-trace entry (6*(n+1)^2*36 per cell)
};

// Coarse faces f = 2*ax + side (side 1: coordinate n); face coordinates
// (u, v) run along the two other axes in increasing order.  Coarse edges
// e = 4*dir + b1 + 2*b2, b1 / b2 the sides on the lower / higher other axis.
__host__ __device__ __forceinline__ void face_axes(int ax, int& A1, int& A2) {
  A1 = ax == 0 ? 1 : 0;
  A2 = ax == 2 ? 1 : 2;
}
__host__ __device__ __forceinline__ int edge_id(int dir, int axa, int sa, int axb, int sb) {
  // sides sa on axis axa and sb on axis axb (the two axes other than dir)
  return axa < axb ? 4 * dir + sa + 2 * sb : 4 * dir + sb + 2 * sa;
}
constexpr int FT_NODE = 36;  // face-trace values per face node: [a][lc][d]

// Boundary value (3 components) of basis q = 3c + d at boundary node
// (i, j, k) of an n-cell from its face traces: any face holding both the
// node and corner c (they agree on shared edges), 0 when none does.
__host__ __device__ __forceinline__ void face_trace_at(const double* ft_cell, int n, int i, int j, int k, int q,
                                                       double (&g)[3]) {
  const int c = q / 3, d = q % 3;
  const int idx[3] = {i, j, k};
  g[0] = g[1] = g[2] = 0.0;
  for (int ax = 0; ax < 3; ax++) {
    const int side = idx[ax] == 0 ? 0 : (idx[ax] == n ? 1 : -1);
    if (side < 0 || ((c >> ax) & 1) != side) continue;
    int A1, A2;
    face_axes(ax, A1, A2);
    const int lc = ((c >> A1) & 1) + 2 * ((c >> A2) & 1);
    const int64_t w = idx[A1] + (int64_t)(n + 1) * idx[A2];
    const double* base = ft_cell + (((int64_t)(2 * ax + side) * (n + 1) * (n + 1) + w) * FT_NODE) + 3 * lc + d;
    for (int a = 0; a < 3; a++) g[a] = base[a * 12];
    return;
  }
}

// Q1 boundary value hat_c(x) e_d.
__host__ __device__ __forceinline__ void hat_trace_at(int n, int i, int j, int k, int q, double (&g)[3]) {
  const int c = q / 3, d = q % 3;
  const double X = (double)i / n, Y = (double)j / n, Z = (double)k / n;
  const double hat = ((c & 1) ? X : 1 - X) * ((c & 2) ? Y : 1 - Y) * ((c & 4) ? Z : 1 - Z);
  g[0] = g[1] = g[2] = 0.0;
  g[d] = hat;
}

struct Tmpl {  // per distinct n: gather lists of interior rows x stencil
  int n;
  int64_t ptr0;  // offset into g_ptr (nint*NS+1 entries)
  int64_t ent0;  // offset into g_ent
};

// Task queues: one per XCD.  Workgroups are dispatched round-robin over the
// 8 XCDs, so workgroup b pulls from queue b % 8 first; all three tasks of a
// cell sit in the same queue, so the cell matrix they stream stays in one
// XCD's L2.  An exhausted queue sends the workgroup on to the next one.
constexpr int NQ = 8;
constexpr int QSTRIDE = 32;  // counters 128 bytes apart

struct SolveArgs {
  const Cell* cells;
  const int* task_cell;
  const int* task_group;
  int ntasks;
  int seg[NQ + 1];   // tasks [seg[q], seg[q+1]) belong to queue q
  unsigned* queue;   // NQ counters, QSTRIDE apart
  const float* K32;
  const double* K64;
  const float* dinv32;   // [int0 + v][3]
  const double* dinv64;
  const double* b64;     // [int0 + v][3][24]
  double* x64;
  float* scr32;          // per-WG scratch: 3 vectors of maxint*3*RG floats
  double* scr64;         // per-WG scratch (FP64 inner mode)
  int64_t scr_stride;    // entries per vector
  int* iters;            // [cell][24]
  double* relres;        // [cell][24]
  int* flags;            // [cell][24]: bit0 not converged, bit1 guess rejected,
                         // bit2 FP32 inner solve failed, bit3 handed to the FP64 pass,
                         // bit4 FP64 solve from the guess failed and was redone from zero
  // FP32 launch: tasks with a failed right-hand side are appended to the
  // FP64 pass's list (count, then cells, then groups); FP64 pass: its task
  // list and count (written by the FP32 launch), queue 0 only
  int* fb_count;
  int* fb_cell;
  int* fb_group;
  const int* ntasks_dev;
  double tol, inner_tol;
  int maxit;
  int inner_cap;  // iterations one inner solve may take before the RHS is
                  // declared failed in this precision (FP32: handed to FP64)
  int use_guess;
  int max_nint;  // largest cell of the batch (rows of the LDS vectors)
  int stage_x;   // register variants: the residual pass stages the FP64 iterate in LDS
};

__device__ __forceinline__ int interior_index(int i, int j, int k, int m) {
  return (i - 1) + m * ((j - 1) + m * (k - 1));
}

// ---------------------------------------------------------------------------
// Cell assembly: one thread per (cell, interior node).  For each stencil
// block, sum the element contributions (template gather lists, ascending
// tet order) with gradients recomputed from the structured fine geometry.
// Interior-neighbour blocks go to K (FP64 and FP32); boundary-neighbour
// blocks update b = -K_IB g for the 24 bases.
// ---------------------------------------------------------------------------
__device__ void tet_nodes(int t, int n, int (&ijk)[4][3]) {
  const int hex = t / 6, p = t % 6;
  int c[3] = {hex % n, (hex / n) % n, hex / (n * n)};
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  ijk[0][0] = c[0];
  ijk[0][1] = c[1];
  ijk[0][2] = c[2];
  for (int s = 0; s < 3; s++) {
    c[perm[p][s]]++;
    ijk[s + 1][0] = c[0];
    ijk[s + 1][1] = c[1];
    ijk[s + 1][2] = c[2];
  }
}

__device__ void tet_grads(const int (&ijk)[4][3], double hf, double (&g)[4][3], double& V) {
  double J[3][3];
  for (int r = 0; r < 3; r++)
    for (int a = 0; a < 3; a++) J[r][a] = hf * (double)(ijk[r + 1][a] - ijk[0][a]);
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
  g[1][0] = c00 * inv; g[1][1] = c01 * inv; g[1][2] = c02 * inv;
  g[2][0] = c10 * inv; g[2][1] = c11 * inv; g[2][2] = c12 * inv;
  g[3][0] = c20 * inv; g[3][1] = c21 * inv; g[3][2] = c22 * inv;
  for (int a = 0; a < 3; a++) g[0][a] = -(g[1][a] + g[2][a] + g[3][a]);
  V = fabs(det) / 6.0;
}

// Step index of axis a in the tet's axis order (xyz,xzy,yxz,yzx,zxy,zyx).
__device__ constexpr int kStepOf[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {2, 0, 1}, {1, 2, 0}, {2, 1, 0}};

// Elastic block (I=i, J=j) of a tet: textbook V(lam dI_a dJ_b + mu dI_b dJ_a +
// mu delta dI.dJ); the reference convention is its transpose.
__device__ __forceinline__ void elastic_block(const double* dI, const double* dJ, double vl, double vm, int conv,
                                              double (&K)[3][3]) {
  const double dd = dI[0] * dJ[0] + dI[1] * dJ[1] + dI[2] * dJ[2];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double v = conv == MFS_ELEM_TEXTBOOK ? vl * dI[a] * dJ[b] + vm * dI[b] * dJ[a]
                                           : vl * dI[b] * dJ[a] + vm * dI[a] * dJ[b];
      if (a == b) v += vm * dd;
      K[a][b] += v;
    }
}

__global__ __launch_bounds__(256) void k_cell_assemble(const Cell* __restrict__ cells, int ncells,
                                                       const int64_t* __restrict__ node_cell_ptr,
                                                       const Tmpl* __restrict__ tmpls,
                                                       const int64_t* __restrict__ g_ptr,
                                                       const uint32_t* __restrict__ g_ent,
                                                       const double* __restrict__ E, double nu, double h, int conv,
                                                       const double* __restrict__ ft, int64_t total_int,
                                                       float* __restrict__ K32,
                                                       double* __restrict__ K64, float* __restrict__ dinv32,
                                                       double* __restrict__ dinv64, double* __restrict__ b64) {
  for (int64_t gv = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; gv < total_int;
       gv += (int64_t)gridDim.x * blockDim.x) {
    // find the cell (binary search over cumulative interior counts)
    int lo = 0, hi = ncells - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (node_cell_ptr[mid] <= gv) lo = mid;
      else hi = mid - 1;
    }
    const Cell C = cells[lo];
    const Tmpl T = tmpls[C.tmpl];
    const int n = C.n, m = n - 1, nint = C.nint;
    const int v = (int)(gv - C.int0);
    const int i = v % m + 1, j = (v / m) % m + 1, k = v / (m * m) + 1;
    const double hf = h / n;
    const double lamf = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), muf = 1.0 / (2.0 * (1.0 + nu));
    double bacc[3][NB];
    for (int a = 0; a < 3; a++)
      for (int q = 0; q < NB; q++) bacc[a][q] = 0.0;
    float* K32c = K32 + C.int0 * (NS * 9);
    double* K64c = K64 + C.int0 * (NS * 9);
    for (int s = 0; s < NS; s++) {
      const int ni = i + c_stencil[s][0], nj = j + c_stencil[s][1], nk = k + c_stencil[s][2];
      double Kb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      const int64_t slot = (int64_t)v * NS + s;
      for (int64_t q = g_ptr[T.ptr0 + slot]; q < g_ptr[T.ptr0 + slot + 1]; q++) {
        const uint32_t ce = g_ent[T.ent0 + q];
        const int t = (int)(ce >> 4), li = (ce >> 2) & 3, lj = ce & 3;
        int ijk[4][3];
        tet_nodes(t, n, ijk);
        double g[4][3], V;
        tet_grads(ijk, hf, g, V);
        const double Et = E[C.tet0 + t];
        elastic_block(g[li], g[lj], V * Et * lamf, V * Et * muf, conv, Kb);
      }
      const bool interior = ni >= 1 && ni <= m && nj >= 1 && nj <= m && nk >= 1 && nk <= m;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          const int64_t at = (int64_t)(s * 9 + 3 * a + b) * nint + v;
          K64c[at] = interior ? Kb[a][b] : 0.0;
          K32c[at] = interior ? (float)Kb[a][b] : 0.0f;
        }
      if (!interior) {
        if (ft == nullptr) {
          // g for basis (c, d) at this boundary node: hat_c(x) e_d
          const double X = (double)ni / n, Y = (double)nj / n, Z = (double)nk / n;
          for (int c = 0; c < 8; c++) {
            const double hat = ((c & 1) ? X : 1 - X) * ((c & 2) ? Y : 1 - Y) * ((c & 4) ? Z : 1 - Z);
            if (hat == 0.0) continue;
            for (int d = 0; d < 3; d++)
              for (int a = 0; a < 3; a++) bacc[a][3 * c + d] -= Kb[a][d] * hat;
          }
        } else {
          // g from the face traces (reference-trace mode): full vectors
#pragma unroll
          for (int q = 0; q < NB; q++) {
            double g[3];
            face_trace_at(ft + C.ft0, n, ni, nj, nk, q, g);
            for (int a = 0; a < 3; a++) bacc[a][q] -= Kb[a][0] * g[0] + Kb[a][1] * g[1] + Kb[a][2] * g[2];
          }
        }
      }
      if (s == S_CENTER)
        for (int a = 0; a < 3; a++) {
          dinv64[(C.int0 + v) * 3 + a] = 1.0 / Kb[a][a];
          dinv32[(C.int0 + v) * 3 + a] = (float)(1.0 / Kb[a][a]);
        }
    }
    double* bv = b64 + (C.int0 + v) * (3 * NB);
    for (int a = 0; a < 3; a++)
      for (int q = 0; q < NB; q++) bv[a * NB + q] = bacc[a][q];
  }
}

// ---------------------------------------------------------------------------
// Stencil SpMM for one interior node and RG right-hand sides:
// acc[a][r] = sum_s sum_b K[s][a][b] * P(nbr_s)[b][r]
// ---------------------------------------------------------------------------
// Stencil offsets packed two bits per point ((d + 1) << 2s) for the loop below.
constexpr uint32_t pack_offsets(int dim) {
  uint32_t w = 0;
  for (int s = 0; s < NS; s++) w |= (uint32_t)(kStencil[s][dim] + 1) << (2 * s);
  return w;
}
constexpr uint32_t kOffI = pack_offsets(0), kOffJ = pack_offsets(1), kOffK = pack_offsets(2);

// Loads through a raw buffer resource (gfx9 descriptor word 3; the range
// check returns 0 beyond num_records).
constexpr int kBufferRsrcWord3 = 0x00020000;
template <typename KT>
__device__ __forceinline__ KT buffer_load(__amdgpu_buffer_rsrc_t rs, int voff, int soff);
template <>
__device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}
template <>
__device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
}

// Row offset of stencil point s in the cell's interior numbering
// (x fastest, m interior nodes per edge): di + m (dj + m dk).
__device__ __forceinline__ int stencil_offset(int s, int m) {
  const int di = (int)((kOffI >> (2 * s)) & 3u) - 1;
  const int dj = (int)((kOffJ >> (2 * s)) & 3u) - 1;
  const int dk = (int)((kOffK >> (2 * s)) & 3u) - 1;
  return di + m * (dj + m * dk);
}
// The row whose p (or x) a stencil point reads.  An interior neighbour is
// row v + off.  A neighbour on the cell boundary has a zero K block (the
// cell assembly stores 0 there), so any row of the cell may be read for it:
// the sum clamped to the cell's rows is one instruction instead of the
// neighbour's coordinates, its bounds test and a select.
__device__ __forceinline__ int neighbour_row(int v, int off, int nint) { return min(max(v + off, 0), nint - 1); }

// The same product for RPT rows of one thread at once (rows v[k]; rows with
// ok[k] false are computed on row 0 and discarded by the caller).  The K
// values of all RPT rows for the next stencil point are loaded while the
// current point is applied, so a wave keeps RPT*9 loads in flight and the
// row passes of one thread do not queue behind each other.
template <typename KT, typename VT, int RPT, typename Load>
__device__ __forceinline__ void stencil_rows(const KT* __restrict__ Kc, int nint, const int (&v)[RPT],
                                             const bool (&ok)[RPT], int m, Load load, VT (&acc)[RPT][3][RG]) {
  int vi[RPT];
  const KT* kp[RPT];
  KT kn[RPT][9];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    vi[q] = ok[q] ? v[q] : 0;
    kp[q] = Kc + vi[q];
#pragma unroll
    for (int e = 0; e < 9; e++) kn[q][e] = kp[q][(int64_t)e * nint];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int r = 0; r < RG; r++) acc[q][a][r] = 0;
  }
#pragma unroll 1
  for (int s = 0; s < NS; s++) {
    VT kv[RPT][9];
#pragma unroll
    for (int q = 0; q < RPT; q++)
#pragma unroll
      for (int e = 0; e < 9; e++) kv[q][e] = (VT)kn[q][e];
    if (s + 1 < NS) {
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        kp[q] += (int64_t)9 * nint;
#pragma unroll
        for (int e = 0; e < 9; e++) kn[q][e] = kp[q][(int64_t)e * nint];
      }
    }
    const int off = stencil_offset(s, m);
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int u = neighbour_row(vi[q], off, nint);
#pragma unroll
      for (int b = 0; b < 3; b++) {
        VT pv[RG];
        load(u, b, pv);
#pragma unroll
        for (int r = 0; r < RG; r++) {
          acc[q][0][r] += kv[q][0 * 3 + b] * pv[r];
          acc[q][1][r] += kv[q][1 * 3 + b] * pv[r];
          acc[q][2][r] += kv[q][2 * 3 + b] * pv[r];
        }
      }
    }
  }
}

// stencil_rows for the slab variants: the same products in the same order,
// with two K register sets used in turn (the next point's K is requested
// into one while the other is applied, two points per loop trip), so no set
// is copied from one point to the next.
template <typename KT, typename VT, int RPT, typename Load>
__device__ __forceinline__ void stencil_rows_pp(const KT* __restrict__ Kc, int nint, const int (&v)[RPT],
                                                const bool (&ok)[RPT], int m, Load load, VT (&acc)[RPT][3][RG]) {
  int vi[RPT];
  const KT* kp[RPT];
  KT ka[RPT][9], kb[RPT][9];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    vi[q] = ok[q] ? v[q] : 0;
    kp[q] = Kc + vi[q];
#pragma unroll
    for (int e = 0; e < 9; e++) ka[q][e] = kp[q][(int64_t)e * nint];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int r = 0; r < RG; r++) acc[q][a][r] = 0;
  }
  auto request = [&](KT (&kn)[RPT][9]) {
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      kp[q] += (int64_t)9 * nint;
#pragma unroll
      for (int e = 0; e < 9; e++) kn[q][e] = kp[q][(int64_t)e * nint];
    }
  };
  auto apply = [&](int s, const KT (&kv)[RPT][9]) {
    const int off = stencil_offset(s, m);
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int u = neighbour_row(vi[q], off, nint);
#pragma unroll
      for (int b = 0; b < 3; b++) {
        VT pv[RG];
        load(u, b, pv);
#pragma unroll
        for (int r = 0; r < RG; r++) {
          acc[q][0][r] += (VT)kv[q][0 * 3 + b] * pv[r];
          acc[q][1][r] += (VT)kv[q][1 * 3 + b] * pv[r];
          acc[q][2][r] += (VT)kv[q][2 * 3 + b] * pv[r];
        }
      }
    }
  };
  static_assert(NS % 2 == 1, "the last stencil point is applied after the paired loop");
#pragma unroll 1
  for (int s = 0; s + 1 < NS; s += 2) {
    request(kb);  // point s + 1
    apply(s, ka);
    request(ka);  // point s + 2 (NS is odd, so it exists)
    apply(s + 1, kb);
  }
  apply(NS - 1, ka);
}

// Inner-iteration vector phases over RPT rows of one thread at a time
// (separate restrict-qualified arrays, so the loads of the RPT rows are
// issued together).
// q = K p into Q, partial p.q
template <typename T, int RPT>
__device__ __forceinline__ void phase_spmv(const float* __restrict__ K32, const double* __restrict__ K64, int nint,
                                           int m, const T* P, T* __restrict__ Q, double (&w)[RG]) {
  constexpr bool fp64 = std::is_same<T, double>::value;
  const int tid = threadIdx.x, nth = blockDim.x;
  auto ld = [&](int u, int b, T (&pv)[RG]) {
#pragma unroll
    for (int r = 0; r < RG; r++) pv[r] = P[lds_at<T, false>(u, b, r)];
  };
  constexpr int RP = RPT > 0 ? RPT : 1;
  for (int v0 = tid; v0 < nint; v0 += RP * nth) {
    int vr[RP];
    bool ok[RP];
#pragma unroll
    for (int q = 0; q < RP; q++) {
      vr[q] = v0 + q * nth;
      ok[q] = vr[q] < nint;
    }
    T acc[RP][3][RG];
    if constexpr (fp64)
      stencil_rows_pp<double, double, RP>(K64, nint, vr, ok, m, ld, acc);
    else
      stencil_rows_pp<float, float, RP>(K32, nint, vr, ok, m, ld, acc);
#pragma unroll
    for (int q = 0; q < RP; q++) {
      if (!ok[q]) continue;
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int r = 0; r < RG; r++) {
          const int64_t e = ((int64_t)vr[q] * 3 + a) * RG + r;
          Q[e] = acc[q][a][r];
          w[r] += (double)(acc[q][a][r] * P[lds_at<T, false>(vr[q], a, r)]);
        }
    }
  }
}

// D += alpha p, R -= alpha q, z = dinv R; partials r.z, r.r
template <typename T, int RPT>
__device__ __forceinline__ void phase_update(int nint, const T* __restrict__ dinv, const T* P, const T* __restrict__ Q,
                                             T* __restrict__ R, T* __restrict__ D, const T (&alpha)[RG],
                                             double (&w)[2 * RG]) {
  const int tid = threadIdx.x, nth = blockDim.x;
  constexpr int RP = RPT > 0 ? RPT : 1;
  for (int v0 = tid; v0 < nint; v0 += RP * nth) {
#pragma unroll
    for (int q = 0; q < RP; q++) {
      const int v = v0 + q * nth;
      if (v >= nint) continue;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const T dv = dinv[v * 3 + a];
#pragma unroll
        for (int r = 0; r < RG; r++) {
          const int64_t e = ((int64_t)v * 3 + a) * RG + r;
          const T pe = P[lds_at<T, false>(v, a, r)];
          D[e] += alpha[r] * pe;
          const T rv = R[e] - alpha[r] * Q[e];
          R[e] = rv;
          const T z = dv * rv;
          w[r] += (double)(rv * z);
          w[RG + r] += (double)(rv * rv);
        }
      }
    }
  }
}

// p = z + beta p for the right-hand sides still iterating (mask `on`)
template <typename T, int RPT>
__device__ __forceinline__ void phase_direction(int nint, const T* __restrict__ dinv, T* P, const T* __restrict__ R,
                                                const T (&beta)[RG], int on) {
  const int tid = threadIdx.x, nth = blockDim.x;
  constexpr int RP = RPT > 0 ? RPT : 1;
  for (int v0 = tid; v0 < nint; v0 += RP * nth) {
#pragma unroll
    for (int q = 0; q < RP; q++) {
      const int v = v0 + q * nth;
      if (v >= nint) continue;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const T dv = dinv[v * 3 + a];
#pragma unroll
        for (int r = 0; r < RG; r++) {
          if (!((on >> r) & 1)) continue;
          const int64_t e = ((int64_t)v * 3 + a) * RG + r, ep = lds_at<T, false>(v, a, r);
          P[ep] = dv * R[e] + beta[r] * P[ep];
        }
      }
    }
  }
}

// K of one cell through a buffer resource: the row's byte offset is the
// (per-thread) VGPR offset and the entry's byte offset (s*9+e)*nint*sizeof(KT)
// the (uniform) SGPR offset, so a load costs no vector address arithmetic.
// The stride is redefined for the compiler on every call, so the 135 entry
// offsets are formed where they are used instead of being hoisted out of
// the iteration loop.
template <typename KT, int RPT>
struct KRows {
  __amdgpu_buffer_rsrc_t rs;
  int vo[RPT];
  int stride;
  __device__ __forceinline__ KRows(const KT* Kc, int nint, const int (&vi)[RPT]) {
    rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<KT*>(Kc), (short)0, NS * 9 * nint * (int)sizeof(KT),
                                           kBufferRsrcWord3);
#pragma unroll
    for (int q = 0; q < RPT; q++) vo[q] = vi[q] * (int)sizeof(KT);
    stride = nint * (int)sizeof(KT);
    asm volatile("" : "+s"(stride));
  }
  __device__ __forceinline__ KT at(int q, int s, int e) const { return buffer_load<KT>(rs, vo[q], (s * 9 + e) * stride); }
};

template <int RPT>
__device__ __forceinline__ void clamp_rows(const int (&v)[RPT], const bool (&ok)[RPT], int (&vi)[RPT]) {
#pragma unroll
  for (int q = 0; q < RPT; q++) vi[q] = ok[q] ? v[q] : 0;
}

// Requests the K values of stencil points 0 .. PF-1 into the ring that
// stencil_rows_pf consumes.
template <typename KT, int RPT, int PF>
__device__ __forceinline__ void k_ring_request(const KT* __restrict__ Kc, int nint, const int (&v)[RPT],
                                               const bool (&ok)[RPT], KT (&ring)[PF][RPT][9]) {
  int vi[RPT];
  clamp_rows<RPT>(v, ok, vi);
  const KRows<KT, RPT> kr(Kc, nint, vi);
#pragma unroll
  for (int p = 0; p < PF; p++)
#pragma unroll
    for (int q = 0; q < RPT; q++)
#pragma unroll
      for (int e = 0; e < 9; e++) ring[p][q][e] = kr.at(q, p, e);
}

// stencil_rows with the K values of PF stencil points in flight: the ring
// holds points 0 .. PF-1 on entry (k_ring_request, or the previous call),
// and point (s + PF) mod 15 is requested right after point s is applied,
// into the slot point s used (PF x RPT x 9 registers).  The ring thus wraps:
// the last points request the first ones for the next product with the same
// K, which arrive while the caller runs its update and reductions.  The
// point loop is unrolled, so the ring is indexed statically and needs no
// copies.  A wave then waits for a K load once per PF points instead of
// once per point; PF x RPT x 9 stays below the 63 vector loads a wave can
// have outstanding.  Same arithmetic, in the same order, as stencil_rows.
template <typename KT, typename VT, int RPT, int PF, typename Load>
__device__ __forceinline__ void stencil_rows_pf(const KT* __restrict__ Kc, int nint, const int (&v)[RPT],
                                                const bool (&ok)[RPT], int m, Load load, KT (&ring)[PF][RPT][9],
                                                VT (&acc)[RPT][3][RG]) {
  static_assert(PF >= 1 && NS % PF == 0 && PF * RPT * 9 <= 63, "prefetch ring: PF must divide 15, <= 63 loads");
  int vi[RPT];
  clamp_rows<RPT>(v, ok, vi);
#pragma unroll
  for (int q = 0; q < RPT; q++)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int r = 0; r < RG; r++) acc[q][a][r] = 0;
  const KRows<KT, RPT> kr(Kc, nint, vi);
#pragma unroll
  for (int s0 = 0; s0 < NS; s0 += PF) {
#pragma unroll
    for (int p = 0; p < PF; p++) {
      const int s = s0 + p;
      const int off = stencil_offset(s, m);
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int u = neighbour_row(vi[q], off, nint);
#pragma unroll
        for (int b = 0; b < 3; b++) {
          VT pv[RG];
          load(u, b, pv);
#pragma unroll
          for (int r = 0; r < RG; r++) {
            acc[q][0][r] += (VT)ring[p][q][0 * 3 + b] * pv[r];
            acc[q][1][r] += (VT)ring[p][q][1 * 3 + b] * pv[r];
            acc[q][2][r] += (VT)ring[p][q][2 * 3 + b] * pv[r];
          }
        }
      }
      // the slot just used is refilled with the point PF later, modulo 15
      // (requesting it after the point's products needs no copy of the slot)
      const int sn = (s + PF) % NS;
#pragma unroll
      for (int q = 0; q < RPT; q++)
#pragma unroll
        for (int e = 0; e < 9; e++) ring[p][q][e] = kr.at(q, sn, e);
    }
  }
}

// Register-resident forms of the three phases: thread t owns rows
// t + q * blockDim.x (q < RPT) for the whole task, and keeps their r, q and
// diagonal inverses in registers (d in registers or in LDS behind p), so an
// inner iteration reads only K from memory (p stays in LDS: the stencil
// reads the neighbours' entries).  Rows at or beyond nint (ok false) carry
// values that are never used.
template <typename T, int RPT, int PF>
__device__ __forceinline__ void phase_spmv_reg(const float* __restrict__ K32, const double* __restrict__ K64, int nint,
                                               int m, const T* P, T (&ring)[PF][RPT][9], T (&Qr)[RPT][3][RG],
                                               double (&w)[RG]) {
  constexpr bool fp64 = std::is_same<T, double>::value;
  auto ld = [&](int u, int b, T (&pv)[RG]) {
#pragma unroll
    for (int r = 0; r < RG; r++) pv[r] = P[lds_at<T, true>(u, b, r)];
  };
  int vr[RPT];
  bool ok[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    vr[q] = threadIdx.x + q * blockDim.x;
    ok[q] = vr[q] < nint;
  }
  if constexpr (fp64)
    stencil_rows_pf<double, double, RPT, PF>(K64, nint, vr, ok, m, ld, ring, Qr);
  else
    stencil_rows_pf<float, float, RPT, PF>(K32, nint, vr, ok, m, ld, ring, Qr);
  // the row's three components are summed in T before the FP64 partial
  // (one conversion and FP64 add per right-hand side instead of three)
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    if (!ok[q]) continue;
#pragma unroll
    for (int r = 0; r < RG; r++) {
      T t = 0;
#pragma unroll
      for (int a = 0; a < 3; a++) t += Qr[q][a][r] * P[lds_at<T, true>(vr[q], a, r)];
      w[r] += (double)t;
    }
  }
}

template <typename T, int RPT, bool DLDS>
__device__ __forceinline__ void phase_update_reg(int nint, const T (&dv)[RPT][3], const T* P, const T (&Qr)[RPT][3][RG],
                                                 T (&Rr)[RPT][3][RG], T* DL, T (&Dr)[RPT][3][RG],
                                                 const T (&alpha)[RG], double (&w)[2 * RG]) {
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int v = threadIdx.x + q * blockDim.x;
    if (v >= nint) continue;
    T tz[RG], tr[RG];  // the row's three components summed in T first
#pragma unroll
    for (int r = 0; r < RG; r++) tz[r] = tr[r] = 0;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int r = 0; r < RG; r++) {
        const T pe = P[lds_at<T, true>(v, a, r)];
        if constexpr (DLDS)
          DL[lds_at<T, true>(v, a, r)] += alpha[r] * pe;
        else
          Dr[q][a][r] += alpha[r] * pe;
        const T rv = Rr[q][a][r] - alpha[r] * Qr[q][a][r];
        Rr[q][a][r] = rv;
        const T z = dv[q][a] * rv;
        tz[r] += rv * z;
        tr[r] += rv * rv;
      }
#pragma unroll
    for (int r = 0; r < RG; r++) {
      w[r] += (double)tz[r];
      w[RG + r] += (double)tr[r];
    }
  }
}

template <typename T, int RPT>
__device__ __forceinline__ void phase_direction_reg(int nint, const T (&dv)[RPT][3], T* P, const T (&Rr)[RPT][3][RG],
                                                    const T (&beta)[RG], int on) {
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int v = threadIdx.x + q * blockDim.x;
    if (v >= nint) continue;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int r = 0; r < RG; r++) {
        if (!((on >> r) & 1)) continue;
        const int64_t e = lds_at<T, true>(v, a, r);
        P[e] = dv[q][a] * Rr[q][a][r] + beta[r] * P[e];
      }
  }
}

// ---------------------------------------------------------------------------
// Workgroup all-reduce of N per-thread values (N a power of two, 8..32):
// a butterfly reduce-scatter inside the wave (log2 N + remaining steps of
// shuffles instead of N full wave reductions), one barrier, then lane k sums
// value k over the waves and every thread reads all N totals with readlane.
// The summation order is fixed, so every thread obtains identical totals.
// part: [waves][N] doubles in LDS; its previous contents must no longer be
// needed by any thread (a barrier separates two uses of one buffer).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Exchange with a partner lane that differs in bit log2(o) and agrees in the
// bits above it: DPP moves inside a row of 16 lanes for o <= 8 (quad
// permutes for 1 and 2, half-row and row mirrors for 4 and 8; a mirror
// also flips the lower bits, which the butterfly below does not rely on),
// ds_bpermute for 16 and 32.  Every pairing is an involution.
__device__ __forceinline__ int dpp_partner_i32(int v, int o) {
  switch (o) {
    case 1: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 2: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 4: return __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    default: return __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false); // row_mirror (o = 8)
  }
}
// lane l of every row of 16 gets lane l ^ 8's value (DPP row_ror:8)
__device__ __forceinline__ double row_ror8_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x128, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x128, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double partner_f64(double v, int o) {
  if (o >= 16) return __shfl_xor(v, o, 64);
  const int lo = dpp_partner_i32(__double2loint(v), o);
  const int hi = dpp_partner_i32(__double2hiint(v), o);
  return __hiloint2double(hi, lo);
}

// wg_reduce_to_lanes: the same sums, left in lane k (< N) of every wave
// (other lanes hold 0), for per-value work done one value per lane.
template <int N>
__device__ __forceinline__ double wg_reduce_to_lanes(double (&v)[N], double* part) {
  static_assert(N >= 2 && N <= 32 && (N & (N - 1)) == 0, "N must be a power of two in 2..32");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  constexpr int LOG = N == 2 ? 1 : N == 4 ? 2 : N == 8 ? 3 : N == 16 ? 4 : 5;
  double a[N];
#pragma unroll
  for (int i = 0; i < N; i++) a[i] = v[i];
  // after step t (offset 32 >> t) a lane keeps the half selected by its bit
#pragma unroll
  for (int t = 0; t < LOG; t++) {
    const int o = 32 >> t, h = N >> (t + 1);
    const bool up = (lane & o) != 0;
    if (o >= 16) {
      // v_permlane32_swap / v_permlane16_swap of (a[i], a[h+i]) leaves, in
      // every lane, the value it keeps in one output and its partner's copy
      // of that value in the other: one swap per 32-bit half and an add,
      // with no selects and no LDS round trip
#pragma unroll
      for (int i = 0; i < h; i++) {
        const unsigned al = (unsigned)__double2loint(a[i]), ah = (unsigned)__double2hiint(a[i]);
        const unsigned bl = (unsigned)__double2loint(a[h + i]), bh = (unsigned)__double2hiint(a[h + i]);
        double x0, x1;
        if (o == 32) {
          const auto l = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
          const auto hh = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
          x0 = __hiloint2double((int)hh[0], (int)l[0]);
          x1 = __hiloint2double((int)hh[1], (int)l[1]);
        } else {
          const auto l = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
          const auto hh = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
          x0 = __hiloint2double((int)hh[0], (int)l[0]);
          x1 = __hiloint2double((int)hh[1], (int)l[1]);
        }
        a[i] = x0 + x1;
      }
      continue;
    }
#pragma unroll
    for (int i = 0; i < h; i++) {
      // the pair is made opaque first: otherwise the compiler turns
      // "up ? a[i] : a[h + i]" into a[up ? i : h + i], an array read with a
      // lane-dependent index, and lowers it to a compare-and-select chain
      // over all N values (about 1,000 extra VALU instructions per
      // reduction of 16 values)
      double lo = a[i], hi = a[h + i];
      asm volatile("" : "+v"(lo), "+v"(hi));
      const double send = up ? lo : hi;
      const double keep = up ? hi : lo;
      a[i] = keep + partner_f64(send, o);
    }
  }
#pragma unroll
  for (int o = 32 >> LOG; o > 0; o >>= 1) a[0] += partner_f64(a[0], o);
  // value index held by this lane: its top LOG bits (bit 5 most significant)
  const int idx = lane >> (6 - LOG);
  if ((lane & ((64 >> LOG) - 1)) == 0) part[wave * N + idx] = a[0];
  __syncthreads();
  double s = 0.0;
  if (lane < N)
    for (int w = 0; w < nw; w++) s += part[w * N + lane];
  return s;
}

template <int N>
__device__ __forceinline__ void wg_allreduce(double (&v)[N], double* part, double (&tot)[N]) {
  const double s = wg_reduce_to_lanes<N>(v, part);
#pragma unroll
  for (int k = 0; k < N; k++) tot[k] = readlane_f64(s, k);
}

__device__ __forceinline__ float readlane_t(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane_t(double v, int l) { return readlane_f64(v, l); }

// ---------------------------------------------------------------------------
// Persistent batched solve.  T = inner precision (float: FP32 inner with FP64
// refinement; double: plain FP64 PCG).  Dynamic LDS: p [nint][3][RG] of T.
// Threads own interior nodes (all RG right-hand sides each).  Per-iteration
// scalars (alpha, beta, the active set) are computed redundantly by every
// thread from the all-reduced totals, so an inner iteration has three
// barriers (two reductions and the p update).  Per-RHS values that persist
// across iterations are written by thread 0 only, double-buffered where they
// are read in the same phase.
//
// After an FP64 residual replacement the FP32 iteration is continued, not
// restarted: the search direction kept in LDS is rescaled to the new units
// and the first beta uses the previous iteration's r.z.
// ---------------------------------------------------------------------------
struct RhsState {
  double bb[RG], rr64[RG], prev[RG], scale[RG], itol2[RG];
  double rz[2][RG];   // r.z of the current iterate, by iteration parity
  double rzkeep[RG];  // r.z of the iteration before the last (continuation)
  double rrpred[RG];  // recursive r.r at the end of the inner solve (scaled)
  double rrmin[2][RG];  // stagnation watch: last r.r that was 4x below the
  int itmin[2][RG];     // previous mark, and the inner iteration it was seen
  double ratio[RG];   // previous scale / new scale
  int done[RG], failed[RG], warm[RG], itbase[RG], it[RG];
  int retried[RG], restart[RG];  // FP64 from a guess: one retry from zero
};

//
// REG: r, d, q and the diagonal inverses of a thread's rows live in its
// registers (rows t + k * blockDim.x, k < RPT; the host launches this form
// only when every cell has nint <= RPT * blockDim.x).  Otherwise they live
// in a per-workgroup scratch slab in global memory.
template <typename T, int RPT, bool REG>
__global__ __launch_bounds__(max_threads(REG, RPT))
__attribute__((amdgpu_waves_per_eu(min_waves(std::is_same<T, double>::value, RPT, REG))))
void k_batch_solve(SolveArgs A) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* P = reinterpret_cast<T*>(smem_raw);
  constexpr int kWaves = max_threads(REG, RPT) / 64;
  __shared__ double partA[kWaves * 2 * RG];
  __shared__ double partB[kWaves * 2 * RG];
  __shared__ RhsState S;
  __shared__ int task_s;
  const int tid = threadIdx.x, nth = blockDim.x, lane = threadIdx.x & 63;
  constexpr bool fp64 = std::is_same<T, double>::value;
  T* R = REG ? nullptr : (fp64 ? (T*)A.scr64 : (T*)A.scr32) + (int64_t)blockIdx.x * 3 * A.scr_stride;
  T* D = REG ? nullptr : R + A.scr_stride;  // inner correction
  T* Q = REG ? nullptr : D + A.scr_stride;  // A p
  constexpr int RQ = REG ? RPT : 1;
  T Rr[RQ][3][RG], Qr[RQ][3][RG], dvr[RQ][3];
  // REG: the inner correction d of the owned rows, in LDS behind p (each
  // thread touches only its own rows' entries) or in registers
  constexpr bool kDInLds = REG && d_in_lds(fp64, RPT);
  T Dr[RQ][3][RG];  // unused (and removed) when d is in LDS
  T* DL = kDInLds ? P + (int64_t)A.max_nint * lds_row<T, REG>() : nullptr;
  // REG with stage_x: the FP64 iterate of the task's rows for the residual
  // pass, in the FP64 d region (free during that pass) or behind p
  double* XS = !REG ? nullptr
                    : (kDInLds && fp64) ? reinterpret_cast<double*>(DL)
                                        : reinterpret_cast<double*>(P + (int64_t)A.max_nint * lds_row<T, REG>() *
                                                                            (kDInLds ? 2 : 1));
  // stencil points whose K values are in flight in the register forms
  constexpr int kPrefetch = fp64 ? 1 : 3;

  int tried = 0;  // queues found empty (thread 0)
  const int home = blockIdx.x % NQ;
  for (;;) {
    if (tid == 0) {
      int t = -1;
      if (A.ntasks_dev) {
        // FP64 pass over the FP32 launch's failures: one queue
        const int k = (int)__hip_atomic_fetch_add(A.queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k < *A.ntasks_dev) t = k;
      } else {
        while (tried < NQ) {
          const int q = (home + tried) % NQ;
          const int k = (int)__hip_atomic_fetch_add(A.queue + q * QSTRIDE, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
          if (A.seg[q] + k < A.seg[q + 1]) {
            t = A.seg[q] + k;
            break;
          }
          tried++;
        }
      }
      task_s = t;
    }
    __syncthreads();
    // uniform across the workgroup: keep the task, its cell record and every
    // pointer derived from them in scalar registers
    const int task = __builtin_amdgcn_readfirstlane(task_s);
    if (task < 0) break;
    const int cell = __builtin_amdgcn_readfirstlane(A.task_cell[task]);
    const int grp = __builtin_amdgcn_readfirstlane(A.task_group[task]);
    const Cell C = A.cells[cell];
    const int n = C.n, m = n - 1, nint = C.nint;
    const float* K32 = A.K32 + C.int0 * (NS * 9);
    const double* K64 = A.K64 + C.int0 * (NS * 9);
    const double* b64 = A.b64 + C.int0 * (3 * NB);
    double* x64 = A.x64 + C.int0 * (3 * NB);
    const T* dinv = fp64 ? (const T*)(A.dinv64 + C.int0 * 3) : (const T*)(A.dinv32 + C.int0 * 3);
    if constexpr (REG) {
#pragma unroll
      for (int q = 0; q < RQ; q++) {
        const int v = tid + q * nth;
#pragma unroll
        for (int a = 0; a < 3; a++) dvr[q][a] = v < nint ? dinv[v * 3 + a] : (T)0;
      }
    }
    const int q0 = grp * RG;
    // the FP64 pass re-solves a task: its outcome replaces the FP32 flags
    // (bit 1, a rejected guess, and bit 3, the hand-over, stay)
    if (A.ntasks_dev && tid < RG) A.flags[cell * NB + q0 + tid] &= ~5;
    // FP32 inner solves that need more than a few times the FP64 count are
    // handed to the FP64 pass (CG iterations grow like n for these cells)
    const int inner_cap = min(A.inner_cap, fp64 ? A.inner_cap : 20 * n + 60);
    // an inner solve whose recursive r.r has not dropped 4x within this many
    // iterations is stagnating: FP32 cannot resolve the problem, or (FP64)
    // CG is stalled on an indefinite operator (the reference's element
    // convention gives local operators with negative eigenvalues)
    constexpr int kStagWindow = fp64 ? 1000 : 60;

    // ||b||^2 per RHS
    {
      double w[RG], tot[RG];
#pragma unroll
      for (int r = 0; r < RG; r++) w[r] = 0;
      for (int v = tid; v < nint; v += nth)
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int r = 0; r < RG; r++) {
            const double bv = b64[(int64_t)v * 3 * NB + a * NB + q0 + r];
            w[r] += bv * bv;
          }
      wg_allreduce<RG>(w, partA, tot);
      if (tid == 0)
        for (int r = 0; r < RG; r++) {
          S.bb[r] = tot[r] > 0 ? tot[r] : 1.0;
          S.done[r] = 0;
          S.failed[r] = 0;
          S.warm[r] = 0;
          S.it[r] = 0;
          S.prev[r] = 1e300;
          S.retried[r] = 0;
          S.scale[r] = 1.0;
        }
    }
    if (!A.use_guess)
      for (int v = tid; v < nint; v += nth)
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int r = 0; r < RG; r++) x64[(int64_t)v * 3 * NB + a * NB + q0 + r] = 0.0;
    __syncthreads();
    bool guess_checked = !A.use_guess;
    for (int outer = 0; outer < 64; outer++) {
      // true residual r64 = b - K64 x64 (FP64) into R (unscaled)
      double rr64[RG];
      {
        double w[RG];
#pragma unroll
        for (int r = 0; r < RG; r++) w[r] = 0;
        // two rows per thread interleaved in the variants that already use
        // the whole register file (one wave per SIMD); the others keep one
        // row here so that they stay at two waves per SIMD
        constexpr int RR = (RPT >= 3 || (fp64 && RPT >= 2)) ? 2 : 1;
        auto ld = [&](int u, int b, double (&pv)[RG]) {
#pragma unroll
          for (int r = 0; r < RG; r++) pv[r] = x64[(int64_t)u * 3 * NB + b * NB + q0 + r];
        };
        if constexpr (REG) {
          // when the launch has room for it, the task's FP64 iterate is
          // staged in LDS (own rows, one barrier), so the stencil reads its
          // neighbours' values from there instead of from global memory
          if (A.stage_x) {
#pragma unroll
            for (int q = 0; q < RQ; q++) {
              const int v = tid + q * nth;
              if (v >= nint) continue;
#pragma unroll
              for (int a = 0; a < 3; a++)
#pragma unroll
                for (int r = 0; r < RG; r++)
                  XS[lds_at<double, true>(v, a, r)] = x64[(int64_t)v * 3 * NB + a * NB + q0 + r];
            }
            __syncthreads();
          }
          auto lds_ld = [&](int u, int b, double (&pv)[RG]) {
#pragma unroll
            for (int r = 0; r < RG; r++) pv[r] = XS[lds_at<double, true>(u, b, r)];
          };
          // one owned row at a time (the accumulators are FP64)
#pragma unroll
          for (int q = 0; q < RQ; q++) {
            const int vr[1] = {tid + q * nth};
            const bool ok[1] = {vr[0] < nint};
            double acc[1][3][RG];
            if (A.stage_x) {
              // unrolled, with the next point's K in flight
              double kr1[1][1][9];
              k_ring_request<double, 1, 1>(K64, nint, vr, ok, kr1);
              stencil_rows_pf<double, double, 1, 1>(K64, nint, vr, ok, m, lds_ld, kr1, acc);
            } else
              stencil_rows<double, double, 1>(K64, nint, vr, ok, m, ld, acc);
            if (!ok[0]) continue;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int r = 0; r < RG; r++) {
                const double rv = b64[(int64_t)vr[0] * 3 * NB + a * NB + q0 + r] - acc[0][a][r];
                w[r] += rv * rv;
                Rr[q][a][r] = (T)rv;
              }
          }
        } else
        for (int v0 = tid; v0 < nint; v0 += RR * nth) {
          int vr[RR];
          bool ok[RR];
#pragma unroll
          for (int q = 0; q < RR; q++) {
            vr[q] = v0 + q * nth;
            ok[q] = vr[q] < nint;
          }
          double acc[RR][3][RG];
          stencil_rows_pp<double, double, RR>(K64, nint, vr, ok, m, ld, acc);
#pragma unroll
          for (int q = 0; q < RR; q++) {
            if (!ok[q]) continue;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int r = 0; r < RG; r++) {
                const double rv = b64[(int64_t)vr[q] * 3 * NB + a * NB + q0 + r] - acc[q][a][r];
                w[r] += rv * rv;
                R[((int64_t)vr[q] * 3 + a) * RG + r] = (T)rv;
              }
          }
        }
        wg_allreduce<RG>(w, partB, rr64);
      }
      if (!guess_checked) {
        // a guess whose residual exceeds the zero guess's (= ||b||) is dropped
        guess_checked = true;
        int bad = 0;
#pragma unroll
        for (int r = 0; r < RG; r++) bad |= (!(rr64[r] <= S.bb[r])) << r;
        if (bad) {
          if constexpr (REG) {
#pragma unroll
            for (int q = 0; q < RQ; q++) {
              const int v = tid + q * nth;
              if (v >= nint) continue;
#pragma unroll
              for (int a = 0; a < 3; a++)
#pragma unroll
                for (int r = 0; r < RG; r++)
                  if ((bad >> r) & 1) {
                    x64[(int64_t)v * 3 * NB + a * NB + q0 + r] = 0.0;
                    Rr[q][a][r] = (T)b64[(int64_t)v * 3 * NB + a * NB + q0 + r];
                  }
            }
          } else
          for (int v = tid; v < nint; v += nth)
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int r = 0; r < RG; r++)
                if ((bad >> r) & 1) {
                  x64[(int64_t)v * 3 * NB + a * NB + q0 + r] = 0.0;
                  R[((int64_t)v * 3 + a) * RG + r] = (T)b64[(int64_t)v * 3 * NB + a * NB + q0 + r];
                }
          if (tid < RG && ((bad >> tid) & 1)) A.flags[cell * NB + q0 + tid] |= 2;
#pragma unroll
          for (int r = 0; r < RG; r++)
            if ((bad >> r) & 1) rr64[r] = S.bb[r];
        }
      }
      if (tid == 0)
        for (int r = 0; r < RG; r++) {
          double rr = rr64[r];
          // an outer step that did not reduce the true residual means the
          // inner precision cannot solve this problem: stop it and report it
          if (!S.done[r] && outer > 0 && !(rr < 0.9 * S.prev[r])) S.failed[r] = 1;
          // FP64 from a guess: a failed problem is solved once more from
          // zero (CG on an indefinite operator can stall from one start and
          // converge from another); its residual is then b
          S.restart[r] = 0;
          if (fp64 && A.use_guess && S.failed[r] && !S.done[r] && !S.retried[r]) {
            S.failed[r] = 0;
            S.retried[r] = 1;
            S.restart[r] = 1;
            S.warm[r] = 0;
            rr = S.bb[r];
          }
          S.rr64[r] = rr;
          S.prev[r] = rr;
          // continue the FP32 iteration only if its recursion stayed
          // consistent with the true residual (within 2x in norm)
          if (S.warm[r] && !(rr <= 4.0 * S.rrpred[r] * S.scale[r] * S.scale[r])) S.warm[r] = 0;
          S.done[r] = S.done[r] || S.failed[r] || (rr <= A.tol * A.tol * S.bb[r]) || !(rr == rr);
          const double sc = rr > 0 ? sqrt(rr) : 1.0;
          S.ratio[r] = S.scale[r] / sc;
          S.scale[r] = sc;
          const double rel = sqrt(rr / S.bb[r]);
          double it = fp64 ? A.tol / rel : fmax(A.inner_tol, 0.5 * A.tol / rel);
          it = fmin(it, 0.5);
          S.itol2[r] = it * it;
          S.itbase[r] = S.it[r];
        }
      __syncthreads();
      if constexpr (fp64) {
        int rsm = 0;
#pragma unroll
        for (int r = 0; r < RG; r++) rsm |= (S.restart[r] != 0) << r;
        if (rsm) {
          // x = 0 and r = b for the restarted problems (own rows / all rows)
          if constexpr (REG) {
#pragma unroll
            for (int q = 0; q < RQ; q++) {
              const int v = tid + q * nth;
              if (v >= nint) continue;
#pragma unroll
              for (int a = 0; a < 3; a++)
#pragma unroll
                for (int r = 0; r < RG; r++)
                  if ((rsm >> r) & 1) {
                    x64[(int64_t)v * 3 * NB + a * NB + q0 + r] = 0.0;
                    Rr[q][a][r] = (T)b64[(int64_t)v * 3 * NB + a * NB + q0 + r];
                  }
            }
          } else {
            for (int v = tid; v < nint; v += nth)
              for (int a = 0; a < 3; a++)
#pragma unroll
                for (int r = 0; r < RG; r++)
                  if ((rsm >> r) & 1) {
                    x64[(int64_t)v * 3 * NB + a * NB + q0 + r] = 0.0;
                    R[((int64_t)v * 3 + a) * RG + r] = (T)b64[(int64_t)v * 3 * NB + a * NB + q0 + r];
                  }
          }
        }
      }
      int live = 0, left = 0;
#pragma unroll
      for (int r = 0; r < RG; r++)
        if (!S.done[r]) {
          live |= 1 << r;
          left = max(left, A.maxit - S.it[r]);
        }
      if (!live || left <= 0) break;
      // start: R <- R/scale, D = 0, z = dinv R; the direction is z (fresh)
      // or z + beta ratio p_kept (continuation)
      int act = 0;
      int par = 0;
      {
        double w[2 * RG], tot[2 * RG];
#pragma unroll
        for (int r = 0; r < 2 * RG; r++) w[r] = 0;
        T inv[RG];
#pragma unroll
        for (int r = 0; r < RG; r++) inv[r] = ((live >> r) & 1) ? (T)(1.0 / S.scale[r]) : (T)0;
        if constexpr (REG) {
#pragma unroll
          for (int q = 0; q < RQ; q++) {
            if (tid + q * nth >= nint) continue;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int r = 0; r < RG; r++) {
                const T rv = ((live >> r) & 1) ? (T)((double)Rr[q][a][r] * (double)inv[r]) : (T)0;
                Rr[q][a][r] = rv;
                if constexpr (kDInLds)
                  DL[lds_at<T, true>(tid + q * nth, a, r)] = 0;
                else
                  Dr[q][a][r] = 0;
                const T z = dvr[q][a] * rv;
                w[r] += (double)(rv * z);
                w[RG + r] += (double)(rv * rv);
              }
          }
        } else
        for (int v = tid; v < nint; v += nth)
          for (int a = 0; a < 3; a++) {
            const T dv = dinv[v * 3 + a];
#pragma unroll
            for (int r = 0; r < RG; r++) {
              const int64_t e = ((int64_t)v * 3 + a) * RG + r;
              const T rv = ((live >> r) & 1) ? (T)((double)R[e] * (double)inv[r]) : (T)0;
              R[e] = rv;
              D[e] = 0;
              const T z = dv * rv;
              w[r] += (double)(rv * z);
              w[RG + r] += (double)(rv * rv);
            }
          }
        wg_allreduce<2 * RG>(w, partA, tot);
        T beta[RG];
#pragma unroll
        for (int r = 0; r < RG; r++) {
          const bool on = ((live >> r) & 1) && tot[RG + r] > S.itol2[r];
          act |= on << r;
          const bool keep = on && S.warm[r] && S.rzkeep[r] > 0.0 && S.ratio[r] > 0.0;
          beta[r] = keep ? (T)(tot[r] / (S.rzkeep[r] * S.ratio[r])) : (T)0;
        }
        if (tid == 0)
          for (int r = 0; r < RG; r++) {
            S.rz[0][r] = tot[r];
            S.rrmin[0][r] = tot[RG + r];
            S.itmin[0][r] = 0;
          }
        if constexpr (REG) {
#pragma unroll
          for (int q = 0; q < RQ; q++) {
            const int v = tid + q * nth;
            if (v >= nint) continue;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int r = 0; r < RG; r++) {
                const int64_t e = lds_at<T, true>(v, a, r);
                const T z = dvr[q][a] * Rr[q][a][r];
                P[e] = beta[r] != (T)0 ? z + beta[r] * P[e] : z;
              }
          }
        } else
        for (int v = tid; v < nint; v += nth)
          for (int a = 0; a < 3; a++) {
            const T dv = dinv[v * 3 + a];
#pragma unroll
            for (int r = 0; r < RG; r++) {
              const int64_t e = ((int64_t)v * 3 + a) * RG + r, ep = lds_at<T, false>(v, a, r);
              const T z = dv * R[e];
              P[ep] = beta[r] != (T)0 ? z + beta[r] * P[ep] : z;
            }
          }
        __syncthreads();
      }
      int inner_it = 0, capfail = 0;
      // tested at the bottom: r and d then stay in one set of registers
      // across iterations (a test at the top made the compiler carry them
      // in two sets and copy 48 values per iteration)
      // K of the first stencil points, carried from one product to the next
      T kring[REG ? kPrefetch : 1][RQ][9];
      if constexpr (REG) {
        if (act && inner_cap > 0) {
          int vr[RQ];
          bool ok[RQ];
#pragma unroll
          for (int q = 0; q < RQ; q++) {
            vr[q] = tid + q * nth;
            ok[q] = vr[q] < nint;
          }
          if constexpr (fp64)
            k_ring_request<double, RQ, kPrefetch>(K64, nint, vr, ok, kring);
          else
            k_ring_request<float, RQ, kPrefetch>(K32, nint, vr, ok, kring);
        }
      }
      if (act && inner_cap > 0) do {
        // q = K p, p.q
        T alpha[RG];
        {
          double w[RG];
#pragma unroll
          for (int r = 0; r < RG; r++) w[r] = 0;
          if constexpr (REG)
            phase_spmv_reg<T, RQ, kPrefetch>(K32, K64, nint, m, P, kring, Qr, w);
          else
            phase_spmv<T, RPT>(K32, K64, nint, m, P, Q, w);
          // alpha of right-hand side r computed in lane r, then broadcast
          // (the state is read before the reduction's barrier: it was
          // written before the previous iteration's last barrier)
          const double rzp = lane < RG ? S.rz[par][lane] : 0.0;
          const double pql = wg_reduce_to_lanes<RG>(w, partB);
          T al = 0;
          if (lane < RG && ((act >> lane) & 1) && pql != 0.0) al = (T)(rzp / pql);
#pragma unroll
          for (int r = 0; r < RG; r++) alpha[r] = readlane_t(al, r);
        }
        // x += alpha p, r -= alpha q, z = dinv r; r.z, r.r
        int act2 = act;
        T beta[RG];
        {
          double w[2 * RG];
#pragma unroll
          for (int r = 0; r < 2 * RG; r++) w[r] = 0;
          if constexpr (REG)
            phase_update_reg<T, RQ, kDInLds>(nint, dvr, P, Qr, Rr, DL, Dr, alpha, w);
          else
            phase_update<T, RPT>(nint, dinv, P, Q, R, D, alpha, w);
          // lane r < RG decides right-hand side r from its r.z (lane r) and
          // r.r (lane RG + r, moved to lane r within the row of 16); the
          // decisions are gathered with ballots.  The state it needs is read
          // before the reduction's barrier (written before an earlier one).
          static_assert(RG == 8, "r.r is moved from lane r + 8 by a row rotation of 8");
          const int sl = lane < RG ? lane : 0;
          const double rz_old = S.rz[par][sl], itol2 = S.itol2[sl], rrmin = S.rrmin[par][sl];
          const int itbase = S.itbase[sl], itmin = S.itmin[par][sl];
          const double rzl = wg_reduce_to_lanes<2 * RG>(w, partA);
          const double rrl = row_ror8_f64(rzl);
          const bool mine = lane < RG && ((act >> lane) & 1);
          bool cont = false, stag = false;
          T be = 0;
          if (mine) {
            const bool stop = !(rrl == rrl) || rrl <= itol2 || itbase + inner_it + 1 >= A.maxit;
            if (!stop) {
              const int itm = rrl < 0.25 * rrmin ? inner_it + 1 : itmin;
              stag = inner_it + 1 - itm > kStagWindow;
            }
            cont = !(stop || stag);
            if (cont) be = (T)(rzl / rz_old);
          }
          act2 = (int)(__ballot(cont) & ((1u << RG) - 1));
          capfail |= (int)(__ballot(stag) & ((1u << RG) - 1));  // recorded after the x update
#pragma unroll
          for (int r = 0; r < RG; r++) beta[r] = readlane_t(be, r);
          if (tid < RG) {
            const int r = tid;
            const bool on = (act >> r) & 1;
            const bool mark = on && rrl < 0.25 * S.rrmin[par][r];
            S.rrmin[par ^ 1][r] = mark ? rrl : S.rrmin[par][r];
            S.itmin[par ^ 1][r] = mark ? inner_it + 1 : S.itmin[par][r];
            if (!on) {
              S.rz[par ^ 1][r] = S.rz[par][r];
            } else if (cont) {
              S.rz[par ^ 1][r] = rzl;
            } else {
              // stopped: keep (p, r.z) of this iteration for the next step
              S.rz[par ^ 1][r] = S.rz[par][r];
              S.rzkeep[r] = S.rz[par][r];
              S.it[r] = S.itbase[r] + inner_it + 1;
              S.warm[r] = rrl <= S.itol2[r];
              S.rrpred[r] = rrl;
            }
          }
        }
        // p = z + beta p (continuing RHS only)
        if constexpr (REG)
          phase_direction_reg<T, RQ>(nint, dvr, P, Rr, beta, act2);
        else
          phase_direction<T, RPT>(nint, dinv, P, R, beta, act2);
        __syncthreads();
        act = act2;
        par ^= 1;
        inner_it++;
      } while (act && inner_it < inner_cap);
      // right-hand sides still active here reached the iteration cap
      capfail |= act;  // recorded by thread 0 after the x update
      act = 0;
      // x64 += scale * d (not for problems whose inner solve failed)
      if constexpr (REG) {
#pragma unroll
        for (int q = 0; q < RQ; q++) {
          const int v = tid + q * nth;
          if (v >= nint) continue;
#pragma unroll
          for (int a = 0; a < 3; a++)
#pragma unroll
            for (int r = 0; r < RG; r++)
              if (((live >> r) & 1) && !((capfail >> r) & 1) && !S.failed[r])
                x64[(int64_t)v * 3 * NB + a * NB + q0 + r] +=
                    S.scale[r] * (double)(kDInLds ? DL[lds_at<T, true>(v, a, r)] : Dr[q][a][r]);
        }
      } else
      for (int v = tid; v < nint; v += nth)
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int r = 0; r < RG; r++)
            if (((live >> r) & 1) && !((capfail >> r) & 1) && !S.failed[r]) {
              const int64_t e = ((int64_t)v * 3 + a) * RG + r;
              x64[(int64_t)v * 3 * NB + a * NB + q0 + r] += S.scale[r] * (double)D[e];
            }
      __syncthreads();
      if (tid == 0)
        for (int r = 0; r < RG; r++)
          if ((capfail >> r) & 1) {
            S.failed[r] = 1;
            S.it[r] = S.itbase[r] + inner_it;
          }
      if (inner_it == 0) break;  // nothing more can be done
    }
    __syncthreads();
    if (tid < RG) {
      const int r = tid;
      A.iters[cell * NB + q0 + r] = S.it[r];
      A.relres[cell * NB + q0 + r] = sqrt(S.rr64[r] / S.bb[r]);
      if (!(S.rr64[r] <= A.tol * A.tol * S.bb[r])) A.flags[cell * NB + q0 + r] |= 1;
      if (S.failed[r]) A.flags[cell * NB + q0 + r] |= A.fb_count ? (4 | 8) : 4;
      if (S.retried[r]) A.flags[cell * NB + q0 + r] |= 16;
    }
    if (A.fb_count && tid == 0) {
      bool any = false;
      for (int r = 0; r < RG; r++) any |= S.failed[r] != 0;
      if (any) {
        const int slot = __hip_atomic_fetch_add(A.fb_count, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        A.fb_cell[slot] = cell;
        A.fb_group[slot] = grp;
      }
    }
    __syncthreads();
  }
}

__global__ void k_batch_reset(unsigned* queue, int* flags, int nflags, int* fb_count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < NQ) queue[t * QSTRIDE] = 0u;
  if (t == 0 && fb_count) *fb_count = 0;
  for (int i = t; i < nflags; i += gridDim.x * blockDim.x) flags[i] = 0;
}

// Coarse Galerkin matrices Kc = Phi^T K Phi (24x24 per cell), FP64, by an
// element loop: one workgroup per cell, tets processed one at a time.
// form MFS_BASIS_COMPONENTWISE: the coarse functions are Psi_{c,a} =
// e_a (sum_d phi_{c,d})_a (the reference's use of one vector basis per
// coarse node, component by component) and the coarse element is the
// reference's block(I,J) = int B_J^T D B_I of their gradients
// (MsFEM_Grid.h:3837-3975), integrated exactly over the fine tets: the
// textbook Galerkin product, blockwise transposed in the reference
// convention.
__global__ __launch_bounds__(256) void k_galerkin(const Cell* __restrict__ cells, const double* __restrict__ E,
                                                  double nu, double h, int conv, const double* __restrict__ x64,
                                                  const double* __restrict__ ft, int form,
                                                  double* __restrict__ Kc) {
  // GT tets per round: three barriers per round instead of per tet
  constexpr int GT = 8;
  __shared__ double phi[GT][12][NB];  // element values of the 24 bases
  __shared__ double KeS[6][12][12];   // unit-modulus element matrix of each of the 6 tet shapes
  __shared__ double KP[GT][12][NB];
  __shared__ double Eg[GT];
  __shared__ int Pg[GT];
  const int cell = blockIdx.x, tid = threadIdx.x;
  const Cell C = cells[cell];
  const int n = C.n, m = n - 1;
  const double hf = h / n;
  const double lamf = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), muf = 1.0 / (2.0 * (1.0 + nu));
  const int econv = form == MFS_BASIS_COMPONENTWISE ? (int)MFS_ELEM_TEXTBOOK : conv;
  // register tiles: wave wv accumulates tets 2wv, 2wv+1 of every round; lane
  // l owns the 3x3 tile (rows 3(l/8).., columns 3(l%8)..) of the 24x24 sum
  const int wv = tid >> 6, ln = tid & 63, ti = ln >> 3, tj = ln & 7;
  double acc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const int ntet = 6 * n * n * n;
  // every tet of the cell is a translate of one of 6 shapes: Ke = E_t KeS[shape]
  if (tid < 6 * 16) {
    const int p = tid / 16, I = (tid / 4) % 4, J = tid % 4;
    int ijk[4][3];
    tet_nodes(p, n, ijk);  // hex 0, shape p
    double gr[4][3], V;
    tet_grads(ijk, hf, gr, V);
    double Kb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    elastic_block(gr[I], gr[J], V * lamf, V * muf, econv, Kb);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) KeS[p][3 * I + a][3 * J + b] = Kb[a][b];
  }
  for (int t0 = 0; t0 < ntet; t0 += GT) {
    // element basis values of GT tets (zero past the last tet)
    for (int q = tid; q < GT * 12 * NB; q += blockDim.x) {
      const int g = q / (12 * NB), p = (q / NB) % 12, bq = q % NB;
      const int t = t0 + g;
      double val = 0.0;
      if (t < ntet) {
        // vertex v of the tet: the hex corner plus one along every axis
        // stepped within the first v steps of the tet's axis order
        const int vtx = p / 3, a = p % 3;
        const int hex = t / 6, pm = t % 6;
        const int i = hex % n + (vtx > kStepOf[pm][0]), j = (hex / n) % n + (vtx > kStepOf[pm][1]),
                  k = hex / (n * n) + (vtx > kStepOf[pm][2]);
        if (i >= 1 && i <= m && j >= 1 && j <= m && k >= 1 && k <= m) {
          val = x64[(C.int0 + interior_index(i, j, k, m)) * 3 * NB + a * NB + bq];
        } else {
          double gv[3];
          if (ft)
            face_trace_at(ft + C.ft0, n, i, j, k, bq, gv);
          else
            hat_trace_at(n, i, j, k, bq, gv);
          val = gv[a];
        }
      }
      phi[g][p][bq] = val;
    }
    // moduli and shapes of the round's tets (modulus 0 past the last tet)
    if (tid < GT) {
      const int t = t0 + tid;
      Eg[tid] = t < ntet ? E[C.tet0 + t] : 0.0;
      Pg[tid] = t < ntet ? t % 6 : 0;
    }
    __syncthreads();
    if (form == MFS_BASIS_COMPONENTWISE) {
      // Psi_{c,a2}[p] = (p's component == a2) * sum_d phi_{c,d}[p]
      for (int q = tid; q < GT * 12 * 8; q += blockDim.x) {
        const int g = q / 96, p = (q / 8) % 12, c = q % 8;
        const double sum = phi[g][p][3 * c] + phi[g][p][3 * c + 1] + phi[g][p][3 * c + 2];
        for (int a2 = 0; a2 < 3; a2++) phi[g][p][3 * c + a2] = (p % 3 == a2) ? sum : 0.0;
      }
      __syncthreads();
    }
    {
      // KP = Ke phi: one 3x3 tile (rows 3pb.., columns 3cb..) of one tet per thread
      static_assert(GT * 4 * 8 == 256, "one KP tile per thread");
      const int g = tid >> 5, pb = (tid >> 3) & 3, cb = tid & 7;
      const int shp = Pg[g];
      const double et = Eg[g];
      double kp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
      for (int u = 0; u < 12; u++) {
        double kr[3], pr[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          kr[r] = KeS[shp][3 * pb + r][u];
          pr[r] = phi[g][u][3 * cb + r];
        }
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) kp[r][c] += kr[r] * pr[c];
      }
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) KP[g][3 * pb + r][3 * cb + c] = et * kp[r][c];
    }
    __syncthreads();
#pragma unroll
    for (int gg = 0; gg < 2; gg++) {
      const int g = 2 * wv + gg;
#pragma unroll
      for (int p = 0; p < 12; p++) {
        double a[3], b[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          a[r] = phi[g][p][3 * ti + r];
          b[r] = KP[g][p][3 * tj + r];
        }
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) acc[r][c] += a[r] * b[c];
      }
    }
    __syncthreads();
  }
  // sum the four waves' tiles (phi is free now), then store
  double* red = &phi[0][0][0];  // 4 x 576 doubles fit in GT*12*NB
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) red[wv * NB * NB + (3 * ti + r) * NB + 3 * tj + c] = acc[r][c];
  __syncthreads();
  const bool bt = form == MFS_BASIS_COMPONENTWISE && conv == MFS_ELEM_REFERENCE;
  for (int pr = tid; pr < NB * NB; pr += blockDim.x) {
    const int bi = pr / NB, bj = pr % NB;
    const double v = (red[pr] + red[NB * NB + pr]) + (red[2 * NB * NB + pr] + red[3 * NB * NB + pr]);
    const int out = bt ? (3 * (bi / 3) + bj % 3) * NB + 3 * (bj / 3) + bi % 3 : pr;
    Kc[(int64_t)cell * NB * NB + out] = v;
  }
}

// ---------------------------------------------------------------------------
// Reference-trace mode, step 1: coarse-edge profiles.
// The tet owning the boundary triangle of square (a, b) of face (ax, side):
// the lower triangle steps A1 then A2 from the square's low corner, the
// upper one A2 then A1; the tet steps the face axis last (side 0) or first
// (side 1).  Perm index in the order xyz,xzy,yxz,yzx,zxy,zyx = 2*o0 + (o1>o2).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int face_tri_tet(int n, int ax, int side, int a, int b, int upper) {
  int A1, A2;
  face_axes(ax, A1, A2);
  int hx[3];
  hx[ax] = side ? n - 1 : 0;
  hx[A1] = a;
  hx[A2] = b;
  const int s1 = upper ? A2 : A1, s2 = upper ? A1 : A2;
  const int o0 = side ? ax : s1, o1 = side ? s1 : s2, o2 = side ? s2 : ax;
  return 6 * (hx[0] + n * (hx[1] + n * hx[2])) + 2 * o0 + (o1 > o2 ? 1 : 0);
}

// One thread per (cell, edge).  A segment's modulus is that of the boundary
// triangle holding it in the first face (lower face axis) that contains the
// edge (the reference reads it from the first neighbouring face,
// MsFEM_Grid.h:4880-4920).  The chain of equal segments with diagonal blocks
// (E_s/L)(mu + (lam+mu) t t^T) is solved exactly per component: the profile
// from end 0 is 1 - C(t)/C(n) with C(t) = sum_{s<t} 1/E_s (FEM_1D with
// Dirichlet 1 / 0 ends, FEM.h:3796-4365), the same for every component.
__global__ void k_edge_profiles(const Cell* __restrict__ cells, int ncells, const double* __restrict__ E,
                                double* __restrict__ et) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncells * 12) return;
  const int cell = g / 12, e = g % 12;
  const Cell C = cells[cell];
  const int n = C.n;
  const int dir = e / 4, b1 = e & 1, b2 = (e >> 1) & 1;
  const int o1 = dir == 0 ? 1 : 0;  // lower of the two other axes
  // first face: axis o1, side b1; the edge's fixed coordinate on the higher
  // other axis is b2*n
  int A1, A2;
  face_axes(o1, A1, A2);
  double* out = et + C.et0 + (int64_t)e * (n + 1) * 2;
  double acc = 0.0;
  out[0] = 0.0;  // C(0)
  for (int t = 0; t < n; t++) {
    int tet;
    if (dir == A1)  // edge along u at v = b2*n
      tet = b2 ? face_tri_tet(n, o1, b1, t, n - 1, 1) : face_tri_tet(n, o1, b1, t, 0, 0);
    else  // edge along v at u = b2*n
      tet = b2 ? face_tri_tet(n, o1, b1, n - 1, t, 0) : face_tri_tet(n, o1, b1, 0, t, 1);
    acc += 1.0 / E[C.tet0 + tet];
    out[2 * (t + 1)] = acc;
  }
  for (int t = 0; t <= n; t++) {
    const double ct = out[2 * t];
    out[2 * t] = (acc - ct) / acc;  // end 0 carries the coarse-node value
    out[2 * t + 1] = ct / acc;
  }
}

// ---------------------------------------------------------------------------
// Reference-trace mode, step 2: face solves.  One workgroup per (cell,
// face): the 2-D elastic problem on the face's (n-1)^2 interior nodes
// (FEM_2D_forElasticDeformation, FEM.h:4367-5178) for 12 right-hand sides
// (4 face corners x 3 drives).  Element: per boundary triangle area *
// block(gI, gJ) with the in-plane hat gradients (FEM_Grid.h:4429-4500) and
// the owning tet's modulus.  The reference iterates (SSOR-PCG to 1e-15);
// here the same discrete system is factored: node-major DOFs 3w+a with
// w = (u-1) + m(v-1) give a half bandwidth of 3m+5, so a banded LDL^T of the
// whole face fits in LDS (<= 113 KB at n = 12) and costs ~N bw^2 / 2 flops
// instead of ~150 iterations of 12 stencil products.  One barrier per
// elimination step: the trailing update and the forward elimination of the
// 12 right-hand sides use the unscaled pivot column; the column is scaled to
// L one step later, when nothing reads it.  A final residual pass (blocks
// recomputed per node) checks every solve against the trace tolerance.
// ---------------------------------------------------------------------------
constexpr int FS = 7;       // face stencil points
constexpr int FRHS = 12;    // face bases: 4 corners x 3 drives
constexpr int F_THREADS = 256;
constexpr int kMaxFaceBw = 3 * MAXN + 2;  // face_bw(MAXN - 1)
static_assert(kMaxFaceBw * (kMaxFaceBw + 1) / 2 <= 3 * F_THREADS, "trailing-update entries per thread");
__device__ constexpr int kFace[FS][2] = {{-1, -1}, {0, -1}, {-1, 0}, {0, 0}, {1, 0}, {0, 1}, {1, 1}};
// The 6 triangles around a face node: vertex offsets (du, dv) from the node,
// and the low corner of their square with the upper-triangle flag.  Lower
// triangle of square (a,b): (a,b),(a+1,b),(a+1,b+1); upper: (a,b),(a,b+1),(a+1,b+1).
__device__ constexpr int kTriV[6][3][2] = {{{0, 0}, {1, 0}, {1, 1}},     {{0, 0}, {0, 1}, {1, 1}},
                                           {{-1, -1}, {0, -1}, {0, 0}},  {{-1, -1}, {-1, 0}, {0, 0}},
                                           {{-1, 0}, {0, 0}, {0, 1}},    {{0, -1}, {0, 0}, {1, 0}}};
__device__ constexpr int kTriSq[6][3] = {{0, 0, 0}, {0, 0, 1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, 0}, {0, -1, 1}};
__device__ constexpr int kTriC[6] = {0, 0, 2, 2, 1, 1};  // the node's vertex index

__host__ __device__ constexpr int face_slot(int du, int dv) {
  return du == -1 ? (dv == -1 ? 0 : 2) : du == 0 ? (dv == -1 ? 1 : dv == 0 ? 3 : 5) : (dv == 0 ? 4 : 6);
}

// Stencil blocks of interior face node (u, v): Kb[slot][3a+b].
__device__ __forceinline__ void face_node_blocks(const double* __restrict__ Ecell, int n, int ax, int side, int A1,
                                                 int A2, double hf, double lamf, double muf, int conv, int u, int v,
                                                 double (&Kb)[FS][9]) {
#pragma unroll
  for (int q = 0; q < FS; q++)
#pragma unroll
    for (int e = 0; e < 9; e++) Kb[q][e] = 0.0;
#pragma unroll
  for (int tr = 0; tr < 6; tr++) {
    // 2-D hat gradients in (u, v) units: g_i = perp(P_{i+2} - P_{i+1}) / (2A)
    const double area2 = (double)((kTriV[tr][1][0] - kTriV[tr][0][0]) * (kTriV[tr][2][1] - kTriV[tr][0][1]) -
                                  (kTriV[tr][1][1] - kTriV[tr][0][1]) * (kTriV[tr][2][0] - kTriV[tr][0][0]));
    double g[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      const double gu = -(double)(kTriV[tr][k2][1] - kTriV[tr][k1][1]) / (area2 * hf);
      const double gv = (double)(kTriV[tr][k2][0] - kTriV[tr][k1][0]) / (area2 * hf);
      // embed: component A1 <- gu, A2 <- gv, face axis <- 0
      g[k][0] = A1 == 0 ? gu : (A2 == 0 ? gv : 0.0);
      g[k][1] = A1 == 1 ? gu : (A2 == 1 ? gv : 0.0);
      g[k][2] = A1 == 2 ? gu : (A2 == 2 ? gv : 0.0);
    }
    const double area = 0.5 * fabs(area2) * hf * hf;
    const double Et =
        Ecell[face_tri_tet(n, ax, side, u + kTriSq[tr][0], v + kTriSq[tr][1], kTriSq[tr][2])];
    const int ic = kTriC[tr];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double blk[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      elastic_block(g[ic], g[k], area * Et * lamf, area * Et * muf, conv, blk);
      const int q = face_slot(kTriV[tr][k][0] - kTriV[tr][ic][0], kTriV[tr][k][1] - kTriV[tr][ic][1]);
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Kb[q][3 * a + b] += blk[a][b];
    }
  }
}

__host__ __device__ __forceinline__ int face_bw(int m) { return 3 * (m + 1) + 2; }
size_t face_lds_bytes(int max_n) {
  const int m = max_n - 1, N = 3 * m * m, bw = face_bw(m);
  return sizeof(double) * ((size_t)N * (bw + 1) + (size_t)N * FRHS + 2 * FRHS);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(F_THREADS) void k_face_traces(const Cell* __restrict__ cells,
                                                           const int* __restrict__ order,
                                                           const double* __restrict__ E, double nu, double h,
                                                           int conv, const double* __restrict__ et,
                                                           double* __restrict__ ft, double tol2,
                                                           unsigned long long* worst, int* nfail) {
  extern __shared__ double lds[];
  const int cell = order[blockIdx.x / 6], f = blockIdx.x % 6;
  const Cell C = cells[cell];
  const int n = C.n, m = n - 1, nf = m * m, s = n + 1, N = 3 * nf, bw = face_bw(m), ld = bw + 1;
  const int ax = f >> 1, side = f & 1;
  int A1, A2;
  face_axes(ax, A1, A2);
  double* B = lds;                    // band, row i: columns i-bw .. i at B[i*ld + j-i+bw]
  double* X = B + (size_t)N * ld;     // [N][FRHS]: right-hand sides, then solutions
  double* red = X + (size_t)N * FRHS; // [2][FRHS]: r.r, b.b
  double* FT = ft + C.ft0 + (int64_t)f * s * s * FT_NODE;
  const double* ET = et + C.et0;
  const double* Ecell = E + C.tet0;
  const double hf = h / n;
  const double lamf = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), muf = 1.0 / (2.0 * (1.0 + nu));
  const int tid = threadIdx.x;

  // (a) face-boundary nodes: edge profiles times e_d; zero the band
  for (int w = tid; w < s * s; w += blockDim.x) {
    const int u = w % s, v = w / s;
    if (u > 0 && u < n && v > 0 && v < n) continue;
    double val[4] = {0.0, 0.0, 0.0, 0.0};
    const bool cu = u == 0 || u == n, cv = v == 0 || v == n;
    if (cu && cv) {
      val[(u == n) + 2 * (v == n)] = 1.0;
    } else if (cv) {  // edge along A1 at v = 0 or n
      const int e = edge_id(A1, ax, side, A2, v == n);
      const int hi = 2 * (v == n);
      val[hi] = ET[((int64_t)e * s + u) * 2];
      val[hi + 1] = ET[((int64_t)e * s + u) * 2 + 1];
    } else {  // edge along A2 at u = 0 or n
      const int e = edge_id(A2, ax, side, A1, u == n);
      const int lo = (u == n);
      val[lo] = ET[((int64_t)e * s + v) * 2];
      val[lo + 2] = ET[((int64_t)e * s + v) * 2 + 1];
    }
    double* o = FT + (int64_t)w * FT_NODE;
    for (int a = 0; a < 3; a++)
      for (int lc = 0; lc < 4; lc++)
        for (int d = 0; d < 3; d++) o[a * 12 + 3 * lc + d] = a == d ? val[lc] : 0.0;
  }
  for (int e = tid; e < N * ld; e += blockDim.x) B[e] = 0.0;
  if (tid < 2 * FRHS) red[tid] = 0.0;
  __syncthreads();  // boundary traces visible to the workgroup; band zeroed

  // (b) band (lower triangle) and right-hand sides, one thread per node
  static_assert((MAXN - 1) * (MAXN - 1) <= F_THREADS, "one face node per thread");
  if (tid < nf) {
    const int w = tid, u = w % m + 1, v = w / m + 1;
    double Kb[FS][9];
    face_node_blocks(Ecell, n, ax, side, A1, A2, hf, lamf, muf, conv, u, v, Kb);
#pragma unroll
    for (int q = 0; q < FS; q++) {
      const int nu_ = u + kFace[q][0], nv = v + kFace[q][1];
      if (!(nu_ > 0 && nu_ < n && nv > 0 && nv < n)) continue;
      const int wn = (nu_ - 1) + m * (nv - 1);
      if (wn > w) continue;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          const int i = 3 * w + a, j = 3 * wn + b;
          if (j <= i) B[(size_t)i * ld + (j - i + bw)] = Kb[q][3 * a + b];
        }
    }
    for (int r0 = 0; r0 < FRHS; r0 += 4) {  // b = -K_IB g, four right-hand sides at a time
      double rhs[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
      for (int q = 0; q < FS; q++) {
        const int nu_ = u + kFace[q][0], nv = v + kFace[q][1];
        if (nu_ > 0 && nu_ < n && nv > 0 && nv < n) continue;
        const double* gv = FT + ((int64_t)nu_ + s * nv) * FT_NODE + r0;
        for (int a = 0; a < 3; a++)
          for (int r = 0; r < 4; r++)
            rhs[a][r] -= Kb[q][3 * a] * gv[r] + Kb[q][3 * a + 1] * gv[12 + r] + Kb[q][3 * a + 2] * gv[24 + r];
      }
      for (int a = 0; a < 3; a++)
        for (int r = 0; r < 4; r++) X[(size_t)(3 * w + a) * FRHS + r0 + r] = rhs[a][r];
    }
  }
  __syncthreads();

  // (c) banded LDL^T with the forward elimination of the 12 right-hand sides
  const int npairs = bw * (bw + 1) / 2;
  int pdi[3], pdj[3];  // this thread's trailing-update entries (di >= dj >= 1)
  for (int c = 0; c < 3; c++) {
    const int p = tid + F_THREADS * c;
    pdi[c] = pdj[c] = 0;
    if (p < npairs) {
      int di = 1, base = 0;
      while (base + di <= p) {
        base += di;
        di++;
      }
      pdi[c] = di;
      pdj[c] = p - base + 1;
    }
  }
  for (int k = 0; k < N; k++) {
    const double dk = B[(size_t)k * ld + bw];
    const double inv = 1.0 / dk;
    for (int c = 0; c < 3; c++) {
      const int i = k + pdi[c], j = k + pdj[c];
      if (pdi[c] == 0 || i >= N) continue;
      B[(size_t)i * ld + (j - i + bw)] -= B[(size_t)i * ld + (k - i + bw)] * B[(size_t)j * ld + (k - j + bw)] * inv;
    }
    for (int t = tid; t < bw * FRHS; t += F_THREADS) {
      const int i = k + 1 + t / FRHS, r = t % FRHS;
      if (i < N) X[(size_t)i * FRHS + r] -= B[(size_t)i * ld + (k - i + bw)] * inv * X[(size_t)k * FRHS + r];
    }
    if (k > 0 && tid < bw) {  // scale column k-1 to L (no one reads it in this step)
      const int i = k + tid;
      if (i < N) B[(size_t)i * ld + (k - 1 - i + bw)] /= B[(size_t)(k - 1) * ld + bw];
    }
    __syncthreads();
  }
  for (int t = tid; t < N * FRHS; t += F_THREADS) X[t] /= B[(size_t)(t / FRHS) * ld + bw];
  __syncthreads();
  for (int k = N - 1; k > 0; k--) {  // L^T x = y
    for (int t = tid; t < bw * FRHS; t += F_THREADS) {
      const int i = k - 1 - t / FRHS, r = t % FRHS;
      if (i >= 0 && i >= k - bw) X[(size_t)i * FRHS + r] -= B[(size_t)k * ld + (i - k + bw)] * X[(size_t)k * FRHS + r];
    }
    __syncthreads();
  }

  // (d) interior face values, residual check r = b - K x, and iterative
  // refinement with the same factor for faces whose residual is above the
  // tolerance (nf <= F_THREADS: one node per thread)
  const bool own = tid < nf;
  const int uo = own ? tid % m + 1 : 1, vo = own ? tid / m + 1 : 1;
  double Kb[FS][9];
  face_node_blocks(Ecell, n, ax, side, A1, A2, hf, lamf, muf, conv, uo, vo, Kb);
  for (int pass = 0;; pass++) {
    if (own) {
      double* o = FT + ((int64_t)uo + s * vo) * FT_NODE;
      for (int a = 0; a < 3; a++)
        for (int r = 0; r < FRHS; r++)
          o[12 * a + r] = (pass ? o[12 * a + r] : 0.0) + X[(size_t)(3 * tid + a) * FRHS + r];
    }
    if (tid < 2 * FRHS) red[tid] = 0.0;
    __syncthreads();  // x complete in FT; X free for the residual
    for (int r0 = 0; r0 < FRHS; r0 += 4) {
      double rr[4] = {0, 0, 0, 0}, bb[4] = {0, 0, 0, 0};
      if (own) {
        double res[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        double bvec[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
        for (int q = 0; q < FS; q++) {
          const int nu_ = uo + kFace[q][0], nv = vo + kFace[q][1];
          const bool in = nu_ > 0 && nu_ < n && nv > 0 && nv < n;
          const double* xn = FT + ((int64_t)nu_ + s * nv) * FT_NODE + r0;
          for (int b = 0; b < 3; b++)
            for (int r = 0; r < 4; r++) {
              const double xv = xn[12 * b + r];
              for (int a = 0; a < 3; a++) {
                res[a][r] -= Kb[q][3 * a + b] * xv;
                if (!in) bvec[a][r] -= Kb[q][3 * a + b] * xv;
              }
            }
        }
        for (int a = 0; a < 3; a++)
          for (int r = 0; r < 4; r++) {
            // res = -K_II x - K_IB g = b - K_II x (b = -K_IB g = bvec)
            X[(size_t)(3 * tid + a) * FRHS + r0 + r] = res[a][r];
            rr[r] += res[a][r] * res[a][r];
            bb[r] += bvec[a][r] * bvec[a][r];
          }
      }
      for (int r = 0; r < 4; r++) {
        rr[r] = wave_sum(rr[r]);
        bb[r] = wave_sum(bb[r]);
      }
      if ((tid & 63) == 0)
        for (int r = 0; r < 4; r++) {
          atomicAdd(&red[r0 + r], rr[r]);
          atomicAdd(&red[FRHS + r0 + r], bb[r]);
        }
    }
    __syncthreads();
    bool bad = false;
    for (int r = 0; r < FRHS; r++) {
      const double rel2 = red[FRHS + r] > 0.0 ? red[r] / red[FRHS + r] : red[r];
      bad = bad || !(rel2 <= tol2);
    }
    if (!bad || pass == 2) break;
    // L D L^T delta = r (X), then x += delta at the top of the next pass
    for (int k = 0; k + 1 < N; k++) {
      for (int t = tid; t < bw * FRHS; t += F_THREADS) {
        const int i = k + 1 + t / FRHS, r = t % FRHS;
        if (i < N) X[(size_t)i * FRHS + r] -= B[(size_t)i * ld + (k - i + bw)] * X[(size_t)k * FRHS + r];
      }
      __syncthreads();
    }
    for (int t = tid; t < N * FRHS; t += F_THREADS) X[t] /= B[(size_t)(t / FRHS) * ld + bw];
    __syncthreads();
    for (int k = N - 1; k > 0; k--) {
      for (int t = tid; t < bw * FRHS; t += F_THREADS) {
        const int i = k - 1 - t / FRHS, r = t % FRHS;
        if (i >= 0 && i >= k - bw) X[(size_t)i * FRHS + r] -= B[(size_t)k * ld + (i - k + bw)] * X[(size_t)k * FRHS + r];
      }
      __syncthreads();
    }
  }
  if (tid < FRHS) {
    const double rr = red[tid], bb = red[FRHS + tid];
    const double rel2 = bb > 0.0 ? rr / bb : rr;
    const bool ok = rel2 == rel2 && rel2 <= tol2;
    if (!ok) atomicAdd(nfail, 1);
    const double rel = rel2 == rel2 ? sqrt(rel2) : 1e300;
    atomicMax(worst, (unsigned long long)__double_as_longlong(rel));
  }
}


}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct mfs_batch_s {
  mfs_ctx ctx = nullptr;
  int ncells = 0;
  double h = 1.0, nu = 0.3;
  int conv = MFS_ELEM_REFERENCE;
  std::vector<Cell> cells;
  std::vector<int64_t> int0;  // per cell
  int64_t total_int = 0, total_tets = 0;
  int max_n = 0, max_nint = 0;
  int ntasks = 0;
  int seg[NQ + 1] = {};  // task ranges of the per-XCD queues
  bool assembled = false;
  int tuned_wg[2] = {0, 0}, tuned_per_cu[2] = {0, 0};  // per precision (0 mixed, 1 FP64)
  int tuned_var[2] = {-1, -1};                        // kernel variant (rows per thread)
  int trace = MFS_TRACE_Q1, form = MFS_BASIS_PER_DIRECTION;
  double trace_tol = 1e-11;  // face solves: FP64 residual floor of ill-conditioned faces is ~1e-12
  int64_t et_total = 0, ft_total = 0;
  double face_worst = 0.0;  // largest relative residual of the face solves
  int face_failed = 0;       // face solves above trace_tol
  double trace_ms = 0.0;       // device time of the edge + face launches
  // device
  Cell* d_cells = nullptr;
  int64_t* d_int0 = nullptr;
  Tmpl* d_tmpls = nullptr;
  int64_t* d_gptr = nullptr;
  uint32_t* d_gent = nullptr;
  double* d_E = nullptr;
  float* d_K32 = nullptr;
  double* d_K64 = nullptr;
  float* d_dinv32 = nullptr;
  double* d_dinv64 = nullptr;
  double* d_b64 = nullptr;
  double* d_x64 = nullptr;
  int* d_iters = nullptr;
  double* d_relres = nullptr;
  int* d_flags = nullptr;
  int* d_task_cell = nullptr;
  int* d_task_group = nullptr;
  int* d_fb = nullptr;  // FP64 pass task list: count, cells[ntasks], groups[ntasks]
  unsigned* d_queue = nullptr;
  void* d_scr = nullptr;
  size_t scr_bytes = 0;
  double* d_et = nullptr;  // reference traces: edge profiles
  double* d_ft = nullptr;  // reference traces: face traces
  int* d_face_order = nullptr;
  unsigned long long* d_fstats = nullptr;  // [0] worst relative residual bits, [1] failures
  double* d_hat = nullptr;  // learned guess: the hat interpolant of one cell's entries
  int64_t hat_len = 0;
  unsigned long long* d_report = nullptr;  // k_batch_report's six sums
};

static int batch_free(mfs_batch B) {
  if (!B) return MFS_OK;
  if (B->ctx) (void)hipStreamSynchronize(B->ctx->stream);
  void* ptrs[] = {B->d_cells, B->d_int0, B->d_tmpls, B->d_gptr, B->d_gent, B->d_E, B->d_K32, B->d_K64,
                  B->d_dinv32, B->d_dinv64, B->d_b64, B->d_x64, B->d_iters, B->d_relres, B->d_flags,
                  B->d_task_cell, B->d_task_group, B->d_fb, B->d_queue, B->d_scr, B->d_et, B->d_ft, B->d_face_order,
                  B->d_fstats, B->d_hat, B->d_report};
  for (void* p : ptrs) dfree(p);
  mfs_ctx c = B->ctx;
  delete B;
  ctx_release(c);
  return MFS_OK;
}

extern "C" int mfs_batch_destroy(mfs_batch B) { return batch_free(B); }

extern "C" int mfs_batch_create(mfs_ctx ctx, const mfs_batch_desc* d, mfs_batch* out) {
  MFS_CHECK_ARG(ctx && d && out, "mfs_batch_create: bad argument");
  MFS_CHECK_ARG(d->n_cells > 0 && d->cell_n && d->cell_origin && d->E, "mfs_batch_create: empty batch");
  MFS_CHECK_ARG(d->cell_h > 0 && d->nu > -1.0 && d->nu < 0.5, "mfs_batch_create: bad cell size or nu");
  MFS_CHECK_ARG(d->convention == MFS_ELEM_REFERENCE || d->convention == MFS_ELEM_TEXTBOOK,
                "mfs_batch_create: unknown element convention");
  MFS_CHECK_ARG(d->trace == MFS_TRACE_Q1 || d->trace == MFS_TRACE_REFERENCE, "mfs_batch_create: unknown trace mode");
  MFS_CHECK_ARG(d->coarse_basis == MFS_BASIS_PER_DIRECTION || d->coarse_basis == MFS_BASIS_COMPONENTWISE,
                "mfs_batch_create: unknown coarse basis form");
  MFS_CHECK_ARG(d->trace_tol >= 0 && d->trace_tol < 1, "mfs_batch_create: trace_tol outside [0, 1)");
  for (int c = 0; c < d->n_cells; c++)
    MFS_CHECK_ARG(d->cell_n[c] >= 2 && d->cell_n[c] <= MAXN, "mfs_batch_create: cell subdivision outside 2..12");
  MFS_HIP(hipSetDevice(ctx->device));
  *out = nullptr;
  mfs_batch B = new mfs_batch_s();
  B->ctx = ctx;
  ctx_retain(ctx);
  B->ncells = d->n_cells;
  B->h = d->cell_h;
  B->nu = d->nu;
  B->conv = d->convention;
  B->trace = d->trace;
  B->form = d->coarse_basis;
  if (d->trace_tol > 0) B->trace_tol = d->trace_tol;
  // templates per distinct n
  std::map<int, int> tmpl_of;
  std::vector<Tmpl> tmpls;
  std::vector<int64_t> gptr_all;
  std::vector<uint32_t> gent_all;
  for (int c = 0; c < d->n_cells; c++) {
    const int n = d->cell_n[c];
    if (tmpl_of.count(n)) continue;
    const int s = n + 1, m = n - 1;
    std::vector<double> xyz(3 * (size_t)s * s * s);
    std::vector<int32_t> tets(4 * 6 * (size_t)n * n * n);
    mfs_mesh_box(n, n, n, 0, 0, 0, 1, 1, 1, xyz.data(), tets.data());
    Pattern P;
    build_pattern((int64_t)s * s * s, 6ll * n * n * n, tets.data(), P);
    GatherLists G;
    build_gather_lists(P, 6ll * n * n * n, tets.data(), G);
    Tmpl T;
    T.n = n;
    T.ptr0 = (int64_t)gptr_all.size();
    T.ent0 = (int64_t)gent_all.size();
    int64_t base = 0;
    gptr_all.push_back(0);
    for (int k = 1; k <= m; k++)
      for (int j = 1; j <= m; j++)
        for (int i = 1; i <= m; i++) {
          const int row = i + s * (j + s * k);
          for (int st = 0; st < NS; st++) {
            const int col = (i + h_stencil[st][0]) + s * ((j + h_stencil[st][1]) + s * (k + h_stencil[st][2]));
            const int32_t* b0 = P.col.data() + P.rowptr[row];
            const int32_t* e0 = P.col.data() + P.rowptr[row + 1];
            const int32_t* f = std::lower_bound(b0, e0, col);
            if (f == e0 || *f != col) {
              set_error("mfs_batch_create: stencil point missing from the cell pattern");
              batch_free(B);
              return MFS_ERR_ARG;
            }
            const int64_t slot = f - P.col.data();
            for (int64_t q = G.ptr[slot]; q < G.ptr[slot + 1]; q++) gent_all.push_back(G.entry[q]);
            base += G.ptr[slot + 1] - G.ptr[slot];
            gptr_all.push_back(base);
          }
        }
    tmpl_of[n] = (int)tmpls.size();
    tmpls.push_back(T);
  }
  // cells
  B->cells.resize(d->n_cells);
  B->int0.resize(d->n_cells);
  int64_t t0 = 0, i0 = 0, e0 = 0, f0 = 0;
  for (int c = 0; c < d->n_cells; c++) {
    Cell& C = B->cells[c];
    C.et0 = e0;
    C.ft0 = f0;
    e0 += 12ll * (d->cell_n[c] + 1) * 2;
    f0 += 6ll * (d->cell_n[c] + 1) * (d->cell_n[c] + 1) * FT_NODE;
    C.n = d->cell_n[c];
    C.nint = (C.n - 1) * (C.n - 1) * (C.n - 1);
    C.tet0 = t0;
    C.int0 = i0;
    C.tmpl = tmpl_of[C.n];
    C.ox = d->cell_origin[3 * c];
    C.oy = d->cell_origin[3 * c + 1];
    C.oz = d->cell_origin[3 * c + 2];
    B->int0[c] = i0;
    t0 += 6ll * C.n * C.n * C.n;
    i0 += C.nint;
    B->max_n = std::max(B->max_n, C.n);
    B->max_nint = std::max(B->max_nint, C.nint);
  }
  B->total_tets = t0;
  B->total_int = i0;
  B->et_total = e0;
  B->ft_total = f0;
  for (int c = 0; c < d->n_cells; c++)
    if (!(d->E[c] > 0) && false) break;
  // tasks: 3 per cell, largest cells first (stable by cell id)
  std::vector<int> order(d->n_cells);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return B->cells[a].n > B->cells[b].n; });
  // cell of rank i goes to queue i % NQ (every queue sees the same size mix);
  // within a queue largest cells first, the 3 tasks of a cell together
  std::vector<int> tc, tg;
  for (int q = 0; q < NQ; q++) {
    B->seg[q] = (int)tc.size();
    for (size_t i = q; i < order.size(); i += NQ)
      for (int g = 0; g < NB / RG; g++) {
        tc.push_back(order[i]);
        tg.push_back(g);
      }
  }
  B->seg[NQ] = (int)tc.size();
  B->ntasks = (int)tc.size();
  int st = MFS_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (st == MFS_OK) st = dmalloc(p, bytes);
  };
  alloc((void**)&B->d_cells, sizeof(Cell) * d->n_cells);
  alloc((void**)&B->d_int0, sizeof(int64_t) * d->n_cells);
  alloc((void**)&B->d_tmpls, sizeof(Tmpl) * tmpls.size());
  alloc((void**)&B->d_gptr, sizeof(int64_t) * gptr_all.size());
  alloc((void**)&B->d_gent, sizeof(uint32_t) * std::max<size_t>(1, gent_all.size()));
  alloc((void**)&B->d_E, sizeof(double) * t0);
  alloc((void**)&B->d_K32, sizeof(float) * NS * 9 * i0);
  alloc((void**)&B->d_K64, sizeof(double) * NS * 9 * i0);
  alloc((void**)&B->d_dinv32, sizeof(float) * 3 * i0);
  alloc((void**)&B->d_dinv64, sizeof(double) * 3 * i0);
  alloc((void**)&B->d_b64, sizeof(double) * 3 * NB * i0);
  alloc((void**)&B->d_x64, sizeof(double) * 3 * NB * i0);
  alloc((void**)&B->d_iters, sizeof(int) * NB * d->n_cells);
  alloc((void**)&B->d_relres, sizeof(double) * NB * d->n_cells);
  alloc((void**)&B->d_flags, sizeof(int) * NB * d->n_cells);
  alloc((void**)&B->d_task_cell, sizeof(int) * B->ntasks);
  alloc((void**)&B->d_task_group, sizeof(int) * B->ntasks);
  alloc((void**)&B->d_fb, sizeof(int) * (1 + 2 * (size_t)B->ntasks));
  alloc((void**)&B->d_queue, sizeof(unsigned) * NQ * QSTRIDE);
  alloc((void**)&B->d_report, sizeof(unsigned long long) * 6);
  if (B->trace == MFS_TRACE_REFERENCE) {
    alloc((void**)&B->d_et, sizeof(double) * B->et_total);
    alloc((void**)&B->d_ft, sizeof(double) * B->ft_total);
    alloc((void**)&B->d_face_order, sizeof(int) * d->n_cells);
    alloc((void**)&B->d_fstats, sizeof(unsigned long long) * 2);
  }
  if (st != MFS_OK) {
    batch_free(B);
    return st;
  }
  hipStream_t s = ctx->stream;
  MFS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_stencil), h_stencil, sizeof(h_stencil)));
  MFS_HIP(hipMemcpyAsync(B->d_cells, B->cells.data(), sizeof(Cell) * d->n_cells, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(B->d_int0, B->int0.data(), sizeof(int64_t) * d->n_cells, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(B->d_tmpls, tmpls.data(), sizeof(Tmpl) * tmpls.size(), hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(B->d_gptr, gptr_all.data(), sizeof(int64_t) * gptr_all.size(), hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(B->d_gent, gent_all.data(), sizeof(uint32_t) * gent_all.size(), hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(B->d_E, d->E, sizeof(double) * t0, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(B->d_task_cell, tc.data(), sizeof(int) * B->ntasks, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(B->d_task_group, tg.data(), sizeof(int) * B->ntasks, hipMemcpyHostToDevice, s));
  if (B->d_face_order)
    MFS_HIP(hipMemcpyAsync(B->d_face_order, order.data(), sizeof(int) * d->n_cells, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemsetAsync(B->d_x64, 0, sizeof(double) * 3 * NB * i0, s));
  MFS_HIP(hipMemsetAsync(B->d_iters, 0, sizeof(int) * NB * d->n_cells, s));
  MFS_HIP(hipMemsetAsync(B->d_relres, 0, sizeof(double) * NB * d->n_cells, s));
  MFS_HIP(hipMemsetAsync(B->d_flags, 0, sizeof(int) * NB * d->n_cells, s));
  MFS_HIP(hipStreamSynchronize(s));
  *out = B;
  return MFS_OK;
}

extern "C" int mfs_batch_cell_dofs(mfs_batch B, int cell, int64_t* n_interior, int64_t* n_total) {
  MFS_CHECK_ARG(B && cell >= 0 && cell < B->ncells, "mfs_batch_cell_dofs: bad argument");
  const int n = B->cells[cell].n;
  if (n_interior) *n_interior = 3ll * (n - 1) * (n - 1) * (n - 1);
  if (n_total) *n_total = 3ll * (n + 1) * (n + 1) * (n + 1);
  return MFS_OK;
}

extern "C" int mfs_batch_assemble(mfs_batch B) {
  MFS_CHECK_ARG(B, "mfs_batch_assemble: NULL batch");
  hipStream_t s = B->ctx->stream;
  if (B->trace == MFS_TRACE_REFERENCE) {
    // edge profiles, then the face solves (boundary data of the volume problems)
    MFS_HIP(hipMemsetAsync(B->d_fstats, 0, sizeof(unsigned long long) * 2, s));
    hipEvent_t ev[2];
    MFS_HIP(hipEventCreate(&ev[0]));
    MFS_HIP(hipEventCreate(&ev[1]));
    MFS_HIP(hipEventRecord(ev[0], s));
    k_edge_profiles<<<(B->ncells * 12 + 63) / 64, 64, 0, s>>>(B->d_cells, B->ncells, B->d_E, B->d_et);
    MFS_HIP(hipGetLastError());
    k_face_traces<<<B->ncells * 6, F_THREADS, face_lds_bytes(B->max_n), s>>>(
        B->d_cells, B->d_face_order, B->d_E, B->nu, B->h, B->conv, B->d_et, B->d_ft, B->trace_tol * B->trace_tol,
        B->d_fstats, (int*)(B->d_fstats + 1));
    MFS_HIP(hipGetLastError());
    MFS_HIP(hipEventRecord(ev[1], s));
    unsigned long long hst[2];
    MFS_HIP(hipMemcpyAsync(hst, B->d_fstats, sizeof(hst), hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    MFS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    B->trace_ms = ms;
    double worst;
    std::memcpy(&worst, &hst[0], sizeof(double));
    B->face_worst = worst;
    B->face_failed = (int)(hst[1] & 0xffffffffu);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  const unsigned grid = (unsigned)std::min<int64_t>(65536, (B->total_int + 255) / 256);
  k_cell_assemble<<<grid, 256, 0, s>>>(B->d_cells, B->ncells, B->d_int0, B->d_tmpls, B->d_gptr, B->d_gent, B->d_E,
                                       B->nu, B->h, B->conv, B->d_ft, B->total_int, B->d_K32, B->d_K64,
                                       B->d_dinv32, B->d_dinv64, B->d_b64);
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipStreamSynchronize(s));
  B->assembled = true;
  return MFS_OK;
}

extern "C" int mfs_batch_trace_status(mfs_batch B, double* max_face_relres, int* n_face_failed,
                                      double* trace_ms) {
  MFS_CHECK_ARG(B, "mfs_batch_trace_status: NULL batch");
  MFS_CHECK_ARG(B->trace == MFS_TRACE_REFERENCE, "mfs_batch_trace_status: batch is not in the reference-trace mode");
  MFS_CHECK_ARG(B->assembled, "mfs_batch_trace_status: call mfs_batch_assemble first");
  if (max_face_relres) *max_face_relres = B->face_worst;
  if (n_face_failed) *n_face_failed = B->face_failed;
  if (trace_ms) *trace_ms = B->trace_ms;
  return MFS_OK;
}

// Host copy of one cell's face traces (reference-trace mode; empty otherwise).
static int cell_face_traces(mfs_batch B, int cell, std::vector<double>& ft) {
  ft.clear();
  if (B->trace != MFS_TRACE_REFERENCE) return MFS_OK;
  const Cell& C = B->cells[cell];
  ft.resize(6ull * (C.n + 1) * (C.n + 1) * FT_NODE);
  MFS_HIP(hipMemcpyAsync(ft.data(), B->d_ft + C.ft0, sizeof(double) * ft.size(), hipMemcpyDeviceToHost,
                         B->ctx->stream));
  MFS_HIP(hipStreamSynchronize(B->ctx->stream));
  return MFS_OK;
}

// The 24 coarse functions at boundary node (i,j,k) of a cell, in the batch's
// coarse-basis form: out[q][a].
static void boundary_functions(const mfs_batch_s* B, const std::vector<double>& ft, int n, int i, int j, int k,
                               double (&out)[NB][3]) {
  for (int q = 0; q < NB; q++) {
    double g[3];
    if (B->trace == MFS_TRACE_REFERENCE)
      face_trace_at(ft.data(), n, i, j, k, q, g);
    else
      hat_trace_at(n, i, j, k, q, g);
    for (int a = 0; a < 3; a++) out[q][a] = g[a];
  }
  if (B->form == MFS_BASIS_COMPONENTWISE)
    for (int c = 0; c < 8; c++)
      for (int a = 0; a < 3; a++) {
        const double sum = out[3 * c][a] + out[3 * c + 1][a] + out[3 * c + 2][a];
        for (int a2 = 0; a2 < 3; a2++) out[3 * c + a2][a] = a2 == a ? sum : 0.0;
      }
}

// p, and for the register-resident variants also d, per workgroup
static size_t lds_bytes(mfs_batch B, int prec, int var) {
  const int rpt = prec == MFS_PREC_F64 ? kRptF64[var] : kRptF32[var];
  const bool dl = kRegVar[var] && d_in_lds(prec == MFS_PREC_F64, rpt);
  const int row = prec == MFS_PREC_F64 ? (kRegVar[var] ? lds_row<double, true>() : lds_row<double, false>())
                                        : (kRegVar[var] ? lds_row<float, true>() : lds_row<float, false>());
  return (size_t)B->max_nint * row * (prec == MFS_PREC_F64 ? 8 : 4) * (dl ? 2 : 1);
}

// The residual pass of the register variants stages the task's FP64 iterate
// in LDS: FP64 in the d region (free during that pass), FP32 one row per
// thread (already one workgroup per CU) in a region behind p, when it fits.
static size_t stage_x_extra(mfs_batch B, int prec, int var) {
  if (!kRegVar[var] || prec == MFS_PREC_F64) return 0;
  return (size_t)B->max_nint * lds_row<double, true>() * 8;
}
static bool stage_x(mfs_batch B, int prec, int var) {
  if (!kRegVar[var]) return false;
  if (prec != MFS_PREC_F64 && kRptF32[var] != 1) return false;
  return lds_bytes(B, prec, var) + stage_x_extra(B, prec, var) <= 160 * 1024 - 8192;
}
static size_t launch_lds(mfs_batch B, int prec, int var) {
  return lds_bytes(B, prec, var) + (stage_x(B, prec, var) ? stage_x_extra(B, prec, var) : 0);
}

// Per-workgroup scratch for a launch shape; must run before any capture.
static int batch_scratch(mfs_batch B, int precision, int per_cu, int var) {
  const int grid = B->ctx->cu_count * per_cu;
  const int64_t stride = (int64_t)B->max_nint * 3 * RG;
  const size_t need = kRegVar[var] ? 0 : (size_t)grid * 3 * stride * (precision == MFS_PREC_F64 ? 8 : 4);
  if (need > B->scr_bytes) {
    MFS_HIP(hipStreamSynchronize(B->ctx->stream));
    dfree(B->d_scr);
    B->d_scr = nullptr;
    B->scr_bytes = 0;
    MFS_TRY(dmalloc(&B->d_scr, need));
    B->scr_bytes = need;
  }
  return MFS_OK;
}

// Launch reset + persistent solve on the context stream (capturable: no
// allocation, no memset; the queue and flags are cleared by a kernel).
static const void* batch_kernel(int precision, int var) {
  if (precision == MFS_PREC_F64) {
    switch (var) {
      case 0: return (const void*)k_batch_solve<double, kRptF64[0], kRegVar[0]>;
      case 1: return (const void*)k_batch_solve<double, kRptF64[1], kRegVar[1]>;
      case 2: return (const void*)k_batch_solve<double, kRptF64[2], kRegVar[2]>;
      default: return (const void*)k_batch_solve<double, kRptF64[3], kRegVar[3]>;
    }
  }
  switch (var) {
    case 0: return (const void*)k_batch_solve<float, kRptF32[0], kRegVar[0]>;
    case 1: return (const void*)k_batch_solve<float, kRptF32[1], kRegVar[1]>;
    case 2: return (const void*)k_batch_solve<float, kRptF32[2], kRegVar[2]>;
    default: return (const void*)k_batch_solve<float, kRptF32[3], kRegVar[3]>;
  }
}

// A variant takes workgroups up to its thread bound; the register-resident
// ones also need every cell's rows covered by the workgroup's threads.
static bool variant_fits(mfs_batch B, int precision, int var, int wg) {
  const int rpt = precision == MFS_PREC_F64 ? kRptF64[var] : kRptF32[var];
  if (wg > max_threads(kRegVar[var], rpt)) return false;
  return !kRegVar[var] || (B->max_nint <= rpt * wg && lds_bytes(B, precision, var) <= 160 * 1024 - 8192);
}

// pass: BL_MAIN = the whole batch, BL_MAIN_LIST = the whole batch in FP32,
// listing its failed tasks for the FP64 pass, BL_FP64_FALLBACK = the FP64
// pass over that list (task count read on the device, so both launches can
// be queued together).
enum { BL_MAIN = 0, BL_MAIN_LIST = 1, BL_FP64_FALLBACK = 2 };
static int batch_launch(mfs_batch B, const mfs_batch_solve_desc* d, int wg, int per_cu, int var, int pass) {
  hipStream_t s = B->ctx->stream;
  const int grid = B->ctx->cu_count * per_cu;
  const int64_t stride = (int64_t)B->max_nint * 3 * RG;
  if (var < 0 || var >= kVariants[d->precision == MFS_PREC_F64 ? 1 : 0] || !variant_fits(B, d->precision, var, wg)) {
    set_error("mfs_batch_solve: variants 3 and 4 (register-resident rows) need nint <= rows per thread x workgroup "
              "size for every cell and twice the LDS; workgroups above 256 threads are for variant 3 only");
    return MFS_ERR_ARG;
  }
  const size_t lds = launch_lds(B, d->precision, var);
  const size_t need = kRegVar[var] ? 0 : (size_t)grid * 3 * stride * (d->precision == MFS_PREC_F64 ? 8 : 4);
  if (need > B->scr_bytes) {
    set_error("mfs_batch_solve: scratch not allocated for this launch shape");
    return MFS_ERR_ARG;
  }
  const bool fb = pass == BL_FP64_FALLBACK;
  if (fb && d->precision != MFS_PREC_F64) {
    set_error("mfs_batch_solve: the fallback pass runs in FP64");
    return MFS_ERR_ARG;
  }
  SolveArgs A;
  A.cells = B->d_cells;
  A.task_cell = fb ? B->d_fb + 1 : B->d_task_cell;
  A.task_group = fb ? B->d_fb + 1 + B->ntasks : B->d_task_group;
  A.ntasks = B->ntasks;
  // MFS_ONE_QUEUE=1 keeps the cell-size order in a single queue (A/B runs)
  static const bool one_queue = std::getenv("MFS_ONE_QUEUE") && std::getenv("MFS_ONE_QUEUE")[0] == '1';
  for (int q = 0; q <= NQ; q++) A.seg[q] = one_queue ? (q == 0 ? 0 : B->ntasks) : B->seg[q];
  A.queue = B->d_queue;
  A.K32 = B->d_K32;
  A.K64 = B->d_K64;
  A.dinv32 = B->d_dinv32;
  A.dinv64 = B->d_dinv64;
  A.b64 = B->d_b64;
  A.x64 = B->d_x64;
  A.scr32 = (float*)B->d_scr;
  A.scr64 = (double*)B->d_scr;
  A.scr_stride = stride;
  A.max_nint = B->max_nint;
  A.stage_x = stage_x(B, d->precision, var) ? 1 : 0;
  A.iters = B->d_iters;
  A.relres = B->d_relres;
  A.flags = B->d_flags;
  const bool list_failures = pass == BL_MAIN_LIST && d->precision == MFS_PREC_MIXED;
  A.fb_count = list_failures ? B->d_fb : nullptr;
  A.fb_cell = list_failures ? B->d_fb + 1 : nullptr;
  A.fb_group = list_failures ? B->d_fb + 1 + B->ntasks : nullptr;
  A.ntasks_dev = fb ? B->d_fb : nullptr;
  A.tol = d->tol;
  A.inner_tol = d->inner_tol > 0 ? d->inner_tol : 1e-4;
  A.maxit = d->maxit > 0 ? d->maxit : 20000;
  A.inner_cap = d->precision == MFS_PREC_F64 ? A.maxit : std::min(A.maxit, 1000);
  A.use_guess = d->use_guess;
  // queue counters (and, for the main launch, the flags and the failure
  // count) are cleared by a kernel: the launch stays capturable
  k_batch_reset<<<std::max(1, (B->ncells * NB + 255) / 256), 256, 0, s>>>(
      B->d_queue, B->d_flags, fb ? 0 : B->ncells * NB, list_failures ? B->d_fb : nullptr);
  void* args[] = {&A};
  MFS_HIP(hipLaunchKernel(batch_kernel(d->precision, var), dim3(grid), dim3(wg), args, lds, s));
  return MFS_OK;
}

static int batch_check_desc(mfs_batch B, const mfs_batch_solve_desc* d) {
  MFS_CHECK_ARG(B && d, "mfs_batch_solve: bad argument");
  MFS_CHECK_ARG(B->assembled, "mfs_batch_solve: call mfs_batch_assemble first");
  MFS_CHECK_ARG(d->precision == MFS_PREC_F64 || d->precision == MFS_PREC_MIXED, "mfs_batch_solve: bad precision");
  MFS_CHECK_ARG(d->tol > 0, "mfs_batch_solve: tol must be positive");
  MFS_CHECK_ARG(d->wg_size == 0 || (d->wg_size % 64 == 0 && d->wg_size >= 64 && d->wg_size <= 512),
                "mfs_batch_solve: wg_size must be a multiple of 64 in 64..512");
  MFS_CHECK_ARG(d->variant >= 0 && d->variant <= kVariants[d->precision == MFS_PREC_F64 ? 1 : 0],
                "mfs_batch_solve: variant must be 0 (autotuned) or 1..4");
  MFS_CHECK_ARG(d->fallback_variant >= 0 && d->fallback_variant <= kVariants[1] &&
                    (d->fallback_wg_size == 0 ||
                     (d->fallback_wg_size % 64 == 0 && d->fallback_wg_size >= 64 && d->fallback_wg_size <= 512)),
                "mfs_batch_solve: fallback_variant must be 0..4 and fallback_wg_size a multiple of 64 in 64..512");
  if (lds_bytes(B, d->precision, 0) > 160 * 1024 - 8192) {
    set_error("mfs_batch_solve: cells too large for the LDS-resident search directions (FP64 inner needs n <= 10)");
    return MFS_ERR_UNSUPPORTED;
  }
  return MFS_OK;
}

// Workgroups of size wg that fit one CU at once (registers and LDS), at
// least 1.
static int resident_per_cu(mfs_batch B, int precision, int wg, int var) {
  const size_t lds = launch_lds(B, precision, var);
  const void* fn = batch_kernel(precision, var);
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, wg, lds) != hipSuccess) nb = 1;
  (void)hipGetLastError();
  return std::max(1, std::min(nb, 8));
}

static int pick_shape(mfs_batch B, const mfs_batch_solve_desc* d, int* wg, int* per_cu, int* var) {
  const int t = d->precision == MFS_PREC_F64 ? 1 : 0;
  const int wgs = d->wg_size ? d->wg_size : (B->tuned_wg[t] ? B->tuned_wg[t] : 256);
  *var = d->variant > 0 ? d->variant - 1 : (B->tuned_var[t] >= 0 ? B->tuned_var[t] : 0);
  *wg = wgs;
  *per_cu = (!d->wg_size && B->tuned_per_cu[t]) ? B->tuned_per_cu[t] : resident_per_cu(B, d->precision, wgs, *var);
  return MFS_OK;
}

// The solve report reduced on the device, so that a solve hands 48 bytes to
// the host instead of three per-problem arrays: [0] largest iteration count,
// [1] sum of the counts, [2] bits of the largest true relative residual
// (non-negative doubles order like their bit patterns; NaN is skipped),
// [3] problems not converged, [4] FP64 retries from zero, [5] problems
// handed from FP32 to FP64.  out is cleared by the caller.
__global__ void k_batch_report(const int* __restrict__ it, const int* __restrict__ fl,
                               const double* __restrict__ rr, int nq, unsigned long long* __restrict__ out) {
  unsigned long long v[6] = {0, 0, 0, 0, 0, 0};
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    const unsigned long long i = (unsigned long long)max(it[q], 0);
    const int f = fl[q];
    const double r = rr[q];
    v[0] = max(v[0], i);
    v[1] += i;
    if (r == r) v[2] = max(v[2], (unsigned long long)__double_as_longlong(r > 0.0 ? r : 0.0));
    v[3] += f & 1;
    v[4] += (f >> 1) & 1;
    v[5] += (f >> 3) & 1;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const unsigned long long w = __shfl_xor(v[k], o, 64);
      v[k] = (k == 0 || k == 2) ? max(v[k], w) : v[k] + w;
    }
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out + 0, v[0]);
    atomicAdd(out + 1, v[1]);
    atomicMax(out + 2, v[2]);
    atomicAdd(out + 3, v[3]);
    atomicAdd(out + 4, v[4]);
    atomicAdd(out + 5, v[5]);
  }
}

static int batch_collect(mfs_batch B, mfs_batch_report* rep) {
  hipStream_t s = B->ctx->stream;
  const int nq = B->ncells * NB;
  unsigned long long h[6];
  MFS_HIP(hipMemsetAsync(B->d_report, 0, sizeof(h), s));
  k_batch_report<<<(unsigned)std::min(1024, (nq + 255) / 256), 256, 0, s>>>(B->d_iters, B->d_flags, B->d_relres, nq,
                                                                            B->d_report);
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipMemcpyAsync(h, B->d_report, sizeof(h), hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  rep->max_iterations = (int)h[0];
  rep->mean_iterations = (double)h[1] / nq;
  double rmax;
  std::memcpy(&rmax, &h[2], sizeof(rmax));
  rep->max_true_relres = rmax;
  rep->n_not_converged = (int)h[3];
  rep->n_fallback = (int)h[4];
  rep->n_precision_fallback = (int)h[5];
  return MFS_OK;
}

extern "C" int mfs_batch_solve(mfs_batch B, const mfs_batch_solve_desc* d, mfs_batch_report* rep) {
  MFS_TRY(batch_check_desc(B, d));
  MFS_CHECK_ARG(rep, "mfs_batch_solve: report is NULL");
  int wg, pc, var;
  MFS_TRY(pick_shape(B, d, &wg, &pc, &var));
  // Problems the FP32 inner iteration cannot solve (the reference's
  // indefinite local operators, see k_batch_solve) are solved again in FP64,
  // as the reference falls back to another solver (FEM/FEM.h:3726-3737),
  // starting from where the FP32 run stopped (a start whose residual exceeds
  // the zero guess's is dropped by the kernel's guess check).  The FP32
  // launch lists those tasks on the device and the FP64 pass, queued right
  // behind it, reads the list: no host round trip between the two.
  // the FP64 pass needs its search directions in LDS (cells with n <= 10)
  const bool mixed = d->precision == MFS_PREC_MIXED && lds_bytes(B, MFS_PREC_F64, 0) <= 160 * 1024 - 8192;
  mfs_batch_solve_desc d64 = *d;
  d64.precision = MFS_PREC_F64;
  d64.use_guess = 1;
  d64.variant = d->fallback_variant;
  d64.wg_size = d->fallback_wg_size;
  int wg64 = 0, pc64 = 0, var64 = 0;
  if (mixed) {
    MFS_TRY(pick_shape(B, &d64, &wg64, &pc64, &var64));
    MFS_TRY(batch_scratch(B, MFS_PREC_F64, pc64, var64));
  }
  MFS_TRY(batch_scratch(B, d->precision, pc, var));
  hipStream_t s = B->ctx->stream;
  hipEvent_t ev[3];
  for (auto& e : ev) MFS_HIP(hipEventCreate(&e));
  // MFS_NO_GRAPH=1 launches the same kernels directly (kernel tracing)
  const char* ng = std::getenv("MFS_NO_GRAPH");
  const bool graph = d->use_graph && !(ng && ng[0] == '1');
  int st = MFS_OK;
  if (graph) {
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t ge[2] = {nullptr, nullptr};
    for (int k = 0; k < (mixed ? 2 : 1) && st == MFS_OK; k++) {
      MFS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      st = k == 0 ? batch_launch(B, d, wg, pc, var, mixed ? BL_MAIN_LIST : BL_MAIN)
                  : batch_launch(B, &d64, wg64, pc64, var64, BL_FP64_FALLBACK);
      hipGraph_t gg = nullptr;
      const hipError_t ce = hipStreamEndCapture(s, &gg);
      g[k] = gg;
      if (st == MFS_OK && ce != hipSuccess) {
        set_error(std::string("mfs_batch_solve: graph capture: ") + hipGetErrorString(ce));
        st = MFS_ERR_HIP;
      }
      if (st == MFS_OK && hipGraphInstantiate(&ge[k], g[k], nullptr, nullptr, 0) != hipSuccess) {
        set_error("mfs_batch_solve: graph instantiation failed");
        st = MFS_ERR_HIP;
      }
    }
    if (st == MFS_OK) {
      MFS_HIP(hipEventRecord(ev[0], s));
      MFS_HIP(hipGraphLaunch(ge[0], s));
      MFS_HIP(hipEventRecord(ev[1], s));
      if (mixed) MFS_HIP(hipGraphLaunch(ge[1], s));
      MFS_HIP(hipEventRecord(ev[2], s));
      MFS_HIP(hipEventSynchronize(ev[2]));
    }
    for (int k = 0; k < 2; k++) {
      if (ge[k]) (void)hipGraphExecDestroy(ge[k]);
      if (g[k]) (void)hipGraphDestroy(g[k]);
    }
  } else {
    MFS_HIP(hipEventRecord(ev[0], s));
    st = batch_launch(B, d, wg, pc, var, mixed ? BL_MAIN_LIST : BL_MAIN);
    MFS_HIP(hipEventRecord(ev[1], s));
    if (st == MFS_OK && mixed) st = batch_launch(B, &d64, wg64, pc64, var64, BL_FP64_FALLBACK);
    MFS_HIP(hipEventRecord(ev[2], s));
    MFS_HIP(hipEventSynchronize(ev[2]));
  }
  float ms_main = 0, ms_all = 0;
  if (st == MFS_OK) {
    MFS_HIP(hipEventElapsedTime(&ms_main, ev[0], ev[1]));
    MFS_HIP(hipEventElapsedTime(&ms_all, ev[0], ev[2]));
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  MFS_TRY(st);
  MFS_TRY(batch_collect(B, rep));
  rep->seconds = ms_all * 1e-3;
  rep->kernel_seconds = ms_main * 1e-3;
  if (rep->n_not_converged > 0) {
    set_error("mfs_batch_solve: " + std::to_string(rep->n_not_converged) + " local problems did not converge");
    return MFS_ERR_NOT_CONVERGED;
  }
  return MFS_OK;
}

// Autotuner: times the whole batched solve for each kernel variant (rows
// carried per thread, slab or registers), each workgroup size the variant
// takes (64..256 threads; up to 512 for one register row per thread) and
// as many workgroups per CU as fit at once (from the kernel's register and
// LDS use) or one fewer, and keeps the fastest shape for this precision.
extern "C" int mfs_batch_autotune(mfs_batch B, const mfs_batch_solve_desc* d, int* best_wg, double* best_ms) {
  MFS_TRY(batch_check_desc(B, d));
  const int t = d->precision == MFS_PREC_F64 ? 1 : 0;
  double best = 1e30;
  int bw = 256, bp = 1, bv = 1;
  for (int var = 0; var < kVariants[t]; var++)
  for (int wg : {64, 128, 192, 256, 320, 384, 448, 512}) {
    if (!variant_fits(B, d->precision, var, wg)) continue;
    const int res = resident_per_cu(B, d->precision, wg, var);
    for (int pc : {res, res - 1}) {
      if (pc < 1 || (pc == res - 1 && res == 1)) continue;
      B->tuned_wg[t] = wg;
      B->tuned_per_cu[t] = pc;
      B->tuned_var[t] = var;
      mfs_batch_solve_desc dd = *d;
      dd.wg_size = 0;
      dd.use_guess = 0;
      mfs_batch_report rep;
      int st = mfs_batch_solve(B, &dd, &rep);
      if (st != MFS_OK && st != MFS_ERR_NOT_CONVERGED) return st;
      if (rep.seconds < best) {
        best = rep.seconds;
        bw = wg;
        bp = pc;
        bv = var;
      }
    }
  }
  B->tuned_wg[t] = bw;
  B->tuned_per_cu[t] = bp;
  B->tuned_var[t] = bv;
  // encoded as wg*100 + per_cu*10 + variant
  if (best_wg) *best_wg = bw * 100 + bp * 10 + bv;
  if (best_ms) *best_ms = best * 1e3;
  return MFS_OK;
}

extern "C" int mfs_batch_set_guess(mfs_batch B, const double* guess) {
  MFS_CHECK_ARG(B && guess, "mfs_batch_set_guess: bad argument");
  MFS_HIP(hipMemcpyAsync(B->d_x64, guess, sizeof(double) * 3 * NB * B->total_int, hipMemcpyHostToDevice,
                         B->ctx->stream));
  MFS_HIP(hipStreamSynchronize(B->ctx->stream));
  return MFS_OK;
}

extern "C" int mfs_batch_get_basis(mfs_batch B, int cell, double* basis) {
  MFS_CHECK_ARG(B && basis && cell >= 0 && cell < B->ncells, "mfs_batch_get_basis: bad argument");
  const Cell& C = B->cells[cell];
  const int n = C.n, m = n - 1, s = n + 1;
  std::vector<double> xi((size_t)C.nint * 3 * NB);
  MFS_HIP(hipMemcpyAsync(xi.data(), B->d_x64 + C.int0 * 3 * NB, sizeof(double) * xi.size(), hipMemcpyDeviceToHost,
                         B->ctx->stream));
  MFS_HIP(hipStreamSynchronize(B->ctx->stream));
  std::vector<double> ft;
  MFS_TRY(cell_face_traces(B, cell, ft));
  const int64_t N = 3ll * s * s * s;
  for (int k = 0; k <= n; k++)
    for (int j = 0; j <= n; j++)
      for (int i = 0; i <= n; i++) {
        const int64_t node = i + s * (j + s * k);
        const bool interior = i >= 1 && i <= m && j >= 1 && j <= m && k >= 1 && k <= m;
        double f[NB][3];
        if (interior) {
          const int u = (i - 1) + m * ((j - 1) + m * (k - 1));
          for (int q = 0; q < NB; q++)
            for (int a = 0; a < 3; a++) f[q][a] = xi[((size_t)u * 3 + a) * NB + q];
          if (B->form == MFS_BASIS_COMPONENTWISE)
            for (int c = 0; c < 8; c++)
              for (int a = 0; a < 3; a++) {
                const double sum = f[3 * c][a] + f[3 * c + 1][a] + f[3 * c + 2][a];
                for (int a2 = 0; a2 < 3; a2++) f[3 * c + a2][a] = a2 == a ? sum : 0.0;
              }
        } else {
          boundary_functions(B, ft, n, i, j, k, f);
        }
        for (int q = 0; q < NB; q++)
          for (int a = 0; a < 3; a++) basis[q * N + 3 * node + a] = f[q][a];
      }
  return MFS_OK;
}

extern "C" int mfs_batch_coarse_matrices(mfs_batch B, double* Kc) {
  MFS_CHECK_ARG(B && Kc, "mfs_batch_coarse_matrices: bad argument");
  MFS_CHECK_ARG(B->assembled, "mfs_batch_coarse_matrices: call mfs_batch_assemble first");
  hipStream_t s = B->ctx->stream;
  double* d_kc;
  MFS_TRY(dmalloc((void**)&d_kc, sizeof(double) * NB * NB * B->ncells));
  k_galerkin<<<B->ncells, 256, 0, s>>>(B->d_cells, B->d_E, B->nu, B->h, B->conv, B->d_x64, B->d_ft, B->form, d_kc);
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipMemcpyAsync(Kc, d_kc, sizeof(double) * NB * NB * B->ncells, hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  dfree(d_kc);
  return MFS_OK;
}

extern "C" int mfs_batch_iterations(mfs_batch B, int32_t* iters) {
  MFS_CHECK_ARG(B && iters, "mfs_batch_iterations: bad argument");
  MFS_HIP(hipMemcpyAsync(iters, B->d_iters, sizeof(int) * NB * B->ncells, hipMemcpyDeviceToHost, B->ctx->stream));
  MFS_HIP(hipStreamSynchronize(B->ctx->stream));
  return MFS_OK;
}

// ---------------------------------------------------------------------------
// Learned-guess glue: features, targets and guesses stay on the device.
// ---------------------------------------------------------------------------
__global__ void k_features(const double* __restrict__ E, int64_t n, double mu, double inv_sigma,
                           float* __restrict__ X) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    X[i] = (float)((log(E[i]) - mu) * inv_sigma);
}

// hat_c(x) e_d at interior node v of an n-cell, for basis q = 3c + d
__device__ __forceinline__ double hat_at(int v, int n, int a, int q) {
  const int m = n - 1;
  const int i = v % m + 1, j = (v / m) % m + 1, k = v / (m * m) + 1;
  const int c = q / 3, d = q % 3;
  if (a != d) return 0.0;
  const double X = (double)i / n, Y = (double)j / n, Z = (double)k / n;
  return ((c & 1) ? X : 1 - X) * ((c & 2) ? Y : 1 - Y) * ((c & 4) ? Z : 1 - Z);
}

// The hat interpolant at every entry (v*3+a)*24+q of one n-cell: the same
// for every cell of a batch of equal subdivisions.
__global__ void k_hat_table(int n, int per_cell, double* __restrict__ hat) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < per_cell) hat[r] = hat_at(r / (3 * NB), n, (r / NB) % 3, r % NB);
}

// T[cell][(v*3+a)*24+q] = x - hat (all cells the same n); mode 1: x = hat + G.
// blockIdx.y walks the cells and blockIdx.x the cell's per_cell entries, two
// adjacent entries per thread (per_cell = 72 (n-1)^3 is even and a cell's
// entries start 16-byte aligned, so they move as float2 / double2); hat
// comes from k_hat_table (L2-resident), so an entry costs a load and a store
// and no divisions.
__global__ void k_targets_guess(const Cell* __restrict__ cells, int ncells, int per_cell,
                                const double* __restrict__ hat_tab, double* __restrict__ x64,
                                float* __restrict__ TG, int mode) {
  const int half = per_cell / 2;
  const double2* ht = reinterpret_cast<const double2*>(hat_tab);
  for (int cell = blockIdx.y; cell < ncells; cell += gridDim.y) {
    double2* xc = reinterpret_cast<double2*>(x64 + cells[cell].int0 * 3 * NB);
    float2* tc = TG ? reinterpret_cast<float2*>(TG + (int64_t)cell * per_cell) : nullptr;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < half; r += gridDim.x * blockDim.x) {
      const double2 hat = ht[r];
      if (mode == 0) {
        const double2 x = xc[r];
        tc[r] = make_float2((float)(x.x - hat.x), (float)(x.y - hat.y));
      } else {
        const float2 g = tc ? tc[r] : make_float2(0.f, 0.f);
        xc[r] = make_double2(hat.x + (double)g.x, hat.y + (double)g.y);
      }
    }
  }
}

static int batch_uniform_n(mfs_batch B) {
  for (const Cell& c : B->cells)
    if (c.n != B->cells[0].n) return -1;
  return B->cells[0].n;
}

extern "C" int mfs_batch_features(mfs_batch B, double mu, double sigma, mfs_vec X) {
  MFS_CHECK_ARG(B && X && sigma > 0, "mfs_batch_features: bad argument");
  const int n = batch_uniform_n(B);
  MFS_CHECK_ARG(n > 0, "mfs_batch_features: all cells must have the same subdivision");
  MFS_CHECK_ARG(X->prec == MFS_F32 && X->n == B->total_tets, "mfs_batch_features: X must be FP32 of ncells*6n^3");
  k_features<<<(unsigned)std::min<int64_t>(65536, (B->total_tets + 255) / 256), 256, 0, B->ctx->stream>>>(
      B->d_E, B->total_tets, mu, 1.0 / sigma, (float*)X->d);
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipStreamSynchronize(B->ctx->stream));
  return MFS_OK;
}

static int targets_guess(mfs_batch B, mfs_vec V, int mode) {
  MFS_CHECK_ARG(B->trace == MFS_TRACE_Q1, "mfs_batch_targets/guess: the learned guess needs the Q1 trace mode");
  const int n = batch_uniform_n(B);
  MFS_CHECK_ARG(n > 0, "mfs_batch_targets/guess: all cells must have the same subdivision");
  const int64_t per_cell = (int64_t)(n - 1) * (n - 1) * (n - 1) * 3 * NB;
  if (V) MFS_CHECK_ARG(V->prec == MFS_F32 && V->n == per_cell * B->ncells, "mfs_batch_targets/guess: bad vector");
  MFS_CHECK_ARG(mode == 1 || V, "mfs_batch_targets: bad vector");
  if (B->hat_len != per_cell) {
    dfree(B->d_hat);
    B->d_hat = nullptr;
    B->hat_len = 0;
    const int st = dmalloc((void**)&B->d_hat, sizeof(double) * per_cell);
    if (st != MFS_OK) return st;
    B->hat_len = per_cell;
  }
  k_hat_table<<<(unsigned)((per_cell + 255) / 256), 256, 0, B->ctx->stream>>>(n, (int)per_cell, B->d_hat);
  MFS_HIP(hipGetLastError());
  // 2,048 entries per workgroup along x (2 x 4 per thread), cells along y
  const dim3 grid((unsigned)((per_cell + 2047) / 2048), (unsigned)std::min(B->ncells, 65535));
  k_targets_guess<<<grid, 256, 0, B->ctx->stream>>>(B->d_cells, B->ncells, (int)per_cell, B->d_hat, B->d_x64,
                                                    V ? (float*)V->d : nullptr, mode);
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipStreamSynchronize(B->ctx->stream));
  return MFS_OK;
}

extern "C" int mfs_batch_targets(mfs_batch B, mfs_vec T) {
  MFS_CHECK_ARG(B && T, "mfs_batch_targets: bad argument");
  return targets_guess(B, T, 0);
}

extern "C" int mfs_batch_guess_from(mfs_batch B, mfs_vec G) {
  MFS_CHECK_ARG(B, "mfs_batch_guess_from: bad argument");
  return targets_guess(B, G, 1);
}

// ---------------------------------------------------------------------------
// Coarse MsFEM system (MsFEM/MsFEM.h:571-1062): the cells form a structured
// ncx*ncy*ncz grid (cell c at grid position given by its origin / h), coarse
// node (I,J,K) = I + (ncx+1)(J + (ncy+1)K).  The 24x24 Galerkin matrices are
// scattered in cell order into a BSR3 matrix with the 27-point coarse
// pattern; fixed coarse nodes are eliminated (FEM.h:505-619 rule).
// ---------------------------------------------------------------------------
static int coarse_grid(mfs_batch B, int* ncx, int* ncy, int* ncz, std::vector<int>& gi) {
  int mx = 0, my = 0, mz = 0;
  gi.resize(3 * B->ncells);
  for (int c = 0; c < B->ncells; c++) {
    const Cell& C = B->cells[c];
    const double o[3] = {C.ox, C.oy, C.oz};
    for (int a = 0; a < 3; a++) {
      const double f = o[a] / B->h;
      const int r = (int)std::lround(f);
      if (std::fabs(f - r) > 1e-6 || r < 0) {
        set_error("mfs_msfem_coarse: cell origins do not lie on a grid of spacing cell_h");
        return MFS_ERR_ARG;
      }
      gi[3 * c + a] = r;
    }
    mx = std::max(mx, gi[3 * c] + 1);
    my = std::max(my, gi[3 * c + 1] + 1);
    mz = std::max(mz, gi[3 * c + 2] + 1);
  }
  if ((int64_t)mx * my * mz != B->ncells) {
    set_error("mfs_msfem_coarse: cells do not tile a full box");
    return MFS_ERR_ARG;
  }
  *ncx = mx;
  *ncy = my;
  *ncz = mz;
  return MFS_OK;
}

extern "C" int mfs_msfem_coarse_system(mfs_batch B, const double* Kc, const double* load, int64_t n_fixed,
                                       const int32_t* fixed_nodes, const double* fixed_values, mfs_mat* A_out,
                                       mfs_vec* b_out, mfs_vec* x0_out) {
  MFS_CHECK_ARG(B && Kc && load && A_out && b_out && x0_out, "mfs_msfem_coarse_system: bad argument");
  MFS_CHECK_ARG(n_fixed == 0 || (fixed_nodes && fixed_values), "mfs_msfem_coarse_system: fixed arrays missing");
  int nx, ny, nz;
  std::vector<int> gi;
  MFS_TRY(coarse_grid(B, &nx, &ny, &nz, gi));
  const int sx = nx + 1, sy = ny + 1, sz = nz + 1;
  const int64_t nn = (int64_t)sx * sy * sz;
  for (int64_t k = 0; k < n_fixed; k++)
    MFS_CHECK_ARG(fixed_nodes[k] >= 0 && fixed_nodes[k] < nn, "mfs_msfem_coarse_system: fixed node out of range");
  // 27-point pattern
  Pattern P;
  P.n = nn;
  P.rowptr.assign(nn + 1, 0);
  for (int64_t v = 0; v < nn; v++) {
    const int i = v % sx, j = (v / sx) % sy, k = v / ((int64_t)sx * sy);
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          const int a = i + di, b = j + dj, c = k + dk;
          if (a < 0 || b < 0 || c < 0 || a >= sx || b >= sy || c >= sz) continue;
          P.col.push_back((int32_t)(a + sx * (b + sy * c)));
        }
    P.rowptr[v + 1] = (int32_t)P.col.size();
  }
  std::vector<double> val(9 * P.col.size(), 0.0);
  auto slot = [&](int64_t r, int64_t c) -> int64_t {
    const int32_t* b0 = P.col.data() + P.rowptr[r];
    const int32_t* e0 = P.col.data() + P.rowptr[r + 1];
    return std::lower_bound(b0, e0, (int32_t)c) - P.col.data();
  };
  for (int c = 0; c < B->ncells; c++) {
    int64_t node[8];
    for (int q = 0; q < 8; q++)
      node[q] = (gi[3 * c] + (q & 1)) + sx * ((gi[3 * c + 1] + ((q >> 1) & 1)) + (int64_t)sy * (gi[3 * c + 2] + (q >> 2)));
    const double* K = Kc + (int64_t)c * NB * NB;
    for (int I = 0; I < 8; I++)
      for (int J = 0; J < 8; J++) {
        const int64_t s = slot(node[I], node[J]);
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) val[9 * s + 3 * a + b] += K[(3 * I + a) * NB + 3 * J + b];
      }
  }
  // elimination
  const int64_t N = 3 * nn;
  std::vector<char> fixed(N, 0);
  std::vector<double> g(N, 0.0), rhs(load, load + N), x0(N, 0.0);
  for (int64_t k = 0; k < n_fixed; k++)
    for (int a = 0; a < 3; a++) {
      fixed[3 * (int64_t)fixed_nodes[k] + a] = 1;
      g[3 * (int64_t)fixed_nodes[k] + a] = fixed_values[3 * k + a];
    }
  for (int64_t r = 0; r < nn; r++)
    for (int a = 0; a < 3; a++) {
      const int64_t i = 3 * r + a;
      for (int64_t s = P.rowptr[r]; s < P.rowptr[r + 1]; s++)
        for (int b = 0; b < 3; b++) {
          const int64_t cj = 3 * (int64_t)P.col[s] + b;
          double& v = val[9 * s + 3 * a + b];
          if (fixed[i]) {
            v = (cj == i) ? 1.0 : 0.0;
          } else if (fixed[cj]) {
            rhs[i] -= v * g[cj];
            v = 0.0;
          }
        }
      if (fixed[i]) {
        rhs[i] = g[i];
        x0[i] = g[i];
      }
    }
  mfs_mat A = nullptr;
  mfs_vec b = nullptr, x = nullptr;
  int st = mfs_mat_create_bsr3(B->ctx, nn, P.rowptr.data(), P.col.data(), val.data(), &A);
  if (st == MFS_OK) st = mfs_vec_create(B->ctx, N, MFS_F64, &b);
  if (st == MFS_OK) st = mfs_vec_upload(b, rhs.data());
  if (st == MFS_OK) st = mfs_vec_create(B->ctx, N, MFS_F64, &x);
  if (st == MFS_OK) st = mfs_vec_upload(x, x0.data());
  if (st != MFS_OK) {
    mfs_mat_destroy(A);
    mfs_vec_destroy(b);
    mfs_vec_destroy(x);
    return st;
  }
  *A_out = A;
  *b_out = b;
  *x0_out = x;
  return MFS_OK;
}

// Coarse load of a uniform traction t on the top face z = max of the cell
// grid: load[3*node + d] = sum over fine top triangles (P1, area/3 per
// vertex) of t_d * hat_node, i.e. Phi^T f with the basis traces.
extern "C" int mfs_msfem_top_traction(mfs_batch B, const double* t, double* load) {
  MFS_CHECK_ARG(B && t && load, "mfs_msfem_top_traction: bad argument");
  int nx, ny, nz;
  std::vector<int> gi;
  MFS_TRY(coarse_grid(B, &nx, &ny, &nz, gi));
  const int sx = nx + 1, sy = ny + 1;
  const int64_t nn = (int64_t)sx * sy * (nz + 1);
  std::fill(load, load + 3 * nn, 0.0);
  for (int c = 0; c < B->ncells; c++) {
    if (gi[3 * c + 2] != nz - 1) continue;
    const int n = B->cells[c].n;
    std::vector<double> ft;
    MFS_TRY(cell_face_traces(B, c, ft));
    const double A = 0.5 * (B->h / n) * (B->h / n);  // fine triangle area
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        // the two top triangles of square (i,j) share the (i,j)-(i+1,j+1) diagonal
        const int tri[2][3][2] = {{{i, j}, {i + 1, j}, {i + 1, j + 1}}, {{i, j}, {i, j + 1}, {i + 1, j + 1}}};
        for (int tt = 0; tt < 2; tt++)
          for (int v = 0; v < 3; v++) {
            double f[NB][3];
            boundary_functions(B, ft, n, tri[tt][v][0], tri[tt][v][1], n, f);
            for (int q = 4; q < 8; q++) {  // top coarse nodes of the cell
              const int64_t node =
                  (gi[3 * c] + (q & 1)) + sx * ((gi[3 * c + 1] + ((q >> 1) & 1)) + (int64_t)sy * nz);
              for (int d = 0; d < 3; d++)
                load[3 * node + d] +=
                    (t[0] * f[3 * q + d][0] + t[1] * f[3 * q + d][1] + t[2] * f[3 * q + d][2]) * A / 3.0;
            }
          }
      }
  }
  return MFS_OK;
}

// Fine displacement of one cell from coarse nodal values (3 per coarse
// node): u = sum_{q,d} u_c[node(q)][d] phi_{q,d}, the downscaling the
// Mechanical driver applies (MechanicalMain.cpp:2176-2181).
// u_fine of one cell from the coarse nodal values, the cell grid positions
// gi known (coarse_grid).
static int downscale_cell(mfs_batch B, const double* u_coarse, int cell, int nx, int ny,
                          const std::vector<int>& gi, double* u_fine) {
  const int sx = nx + 1, sy = ny + 1;
  const int n = B->cells[cell].n;
  const int64_t N = 3ll * (n + 1) * (n + 1) * (n + 1);
  std::vector<double> basis((size_t)NB * N);
  MFS_TRY(mfs_batch_get_basis(B, cell, basis.data()));
  std::fill(u_fine, u_fine + N, 0.0);
  for (int q = 0; q < 8; q++) {
    const int64_t node = (gi[3 * cell] + (q & 1)) + sx * ((gi[3 * cell + 1] + ((q >> 1) & 1)) +
                                                         (int64_t)sy * (gi[3 * cell + 2] + (q >> 2)));
    for (int d = 0; d < 3; d++) {
      const double w = u_coarse[3 * node + d];
      if (w == 0.0) continue;
      const double* ph = basis.data() + (size_t)(3 * q + d) * N;
      for (int64_t e = 0; e < N; e++) u_fine[e] += w * ph[e];
    }
  }
  return MFS_OK;
}

extern "C" int mfs_msfem_downscale(mfs_batch B, const double* u_coarse, int cell, double* u_fine) {
  MFS_CHECK_ARG(B && u_coarse && u_fine && cell >= 0 && cell < B->ncells, "mfs_msfem_downscale: bad argument");
  int nx, ny, nz;
  std::vector<int> gi;
  MFS_TRY(coarse_grid(B, &nx, &ny, &nz, gi));
  return downscale_cell(B, u_coarse, cell, nx, ny, gi, u_fine);
}

// MsFEM-vs-FEM relative error per component (MechanicalMain.cpp:2112-2221):
// one-point rule per fine tet (centroid, weight V), the MsFEM field from the
// downscaled cell solution and the FEM field from the global fine grid.
extern "C" int mfs_msfem_error(mfs_batch B, const double* u_coarse, const double* u_fem, double* err3) {
  MFS_CHECK_ARG(B && u_coarse && u_fem && err3, "mfs_msfem_error: bad argument");
  const int n = batch_uniform_n(B);
  MFS_CHECK_ARG(n > 0, "mfs_msfem_error: all cells must have the same subdivision");
  int nx, ny, nz;
  std::vector<int> gi;
  MFS_TRY(coarse_grid(B, &nx, &ny, &nz, gi));
  const int64_t gx = (int64_t)nx * n + 1, gy = (int64_t)ny * n + 1;
  const int s = n + 1;
  const double hf = B->h / n, V = hf * hf * hf / 6.0;
  double num[3] = {0, 0, 0}, den[3] = {0, 0, 0};
  std::vector<double> uf(3ull * s * s * s);
  for (int c = 0; c < B->ncells; c++) {
    MFS_TRY(downscale_cell(B, u_coarse, c, nx, ny, gi, uf.data()));
    for (int t = 0; t < 6 * n * n * n; t++) {
      const int hex = t / 6, p = t % 6;
      int cc[3] = {hex % n, (hex / n) % n, hex / (n * n)};
      static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
      double ms[3] = {0, 0, 0}, fe[3] = {0, 0, 0};
      for (int v = 0; v < 4; v++) {
        if (v > 0) cc[perm[p][v - 1]]++;
        const int64_t loc = cc[0] + s * (cc[1] + (int64_t)s * cc[2]);
        const int64_t glob = (gi[3 * c] * (int64_t)n + cc[0]) +
                             gx * ((gi[3 * c + 1] * (int64_t)n + cc[1]) + gy * (gi[3 * c + 2] * (int64_t)n + cc[2]));
        for (int a = 0; a < 3; a++) {
          ms[a] += 0.25 * uf[3 * loc + a];
          fe[a] += 0.25 * u_fem[3 * glob + a];
        }
      }
      for (int a = 0; a < 3; a++) {
        num[a] += V * (ms[a] - fe[a]) * (ms[a] - fe[a]);
        den[a] += V * fe[a] * fe[a];
      }
    }
  }
  for (int a = 0; a < 3; a++) {
    const double sd = std::sqrt(den[a]);
    err3[a] = std::sqrt(num[a]) / (sd < 1e-10 ? 1.0 : sd);
  }
  return MFS_OK;
}

extern "C" int mfs_batch_status(mfs_batch B, int32_t* iters, double* relres, int32_t* flags) {
  MFS_CHECK_ARG(B, "mfs_batch_status: NULL batch");
  hipStream_t s = B->ctx->stream;
  const size_t nq = (size_t)NB * B->ncells;
  if (iters) MFS_HIP(hipMemcpyAsync(iters, B->d_iters, sizeof(int) * nq, hipMemcpyDeviceToHost, s));
  if (relres) MFS_HIP(hipMemcpyAsync(relres, B->d_relres, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
  if (flags) MFS_HIP(hipMemcpyAsync(flags, B->d_flags, sizeof(int) * nq, hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  return MFS_OK;
}

// Multi-right-hand-side Jacobi-PCG on gfx950: up to 8 systems A x_j = b_j
// that share A (the homogenisation driver's 6 load cases,
// Mechanical/MechanicalMain.cpp:243-420, which solve one CSSD system per
// STAGE with identical rows and different right-hand sides) run as one
// iteration stream.  Each iteration reads the matrix once for all of them
// (SELL SpMM); the per-system scalars (alpha, beta, norms, stop flags) are
// vectors of the same length and a finished system is frozen (alpha = 0)
// while the others continue.
//
// Layout: the vectors of the KR systems are interleaved, element (i, j) at
// v[i*KR + j], so the operand gather of one column is one contiguous 16-,
// 24- or 32-byte read (KR = 4, 6 or 8 FP32 systems).  MIXED precision runs the same FP64 outer refinement as the
// single-system solver (true residual, unit scaling, x += scale d per
// system) around the FP32 iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mfs_device.hpp"

using namespace mfs;

namespace {

constexpr int MR_MAX = 8;

struct MrState {
  double rz1[MR_MAX], rz2[MR_MAX], rr[MR_MAX], pq[MR_MAX], tol2[MR_MAX];
  double rzkeep[MR_MAX];  // r.z of the iteration before a system stopped
  int done[MR_MAX];       // written by block 0 of K1, read by K2
  int it_done[MR_MAX];    // iteration at which a system stopped (block 0 of K1 only)
  int warm[MR_MAX];       // 1: the first iteration continues a kept direction
  int iter1, iter2, maxit, alldone;
};

enum { MR_RUN = 0, MR_CONV = 1, MR_BREAK = 2, MR_MAXIT = 3 };

// Sum of NQ values' per-block partials part[q*G + b] in block order, done by
// every block of the consuming kernel (identical totals).  The first P
// partials of every value are loaded together (NQ*P loads in flight per
// thread; G <= 256*P needs no more), so the NQ sums cost one memory latency
// instead of one each.
template <int NQ, int P>
__device__ __forceinline__ void mr_sum_partials(const double* __restrict__ part, unsigned G, double (&tot)[NQ]) {
  __shared__ double red[NQ][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double pre[NQ][P];
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int u = 0; u < P; u++) {
      const unsigned b = threadIdx.x + u * 256;
      pre[q][u] = b < G ? part[(size_t)q * G + b] : 0.0;
    }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    double s = pre[q][0];
#pragma unroll
    for (int u = 1; u < P; u++) s += pre[q][u];
    for (unsigned b = threadIdx.x + P * 256; b < G; b += 256) s += part[(size_t)q * G + b];
    s = wave_sum(s);
    if (lane == 0) red[q][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; q++) tot[q] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  __syncthreads();
}

// Per-block partials of NQ values into part[q*gridDim + blockIdx].
template <int NQ>
__device__ __forceinline__ void mr_block_partials(double (&v)[NQ], double* __restrict__ part) {
  __shared__ double red[NQ][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const double s = wave_sum(v[q]);
    if (lane == 0) red[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NQ)
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
        (red[threadIdx.x][0] + red[threadIdx.x][1]) + (red[threadIdx.x][2] + red[threadIdx.x][3]);
}

// The KR values of one element move as 16-byte vectors when KR*sizeof(T)
// is a multiple of 16 (elements then 16-byte aligned), else as 8-byte ones
// (6 FP32 systems: 24-byte elements, 8-byte aligned).
template <typename T, int KR>
struct VecWidth {
  static_assert((KR * sizeof(T)) % 8 == 0, "interleaved element must be whole 8-byte words");
  static constexpr int B = (KR * sizeof(T)) % 16 == 0 ? 16 : 8;  // bytes per load
  static constexpr int NV = KR * sizeof(T) / B;                   // loads per element
};
template <int B>
struct VecWord;
template <>
struct VecWord<16> {
  using type = uint4;
};
template <>
struct VecWord<8> {
  using type = uint2;
};

template <typename T, int KR>
__device__ __forceinline__ void vload(const T* __restrict__ p, T (&o)[KR]) {
  constexpr int B = VecWidth<T, KR>::B, NV = VecWidth<T, KR>::NV;
  using W = typename VecWord<B>::type;
  const W* q = reinterpret_cast<const W*>(p);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const W u = q[v];
    __builtin_memcpy(&o[v * (B / sizeof(T))], &u, B);
  }
}
template <typename T, int KR>
__device__ __forceinline__ void vstore(T* __restrict__ p, const T (&o)[KR]) {
  constexpr int B = VecWidth<T, KR>::B, NV = VecWidth<T, KR>::NV;
  using W = typename VecWord<B>::type;
  W* q = reinterpret_cast<W*>(p);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    W u;
    __builtin_memcpy(&u, &o[v * (B / sizeof(T))], B);
    q[v] = u;
  }
}

// SELL-64 block-row product for KR interleaved operands: acc[a][j] =
// sum over the row's blocks of val(a, b) * X_j(BS*col + b), X supplied by
// xl(i, out[KR]).  Indices and values of U entries are loaded before their
// operands are gathered (see sell_row_x).
template <typename T, int BS, int KR, typename XL, typename CI>
__device__ __forceinline__ void sell_row_multi(const int64_t* __restrict__ slice_base,
                                               const int32_t* __restrict__ slice_width,
                                               const CI* __restrict__ col, const T* __restrict__ val,
                                               int64_t slice, int lane, XL xl, T (&acc)[BS][KR]) {
  constexpr int C = 64;
  constexpr int U = 2;
  constexpr int B2 = BS * BS;
#pragma unroll
  for (int a = 0; a < BS; a++)
#pragma unroll
    for (int j = 0; j < KR; j++) acc[a][j] = 0;
  const int64_t sb = slice_base[slice];
  const int w = slice_width[slice];
  const CI* cp = col + sb + lane;
  const int32_t rbase = sizeof(CI) == 2 ? (int32_t)(slice * C + lane) : 0;  // row-relative 16-bit columns
  const T* vp = val + sb * B2 + lane;
  for (int j0 = 0; j0 < w; j0 += U) {
    int32_t c[U];
    T v[U][B2];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int jj = min(j0 + u, w - 1);
      const bool in = j0 + u < w;
      c[u] = (int32_t)__builtin_nontemporal_load(cp + (int64_t)jj * C) + rbase;
#pragma unroll
      for (int e = 0; e < B2; e++) {
        const T t = __builtin_nontemporal_load(vp + ((int64_t)jj * B2 + e) * C);
        v[u][e] = in ? t : (T)0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
#pragma unroll
      for (int b = 0; b < BS; b++) {
        T xv[KR];
        xl((int64_t)BS * c[u] + b, xv);
#pragma unroll
        for (int a = 0; a < BS; a++)
#pragma unroll
          for (int j = 0; j < KR; j++) acc[a][j] += v[u][BS * a + b] * xv[j];
      }
    }
  }
}

// R = B - A X (precision T), partials of r.r and b.b per system.
template <typename T, int BS, int KR>
__global__ __launch_bounds__(256) void k_mr_residual(const int64_t* __restrict__ slice_base,
                                                     const int32_t* __restrict__ slice_width,
                                                     const int32_t* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ X, const T* __restrict__ Bv,
                                                     T* __restrict__ R, int64_t nb, int64_t n_slices,
                                                     double* __restrict__ part, int xcd) {
  const int64_t slice = (int64_t)xcd_block(xcd) * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const int64_t row = slice * 64 + lane;
  T acc[BS][KR];
  if (slice < n_slices) {
    sell_row_multi<T, BS, KR>(slice_base, slice_width, col, val, slice, lane,
                              [&](int64_t i, T (&o)[KR]) { vload<T, KR>(X + i * KR, o); },
                              acc);
  } else {
#pragma unroll
    for (int a = 0; a < BS; a++)
#pragma unroll
      for (int j = 0; j < KR; j++) acc[a][j] = 0;
  }
  double red[2 * KR];
#pragma unroll
  for (int q = 0; q < 2 * KR; q++) red[q] = 0;
  if (row < nb) {
#pragma unroll
    for (int a = 0; a < BS; a++) {
      const int64_t e0 = (BS * row + a) * KR;
      T bv[KR], rv[KR];
      vload<T, KR>(Bv + e0, bv);
#pragma unroll
      for (int j = 0; j < KR; j++) {
        rv[j] = bv[j] - acc[a][j];
        red[j] += (double)rv[j] * (double)rv[j];
        red[KR + j] += (double)bv[j] * (double)bv[j];
      }
      vstore<T, KR>(R + e0, rv);
    }
  }
  mr_block_partials<2 * KR>(red, part);
}

// Start of an inner run: Z = D^-1 R, P0 = 0, X = 0 (correction), partials
// r.z and r.r per system.
// P0 = ratio_j * P_src (a system's kept direction, rescaled to the units of
// its new residual) or 0 where ratio_j = 0; P_src may alias P0.
template <typename T, int KR>
__global__ __launch_bounds__(256) void k_mr_start(T* __restrict__ X, const T* __restrict__ R, T* __restrict__ Z,
                                                  T* P0, const T* P_src, const double* __restrict__ ratio,
                                                  const T* __restrict__ dinv, int64_t n, double* __restrict__ part) {
  T rat[KR];
#pragma unroll
  for (int j = 0; j < KR; j++) rat[j] = (T)ratio[j];
  double red[2 * KR];
#pragma unroll
  for (int q = 0; q < 2 * KR; q++) red[q] = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T d = dinv ? dinv[i] : (T)1;
    T r[KR], z[KR], zero[KR], ps[KR], p0[KR];
    vload<T, KR>(R + i * KR, r);
    vload<T, KR>(P_src + i * KR, ps);
#pragma unroll
    for (int j = 0; j < KR; j++) {
      z[j] = d * r[j];
      zero[j] = 0;
      p0[j] = rat[j] != (T)0 ? rat[j] * ps[j] : (T)0;
      red[j] += (double)(r[j] * z[j]);
      red[KR + j] += (double)(r[j] * r[j]);
    }
    vstore<T, KR>(Z + i * KR, z);
    vstore<T, KR>(P0 + i * KR, p0);
    vstore<T, KR>(X + i * KR, zero);
  }
  mr_block_partials<2 * KR>(red, part);
}

// tol2[0..8): squared tolerances; tol2[8..16): ratio (0 = fresh start);
// tol2[16..24): r.z before the kept direction, already in the new units
__global__ void k_mr_state(MrState* st, const double* tol2, int k, int maxit) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int j = 0; j < MR_MAX; j++) {
      const bool warm = j < k && tol2[MR_MAX + j] > 0.0;
      st->warm[j] = warm;
      st->rz1[j] = 1.0;
      st->rz2[j] = warm ? tol2[2 * MR_MAX + j] : 1.0;
      st->rzkeep[j] = 0.0;
      st->rr[j] = 0.0;
      st->pq[j] = 1.0;
      // padding systems (j >= k) count as converged from the start
      st->tol2[j] = j < k ? tol2[j] : 1e300;
      st->done[j] = j < k ? MR_RUN : MR_CONV;
      st->it_done[j] = 0;
    }
    st->iter1 = st->iter2 = 0;
    st->maxit = maxit;
    st->alldone = 0;
  }
}

// K1: stop tests and beta per system from K2's partials; P_new = Z + beta
// P_old (own rows written, gathered rows formed on the fly); Q = A P_new;
// partials p.q per system.
template <typename T, int BS, int KR, typename CI>
__global__ __launch_bounds__(256) void k_mr_dir_spmm(const int64_t* __restrict__ slice_base,
                                                     const int32_t* __restrict__ slice_width,
                                                     const CI* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ Z, const T* __restrict__ P_old,
                                                     T* __restrict__ P_new, T* __restrict__ Q, int64_t nb,
                                                     int64_t n_slices, MrState* st, const double* __restrict__ part_rz,
                                                     unsigned g_rz, double* __restrict__ part_pq, int xcd) {
  if (st->alldone) return;
  double t[2 * KR];
  mr_sum_partials<2 * KR, 4>(part_rz, g_rz, t);
  const int it = st->iter2;
  T beta[KR];
  int code[KR];
  int all = 1;
#pragma unroll
  for (int j = 0; j < KR; j++) {
    const double rz = t[j], rr = t[KR + j];
    int c = MR_RUN;
    if (!(rr == rr) || !(rz == rz)) c = MR_BREAK;
    else if (rr <= st->tol2[j]) c = MR_CONV;
    else if (it >= st->maxit) c = MR_MAXIT;
    code[j] = c;
    all &= c != MR_RUN;
    beta[j] = (c == MR_RUN && (it > 0 || st->warm[j])) ? (T)(rz / st->rz2[j]) : (T)0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int j = 0; j < KR; j++) {
      if (code[j] != MR_RUN && st->done[j] == MR_RUN) {
        st->it_done[j] = it;
        st->rzkeep[j] = st->rz2[j];
      }
      st->done[j] = code[j];
      st->rz1[j] = t[j];
      st->rr[j] = t[KR + j];
    }
    st->iter1 = it;
    st->alldone = all;
  }
  if (all) return;
  const int64_t slice = (int64_t)xcd_block(xcd) * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const int64_t row = slice * 64 + lane;
  T acc[BS][KR];
  if (slice < n_slices) {
    sell_row_multi<T, BS, KR>(slice_base, slice_width, col, val, slice, lane,
                              [&](int64_t i, T (&o)[KR]) {
                                T zz[KR], pp[KR];
                                vload<T, KR>(Z + i * KR, zz);
                                vload<T, KR>(P_old + i * KR, pp);
#pragma unroll
                                for (int j = 0; j < KR; j++) o[j] = zz[j] + beta[j] * pp[j];
                              },
                              acc);
  } else {
#pragma unroll
    for (int a = 0; a < BS; a++)
#pragma unroll
      for (int j = 0; j < KR; j++) acc[a][j] = 0;
  }
  double red[KR];
#pragma unroll
  for (int j = 0; j < KR; j++) red[j] = 0;
  if (row < nb) {
#pragma unroll
    for (int a = 0; a < BS; a++) {
      const int64_t e0 = (BS * row + a) * KR;
      T zz[KR], pp[KR], pn[KR];
      vload<T, KR>(Z + e0, zz);
      vload<T, KR>(P_old + e0, pp);
#pragma unroll
      for (int j = 0; j < KR; j++) {
        // a stopped system keeps its last direction (for a continuation)
        pn[j] = code[j] == MR_RUN ? zz[j] + beta[j] * pp[j] : pp[j];
        red[j] += (double)(acc[a][j] * pn[j]);
      }
      vstore<T, KR>(P_new + e0, pn);
      vstore<T, KR>(Q + e0, acc[a]);
    }
  }
  mr_block_partials<KR>(red, part_pq);
}

// K2: alpha per system (0 for stopped systems); X += alpha P, R -= alpha Q,
// Z = D^-1 R; partials r.z and r.r.
template <typename T, int KR>
__global__ __launch_bounds__(256) void k_mr_update(T* __restrict__ X, const T* __restrict__ P, T* __restrict__ R,
                                                   T* __restrict__ Z, const T* __restrict__ Q,
                                                   const T* __restrict__ dinv, int64_t n, MrState* st,
                                                   const double* __restrict__ part_pq, unsigned g_pq,
                                                   double* __restrict__ part_rz, int xcd, unsigned g1, int bs) {
  if (st->alldone) return;
  double pq[KR];
  mr_sum_partials<KR, 5>(part_pq, g_pq, pq);
  T alpha[KR];
#pragma unroll
  for (int j = 0; j < KR; j++) {
    const bool ok = st->done[j] == MR_RUN && pq[j] != 0.0 && isfinite(pq[j]);
    alpha[j] = ok ? (T)(st->rz1[j] / pq[j]) : (T)0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int j = 0; j < KR; j++) {
      st->rz2[j] = st->rz1[j];
      st->pq[j] = pq[j];
    }
    st->iter2 = st->iter1 + 1;
  }
  double red[2 * KR];
#pragma unroll
  for (int q = 0; q < 2 * KR; q++) red[q] = 0;
  // the rows that this XCD's blocks of k_mr_dir_spmm gather next
  int64_t e0, e1;
  update_range(n, xcd, g1, bs, 64, e0, e1);
  for (int64_t i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
    const T d = dinv ? dinv[i] : (T)1;
    T xx[KR], pp[KR], rr[KR], qq[KR], zz[KR];
    vload<T, KR>(X + i * KR, xx);
    vload<T, KR>(P + i * KR, pp);
    vload<T, KR>(R + i * KR, rr);
    vload<T, KR>(Q + i * KR, qq);
#pragma unroll
    for (int j = 0; j < KR; j++) {
      xx[j] += alpha[j] * pp[j];
      rr[j] -= alpha[j] * qq[j];
      zz[j] = d * rr[j];
      red[j] += (double)(rr[j] * zz[j]);
      red[KR + j] += (double)(rr[j] * rr[j]);
    }
    vstore<T, KR>(X + i * KR, xx);
    vstore<T, KR>(R + i * KR, rr);
    vstore<T, KR>(Z + i * KR, zz);
  }
  mr_block_partials<2 * KR>(red, part_rz);
}

// interleave / de-interleave one system's FP64 vector
__global__ void k_mr_pack(const double* __restrict__ v, double* __restrict__ V, int j, int kr, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    V[i * kr + j] = v[i];
}
__global__ void k_mr_unpack(const double* __restrict__ V, double* __restrict__ v, int j, int kr, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = V[i * kr + j];
}
// X64 += X (FP64 corrections of all systems)
__global__ void k_mr_add(double* __restrict__ X64, const double* __restrict__ X, int64_t nk) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nk; i += (int64_t)gridDim.x * blockDim.x)
    X64[i] += X[i];
}
// R32 = R64 * inv_scale (per system)
template <int KR>
__global__ void k_mr_demote(const double* __restrict__ R64, const double* __restrict__ inv_scale,
                            float* __restrict__ R32, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int j = 0; j < KR; j++) R32[i * KR + j] = (float)(R64[i * KR + j] * inv_scale[j]);
}
// X64 += scale * X32 (per system)
template <int KR>
__global__ void k_mr_promote(double* __restrict__ X64, const double* __restrict__ scale, const float* __restrict__ X32,
                             int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int j = 0; j < KR; j++) X64[i * KR + j] += scale[j] * (double)X32[i * KR + j];
}

inline unsigned mr_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256)); }

// Inner iteration engine in precision T for KR interleaved systems.
template <typename T, int KR>
struct MrEngine {
  mfs_mat A = nullptr;
  int64_t n = 0;
  T *X = nullptr, *R = nullptr, *Z = nullptr, *Q = nullptr, *P[2] = {nullptr, nullptr};
  const T* val = nullptr;
  const T* dinv = nullptr;
  MrState* st = nullptr;
  double *part_rz = nullptr, *part_pq = nullptr, *dtol = nullptr;
  unsigned g_spmv = 0, g_vec = 0;
  // large systems: the partials are added once by k_finish_sums
  // (tot_rz: 2*KR values, tot_pq: KR) instead of by every consumer block
  bool fin = false;
  double *tot_rz = nullptr, *tot_pq = nullptr;
  int chunk = 16;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};  // end of the chunks in the two status slots

  int init(mfs_mat A_, bool precond) {
    A = A_;
    n = A->n();
    const bool f64 = std::is_same<T, double>::value;
    if (!f64 && !A->val32_valid) MFS_TRY(mat_refresh_derived(A));
    val = f64 ? (const T*)A->d_val64 : (const T*)A->d_val32;
    dinv = precond ? (f64 ? (const T*)A->d_dinv64 : (const T*)A->d_dinv32) : nullptr;
    for (T** b : {&X, &R, &Z, &Q, &P[0], &P[1]}) MFS_TRY(dmalloc((void**)b, sizeof(T) * n * KR));
    MFS_TRY(dmalloc((void**)&st, sizeof(MrState)));
    g_spmv = (unsigned)spmv_grid(A);
    g_vec = mr_grid(n);
    MFS_TRY(dmalloc((void**)&part_rz, sizeof(double) * 2 * KR * std::max(g_spmv, g_vec)));
    MFS_TRY(dmalloc((void**)&part_pq, sizeof(double) * KR * g_spmv));
    MFS_TRY(dmalloc((void**)&dtol, sizeof(double) * 3 * MR_MAX));
    fin = (uint64_t)g_spmv * g_vec * KR > 1000000ull;
    if (const char* e = std::getenv("MFS_FINISH_SUMS")) fin = e[0] == '1';  // force either way (tests)
    MFS_TRY(dmalloc((void**)&tot_rz, sizeof(double) * 2 * KR));
    MFS_TRY(dmalloc((void**)&tot_pq, sizeof(double) * KR));
    // MFS_NO_GRAPH=1 launches the same kernels directly (kernel tracing)
    const char* e = std::getenv("MFS_NO_GRAPH");
    if (e && e[0] == '1') return MFS_OK;
    hipStream_t s = A->ctx->stream;
    MFS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < chunk; k++) launch_iteration(s, k & 1);
    MFS_HIP(hipStreamEndCapture(s, &graph));
    MFS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return MFS_OK;
  }

  void launch_chunk(hipStream_t s) {
    if (exec) {
      (void)hipGraphLaunch(exec, s);
      return;
    }
    for (int k = 0; k < chunk; k++) launch_iteration(s, k & 1);
  }

  void launch_iteration(hipStream_t s, int parity) {
    const int64_t ns = A->sell.n_slices;
    T* p_old = P[parity];
    T* p_new = P[parity ^ 1];
    const double* rz_in = fin ? tot_rz : part_rz;
    const unsigned rz_g = fin ? 1u : g_vec;
    // 16-bit row-relative columns when the matrix has them
    if (A->d_col16 && A->bs == 1)
      k_mr_dir_spmm<T, 1, KR, int16_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col16, val,
                                                              Z, p_old, p_new, Q, A->nb, ns, st, rz_in, rz_g,
                                                              part_pq, A->xcd_map);
    else if (A->d_col16)
      k_mr_dir_spmm<T, 3, KR, int16_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col16, val,
                                                              Z, p_old, p_new, Q, A->nb, ns, st, rz_in, rz_g,
                                                              part_pq, A->xcd_map);
    else if (A->bs == 1)
      k_mr_dir_spmm<T, 1, KR, int32_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, val, Z,
                                                              p_old, p_new, Q, A->nb, ns, st, rz_in, rz_g, part_pq,
                                                              A->xcd_map);
    else
      k_mr_dir_spmm<T, 3, KR, int32_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, val, Z,
                                                              p_old, p_new, Q, A->nb, ns, st, rz_in, rz_g, part_pq,
                                                              A->xcd_map);
    if (fin) k_finish_sums<<<KR, 256, 0, s>>>(part_pq, g_spmv, tot_pq);
    k_mr_update<T, KR><<<g_vec, 256, 0, s>>>(X, p_new, R, Z, Q, dinv, n, st, fin ? tot_pq : part_pq,
                                             fin ? 1u : g_spmv, part_rz, A->xcd_map, g_spmv, A->bs);
    finish_rz(s);
  }

  void finish_rz(hipStream_t s) {
    if (fin) k_finish_sums<<<2 * KR, 256, 0, s>>>(part_rz, g_vec, tot_rz);
  }

  // continuation data of the last run, per system
  int last_iter = 0;
  double keep_rz[MR_MAX] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool can_keep[MR_MAX] = {false, false, false, false, false, false, false, false};

  // R holds the residuals; X receives the corrections.  tol2[j] absolute
  // squared tolerances of the k live systems; ratio[j] > 0 continues system
  // j's kept direction (its residual is the old recursive one times ratio),
  // 0 starts it fresh.
  int run(int k, const double* tol2, const double* ratio, int maxit, MrState* out) {
    hipStream_t s = A->ctx->stream;
    double* hp = (double*)A->ctx->pinned;
    for (int j = 0; j < MR_MAX; j++) {
      const bool warm = j < k && ratio && ratio[j] > 0.0 && can_keep[j];
      hp[j] = j < k ? tol2[j] : 1e300;
      hp[MR_MAX + j] = warm ? ratio[j] : 0.0;
      hp[2 * MR_MAX + j] = warm ? keep_rz[j] * ratio[j] * ratio[j] : 0.0;
    }
    MFS_HIP(hipMemcpyAsync(dtol, hp, sizeof(double) * 3 * MR_MAX, hipMemcpyHostToDevice, s));
    k_mr_start<T, KR><<<g_vec, 256, 0, s>>>(X, R, Z, P[0], P[last_iter & 1], dtol + MR_MAX, dinv, n, part_rz);
    finish_rz(s);
    k_mr_state<<<1, 64, 0, s>>>(st, dtol, k, maxit);
    MFS_HIP(hipGetLastError());
    MFS_HIP(hipStreamSynchronize(s));
    // chunks of iterations, one queued ahead while some running system's
    // residual is far above its tolerance (as in the single-system engine)
    MrState* hs = (MrState*)A->ctx->pinned;
    for (hipEvent_t* e : {&ev[0], &ev[1]})
      if (!*e) MFS_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    auto submit = [&](int slot) -> int {
      launch_chunk(s);
      MFS_HIP(hipGetLastError());
      MFS_HIP(hipMemcpyAsync(hs + slot, st, sizeof(MrState), hipMemcpyDeviceToHost, s));
      MFS_HIP(hipEventRecord(ev[slot], s));
      return MFS_OK;
    };
    int head = 0, inflight = 1;
    bool ahead = false;
    MFS_TRY(submit(0));
    for (;;) {
      if (ahead && inflight == 1) {
        MFS_TRY(submit(head ^ 1));
        inflight = 2;
      }
      MFS_HIP(hipEventSynchronize(ev[head]));
      inflight--;
      if (hs[head].alldone) break;
      ahead = false;
      for (int j = 0; j < MR_MAX; j++)
        ahead |= hs[head].done[j] == MR_RUN && hs[head].rr[j] > 100.0 * hs[head].tol2[j];
      head ^= 1;
      if (inflight == 0) {
        MFS_TRY(submit(head));
        inflight = 1;
      }
    }
    *out = hs[head];
    hs = out;
    if (inflight) MFS_HIP(hipStreamSynchronize(s));
    // every system's newest direction is in P[iter & 1] (stopped systems
    // copied theirs through the later iterations)
    last_iter = hs->iter1;
    for (int j = 0; j < MR_MAX; j++) {
      keep_rz[j] = hs->rzkeep[j];
      can_keep[j] = j < k && hs->done[j] == MR_CONV && hs->it_done[j] > 0 && std::isfinite(keep_rz[j]) &&
                    keep_rz[j] > 0.0;
    }
    return MFS_OK;
  }

  void destroy() {
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e);
    ev[0] = ev[1] = nullptr;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (T* b : {X, R, Z, Q, P[0], P[1]}) dfree(b);
    dfree(st);
    dfree(part_rz);
    dfree(part_pq);
    dfree(dtol);
    dfree(tot_rz);
    dfree(tot_pq);
  }
};

// FP64 residuals of the interleaved systems: R64 = B64 - A X64, r.r and b.b.
template <int KR>
int mr_true_residual(mfs_mat A, const double* X64, const double* B64, double* R64, double* part, double* rr,
                     double* bb) {
  hipStream_t s = A->ctx->stream;
  const int64_t ns = A->sell.n_slices;
  const unsigned g = (unsigned)spmv_grid(A);
  if (A->bs == 1)
    k_mr_residual<double, 1, KR><<<g, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, A->d_val64, X64, B64,
                                                   R64, A->nb, ns, part, A->xcd_map);
  else
    k_mr_residual<double, 3, KR><<<g, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, A->d_val64, X64, B64,
                                                   R64, A->nb, ns, part, A->xcd_map);
  MFS_HIP(hipGetLastError());
  std::vector<double> hpart((size_t)2 * KR * g);
  MFS_HIP(hipMemcpyAsync(hpart.data(), part, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < KR; q++) {
    double a = 0, b = 0;
    for (unsigned i = 0; i < g; i++) {
      a += hpart[(size_t)q * g + i];
      b += hpart[(size_t)(KR + q) * g + i];
    }
    rr[q] = a;
    bb[q] = b;
  }
  return MFS_OK;
}

template <int KR>
int solve_multi(mfs_mat A, const mfs_solve_desc& d, int k, mfs_vec* b, mfs_vec* x, mfs_solve_report* reps) {
  hipStream_t s = A->ctx->stream;
  const int64_t n = A->n();
  const bool precond = d.method == MFS_PCG_JACOBI;
  const unsigned gv = mr_grid(n);
  auto t0 = std::chrono::steady_clock::now();
  double *X64 = nullptr, *B64 = nullptr, *R64 = nullptr, *part = nullptr, *dsc = nullptr;
  int st = dmalloc((void**)&X64, sizeof(double) * n * KR);
  if (st == MFS_OK) st = dmalloc((void**)&B64, sizeof(double) * n * KR);
  if (st == MFS_OK) st = dmalloc((void**)&R64, sizeof(double) * n * KR);
  if (st == MFS_OK) st = dmalloc((void**)&part, sizeof(double) * 2 * KR * spmv_grid(A));
  if (st == MFS_OK) st = dmalloc((void**)&dsc, sizeof(double) * 2 * KR);
  MrEngine<double, KR> E64;
  MrEngine<float, KR> E32;
  const bool mixed = d.precision == MFS_PREC_MIXED;
  if (st == MFS_OK) st = mixed ? E32.init(A, precond) : E64.init(A, precond);
  auto cleanup = [&]() {
    dfree(X64);
    dfree(B64);
    dfree(R64);
    dfree(part);
    dfree(dsc);
    if (mixed) E32.destroy();
    else E64.destroy();
  };
  if (st != MFS_OK) {
    cleanup();
    return st;
  }
  // interleave; padding systems are zero (converged from the start)
  MFS_HIP(hipMemsetAsync(X64, 0, sizeof(double) * n * KR, s));
  MFS_HIP(hipMemsetAsync(B64, 0, sizeof(double) * n * KR, s));
  for (int j = 0; j < k; j++) {
    k_mr_pack<<<gv, 256, 0, s>>>((const double*)x[j]->d, X64, j, KR, n);
    k_mr_pack<<<gv, 256, 0, s>>>((const double*)b[j]->d, B64, j, KR, n);
  }
  double rr[KR], bb[KR];
  int outer = 0, iters = 0;
  std::vector<int> it_sys(k, 0), status(k, MFS_OK);
  std::vector<double> rel(k), recursive(k, 0.0);
  int rc = mr_true_residual<KR>(A, X64, B64, R64, part, rr, bb);
  auto relres = [&](int j) { return std::sqrt(rr[j] / (bb[j] > 0 ? bb[j] : 1.0)); };
  if (rc == MFS_OK) {
    if (!mixed) {
      // FP64: one inner run on the interleaved residuals, corrections into X64
      double tol2[MR_MAX];
      for (int j = 0; j < k; j++) tol2[j] = d.tol * d.tol * (bb[j] > 0 ? bb[j] : 1.0);
      MFS_HIP(hipMemcpyAsync(E64.R, R64, sizeof(double) * n * KR, hipMemcpyDeviceToDevice, s));
      MrState hs;
      rc = E64.run(k, tol2, nullptr, d.maxit, &hs);
      if (rc == MFS_OK) {
        k_mr_add<<<gv, 256, 0, s>>>(X64, E64.X, n * KR);
        for (int j = 0; j < k; j++) {
          it_sys[j] = hs.done[j] == MR_RUN ? hs.iter1 : hs.it_done[j];
          recursive[j] = std::sqrt(hs.rr[j] / (bb[j] > 0 ? bb[j] : 1.0));
          if (hs.done[j] == MR_BREAK) status[j] = MFS_ERR_NONFINITE;
        }
        iters = hs.iter1;
        outer = 1;
        rc = mr_true_residual<KR>(A, X64, B64, R64, part, rr, bb);
      }
    } else {
      // FP64 refinement around the FP32 interleaved iteration
      // continuation per system, as in the single-system solver: when the
      // FP64 residual after a step is within 2x of what the FP32 recursion
      // predicted, the next step keeps the system's direction
      std::vector<double> prev(k, 1e300), keep(MR_MAX, 0.0), pred(k, 0.0), scale_prev(k, 0.0);
      std::vector<bool> live(k, true);
      while (rc == MFS_OK && outer < 64) {
        bool any = false;
        double tol2[MR_MAX], sc[2 * KR];
        for (int j = 0; j < KR; j++) {
          sc[j] = 1.0;
          sc[KR + j] = 0.0;
        }
        for (int j = 0; j < k; j++) {
          const double rj = relres(j);
          if (live[j] && (rj <= d.tol || !std::isfinite(rj) || !(rj < 0.99 * prev[j]) || iters >= d.maxit)) {
            if (!std::isfinite(rj)) status[j] = MFS_ERR_NONFINITE;
            live[j] = false;
          }
          prev[j] = rj;
          if (!live[j]) {
            tol2[j] = 1e300;  // frozen: converged from the start of the run
            continue;
          }
          any = true;
          const double rn = std::sqrt(rr[j]);
          sc[j] = rn;
          sc[KR + j] = 1.0 / rn;
          const double itol = std::min(0.5, std::max(d.inner_tol, 0.5 * d.tol / rj));
          tol2[j] = itol * itol;
        }
        if (!any) break;
        // scales and inverse scales to the device; frozen systems get 0
        for (int j = 0; j < k; j++)
          if (!live[j]) sc[j] = sc[KR + j] = 0.0;
        MFS_HIP(hipMemcpyAsync(dsc, sc, sizeof(double) * 2 * KR, hipMemcpyHostToDevice, s));
        k_mr_demote<KR><<<gv, 256, 0, s>>>(R64, dsc + KR, E32.R, n);
        for (int j = 0; j < k; j++) {
          const double rj = relres(j);
          keep[j] = (live[j] && outer > 0 && pred[j] > 0.0 && rj <= 2.0 * pred[j]) ? scale_prev[j] / sc[j] : 0.0;
        }
        MrState hs;
        rc = E32.run(k, tol2, keep.data(), d.maxit - iters, &hs);
        if (rc != MFS_OK) break;
        for (int j = 0; j < k; j++)
          if (live[j]) {
            const int ij = hs.done[j] == MR_RUN ? hs.iter1 : hs.it_done[j];
            it_sys[j] += ij;
            recursive[j] = std::sqrt(hs.rr[j]) * relres(j);
            pred[j] = recursive[j];
            scale_prev[j] = sc[j];
          }
        iters += hs.iter1;
        outer++;
        k_mr_promote<KR><<<gv, 256, 0, s>>>(X64, dsc, E32.X, n);
        rc = mr_true_residual<KR>(A, X64, B64, R64, part, rr, bb);
      }
    }
  }
  if (rc == MFS_OK)
    for (int j = 0; j < k; j++) k_mr_unpack<<<gv, 256, 0, s>>>(X64, (double*)x[j]->d, j, KR, n);
  MFS_HIP(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  const double secs = std::chrono::duration<double>(t1 - t0).count();
  cleanup();
  if (rc != MFS_OK) return rc;
  int worst = MFS_OK;
  for (int j = 0; j < k; j++) {
    rel[j] = relres(j);
    int stj = status[j];
    if (stj == MFS_OK && !(rel[j] <= d.tol)) stj = MFS_ERR_NOT_CONVERGED;
    reps[j].iterations = it_sys[j];
    reps[j].outer_steps = outer;
    reps[j].true_relres = rel[j];
    reps[j].recursive_relres = recursive[j];
    reps[j].seconds = secs;
    reps[j].status = stj;
    if (stj != MFS_OK && worst == MFS_OK) worst = stj;
  }
  if (worst == MFS_ERR_NOT_CONVERGED) set_error("mfs_solve_multi: a system did not reach the tolerance");
  if (worst == MFS_ERR_NONFINITE) set_error("mfs_solve_multi: non-finite value in a system");
  return worst;
}

}  // namespace

extern "C" int mfs_solve_multi(mfs_mat A, const mfs_solve_desc* desc, int k, mfs_vec* b, mfs_vec* x,
                               mfs_solve_report* reports) {
  MFS_CHECK_ARG(A && desc && b && x && reports && k >= 1 && k <= MR_MAX, "mfs_solve_multi: bad argument (k in 1..8)");
  MFS_CHECK_ARG(desc->method == MFS_PCG_JACOBI || desc->method == MFS_CG,
                "mfs_solve_multi: only the CG methods have a multi-system form");
  MFS_CHECK_ARG(desc->precision == MFS_PREC_F64 || desc->precision == MFS_PREC_MIXED,
                "mfs_solve_multi: unknown precision");
  MFS_CHECK_ARG(desc->tol > 0 && desc->maxit > 0, "mfs_solve_multi: tol and maxit must be positive");
  for (int j = 0; j < k; j++)
    MFS_CHECK_ARG(b[j] && x[j] && b[j]->n == A->n() && x[j]->n == A->n() && b[j]->prec == MFS_F64 &&
                      x[j]->prec == MFS_F64,
                  "mfs_solve_multi: b and x must be FP64 vectors of the matrix size");
  MFS_HIP(hipSetDevice(A->ctx->device));
  std::memset(reports, 0, sizeof(mfs_solve_report) * k);
  if (k <= 4) return solve_multi<4>(A, *desc, k, b, x, reports);
  if (k <= 6) return solve_multi<6>(A, *desc, k, b, x, reports);
  return solve_multi<8>(A, *desc, k, b, x, reports);
}

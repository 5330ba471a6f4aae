// Krylov solvers of the FEM solve path on gfx950:
//  * Jacobi-PCG / CG, device-resident: scalars (alpha, beta, norms, the stop
//    flag) live in device memory, iterations run in graph-captured chunks and
//    the host reads one small status record per chunk.
//  * MIXED precision: FP64 outer iterative refinement (true residual
//    r = b - A x, x += d, stopping norms in FP64) around FP32 inner solves.
//  * GMRES(m): CGS2 Arnoldi with right Jacobi preconditioning, Givens
//    rotations on the host (one small read-back per Arnoldi step).
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

// Two kernels per iteration.  K1 reads the fields K2 wrote (rz2, iter2) and
// K2 reads the fields K1 wrote (rz1, iter1); each is written by block 0 only,
// so no kernel reads a field it writes.
struct CgState {
  double rz1, rz2, rr, bb, tol2, pq;
  int done, iter1, iter2, maxit;
  int warm;  // 1: the first iteration continues a kept search direction
};
// done codes
enum { CG_RUN = 0, CG_CONVERGED = 1, CG_BREAKDOWN = 2, CG_MAXIT = 3 };

inline unsigned grid_for(int64_t n, int64_t cap) {
  int64_t g = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// Sum of the G per-block partials part[k*G + b] (k < NV) of the previous
// kernel, done redundantly by every block of the consuming kernel: the kernel
// boundary makes the producer's partials visible, so no fences or atomics are
// needed.  Split in two so that a kernel can overlap the partials' load
// latency with its own loads: partials_issue loads the first P partials of
// each value per thread (G <= 256*P needs nothing more), partials_finish adds
// any further ones and reduces over the block.  The order is the same in
// every block, so every block obtains bit-identical totals.
template <int NV, int P>
__device__ __forceinline__ void partials_issue(const double* __restrict__ part, unsigned G, double (&pre)[NV][P]) {
#pragma unroll
  for (int k = 0; k < NV; k++)
#pragma unroll
    for (int u = 0; u < P; u++) {
      const unsigned b = threadIdx.x + u * 256;
      pre[k][u] = b < G ? part[(size_t)k * G + b] : 0.0;
    }
}

template <int NV, int P>
__device__ __forceinline__ void partials_finish(const double* __restrict__ part, unsigned G,
                                                const double (&pre)[NV][P], double (&tot)[NV]) {
  __shared__ double red[NV][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = pre[k][0];
#pragma unroll
    for (int u = 1; u < P; u++) s += pre[k][u];
    for (unsigned b = threadIdx.x + P * 256; b < G; b += 256) s += part[(size_t)k * G + b];
    s = wave_sum(s);
    if (lane == 0) red[k][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) tot[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
  __syncthreads();
}

// Block partial of NV values into part[k*gridDim + blockIdx].
template <typename T, int NV>
__device__ __forceinline__ void block_partials(T (&v)[NV], double* __restrict__ part) {
  __shared__ double red[NV][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum((double)v[k]);
    if (lane == 0) red[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV)
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
        (red[threadIdx.x][0] + red[threadIdx.x][1]) + (red[threadIdx.x][2] + red[threadIdx.x][3]);
}

// K1: from K2's partials (r.z, r.r): stop test, beta = rz / rz_prev;
// p_new = z + beta p_old (own rows written), q = A p_new, per-block partial
// of p_new.q.  Scalar rows compute q = A z + beta A p_old (both products in
// one pass over the matrix): those loads do not depend on beta, so they are
// issued while the partials of the previous kernel are still in flight.
// 3x3 blocks form p_new on the fly for the gathered columns once beta is
// known (two operand gathers per block column would cost 30 more registers).
template <typename T, int BS, typename CI>
__global__ __launch_bounds__(256) void k_cg_dir_spmv(const int64_t* __restrict__ slice_base,
                                                     const int32_t* __restrict__ slice_width,
                                                     const CI* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ z, const T* __restrict__ p_old,
                                                     T* __restrict__ p_new, T* __restrict__ q, int64_t nb,
                                                     int64_t n_slices, CgState* st, const double* __restrict__ part_rz,
                                                     unsigned g_rz, double* __restrict__ part_pq, int xcd) {
  const int was_done = st->done;
  const int64_t slice = (int64_t)xcd_block(xcd) * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const int64_t r = slice * 64 + lane;
  // the slice's start and width are loaded with the stop flag and the
  // partials, ahead of the branch on the flag
  const bool in_range = slice < n_slices;
  const int64_t sb = in_range ? slice_base[slice] : 0;
  const int w = in_range ? slice_width[slice] : 0;
  double pre[2][2];
  partials_issue<2, 2>(part_rz, g_rz, pre);
  if (was_done) return;
  T az[BS], ap[BS], zo[BS], po[BS];
#pragma unroll
  for (int a = 0; a < BS; a++) az[a] = ap[a] = zo[a] = po[a] = 0;
  if constexpr (BS == 1) {
    sell_row_pair<T, BS>(sb, w, col, val, z, p_old, slice, lane, az, ap);
    if (r < nb) {
      zo[0] = z[r];
      po[0] = p_old[r];
    }
  }
  double t[2];
  partials_finish<2, 2>(part_rz, g_rz, pre, t);
  const double rz = t[0], rr = t[1];
  const int iter = st->iter2;
  int done = CG_RUN;
  if (!isfinite(rz) || !isfinite(rr) || (iter > 0 && !(st->pq != 0.0 && isfinite(st->pq))))
    done = CG_BREAKDOWN;
  else if (rr <= st->tol2)
    done = CG_CONVERGED;
  else if (iter >= st->maxit)
    done = CG_MAXIT;
  const T beta = (iter > 0 || st->warm) ? (T)(rz / st->rz2) : (T)0;
  // every thread of this block has read the stop flag (partials_finish
  // ends with a barrier) before block 0 rewrites it
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz1 = rz;
    st->rr = rr;
    st->iter1 = iter;
    st->done = done;
  }
  if (done != CG_RUN) return;
  if constexpr (BS != 1) {
    // az = A p_new with the operand z_j + beta p_j formed on the fly
    if (slice < n_slices)
      sell_row_x<T, BS>(slice_base, slice_width, col, val, slice, lane, nb,
                        [&](int64_t i) { return z[i] + beta * p_old[i]; }, az);
    if (r < nb) {
#pragma unroll
      for (int a = 0; a < BS; a++) {
        zo[a] = z[BS * r + a];
        po[a] = p_old[BS * r + a];
      }
    }
  }
  T red[1] = {0};
  if (r < nb) {
#pragma unroll
    for (int a = 0; a < BS; a++) {
      const int64_t i = BS * r + a;
      const T pn = zo[a] + beta * po[a];
      const T qa = az[a] + beta * ap[a];
      p_new[i] = pn;
      q[i] = qa;
      red[0] += qa * pn;
    }
  }
  block_partials<T, 1>(red, part_pq);
}

// K2: alpha = rz / sum(part_pq); x += alpha p; r -= alpha q; z = dinv r;
// per-block partials of r.z and r.r.  The first pass's vector loads are
// issued before the partials of K1 are reduced.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(T* __restrict__ x, const T* __restrict__ p, T* __restrict__ r,
                                                   T* __restrict__ z, const T* __restrict__ q,
                                                   const T* __restrict__ dinv, int64_t n, CgState* st,
                                                   const double* __restrict__ part_pq, unsigned g_pq,
                                                   double* __restrict__ part_rz, int xcd, unsigned g1, int bs) {
  const int was_done = st->done;
  const double rz = st->rz1;
  double pre[1][5];
  partials_issue<1, 5>(part_pq, g_pq, pre);
  if (was_done) return;
  int64_t e0, e1;
  update_range(n, xcd, g1, bs, 256, e0, e1);
  const int64_t i0 = e0 + threadIdx.x;
  // first pass: four elements per thread, loaded together
  T xv[4], pv[4], rv[4], qv[4], dv[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = i0 + k * 256;
    const bool in = i < e1;
    xv[k] = in ? x[i] : (T)0;
    pv[k] = in ? p[i] : (T)0;
    rv[k] = in ? r[i] : (T)0;
    qv[k] = in ? q[i] : (T)0;
    dv[k] = dinv && in ? dinv[i] : (T)1;
  }
  double pq[1];
  partials_finish<1, 5>(part_pq, g_pq, pre, pq);
  const bool ok = pq[0] != 0.0 && isfinite(pq[0]);
  const T alpha = ok ? (T)(rz / pq[0]) : (T)0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz2 = rz;
    st->pq = pq[0];
    st->iter2 = st->iter1 + 1;
  }
  T red[2] = {0, 0};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = i0 + k * 256;
    if (i >= e1) continue;
    x[i] = xv[k] + alpha * pv[k];
    const T ri = rv[k] - alpha * qv[k];
    r[i] = ri;
    const T zi = dv[k] * ri;
    z[i] = zi;
    red[0] += ri * zi;
    red[1] += ri * ri;
  }
  // further passes when a block's chunk exceeds 1,024 elements
  for (int64_t j0 = i0 + 1024; j0 < e1; j0 += 1024) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int64_t i = j0 + k * 256;
      if (i >= e1) continue;
      x[i] += alpha * p[i];
      const T ri = r[i] - alpha * q[i];
      r[i] = ri;
      const T zi = dinv ? dinv[i] * ri : ri;
      z[i] = zi;
      red[0] += ri * zi;
      red[1] += ri * ri;
    }
  }
  block_partials<T, 2>(red, part_rz);
}

// Start: z = dinv r, p_old = 0, partials of r.z and r.r; scalar state.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_start(T* __restrict__ z, T* __restrict__ p0, const T* __restrict__ r,
                                                  const T* __restrict__ dinv, int64_t n, double* part_rz) {
  T red[2] = {0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = r[i];
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    p0[i] = 0;
    red[0] += ri * zi;
    red[1] += ri * ri;
  }
  block_partials<T, 2>(red, part_rz);
}
// Continuation after an FP64 residual replacement: the new residual is the
// old one times `ratio` (new units), so the kept direction is rescaled the
// same way; p_src may alias p0.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_start_keep(T* __restrict__ z, T* p0, const T* p_src,
                                                       const T* __restrict__ r, const T* __restrict__ dinv, T ratio,
                                                       int64_t n, double* part_rz) {
  T red[2] = {0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = r[i];
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    p0[i] = ratio * p_src[i];
    red[0] += ri * zi;
    red[1] += ri * ri;
  }
  block_partials<T, 2>(red, part_rz);
}
// warm = 0: fresh start (beta = 0 in the first iteration); warm = 1: the
// first beta is r.z / rz_prev with rz_prev the previous iteration's r.z in
// the new units.
__global__ void k_cg_start_state(CgState* st, double tol, double bb, int maxit, int warm, double rz_prev) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->warm = warm;
    st->rz1 = 1.0;
    st->rz2 = warm ? rz_prev : 1.0;
    st->rr = 0.0;
    st->bb = bb;
    st->tol2 = tol * tol * bb;
    st->iter1 = st->iter2 = 0;
    st->maxit = maxit;
    st->pq = 1.0;
    st->done = CG_RUN;
  }
}

// r32 = (float)(r64 * scale); x32 = 0
__global__ void k_demote_scaled(const double* __restrict__ r64, double scale, float* __restrict__ r32,
                                float* __restrict__ x32, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    r32[i] = (float)(r64[i] * scale);
    x32[i] = 0.0f;
  }
}
// x64 += scale * (double)d32 (optionally d = dinv * d first)
__global__ void k_promote_add(double* __restrict__ x64, const float* __restrict__ d32, const float* __restrict__ dinv,
                              double scale, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double d = (double)d32[i];
    if (dinv) d *= (double)dinv[i];
    x64[i] += scale * d;
  }
}

// ---------------------------------------------------------------------------
// PCG engine in precision T on one matrix.
// ---------------------------------------------------------------------------
// Bump allocator over a caller-provided device buffer (mfs_solve_noalloc):
// 256-byte aligned slices; with base == nullptr it only measures.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, off = 0;
  bool ok = true;
  template <typename T>
  T* take(int64_t count) {
    const size_t start = (off + 255) & ~size_t(255);
    const size_t bytes = sizeof(T) * (size_t)std::max<int64_t>(count, 1);
    if (start + bytes > cap) {
      ok = false;
      return nullptr;
    }
    off = start + bytes;
    return base ? reinterpret_cast<T*>(base + start) : nullptr;
  }
};

template <typename T>
struct PcgEngine {
  mfs_mat A = nullptr;
  int64_t n = 0;
  T *x = nullptr, *r = nullptr, *z = nullptr, *q = nullptr;
  T* pbuf[2] = {nullptr, nullptr};  // ping-pong search directions
  const T* dinv = nullptr;
  const T* val = nullptr;
  CgState* st = nullptr;
  ReduceScratch red;
  unsigned g_spmv = 0, g_vec = 0;
  int chunk = 16;
  bool use_graph = true;
  // external buffers (mfs_solve_noalloc): no graph, no events; the host
  // waits for each chunk with a stream synchronisation
  bool sync_only = false;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};  // end of the chunks in the two status slots

  // Large systems: one extra kernel per reduction adds the producer's
  // partials once (tot), instead of every consumer block re-adding them all
  // (their cost grows with the product of the two grids).
  bool fin = false;
  double* tot = nullptr;  // [pq, rz, rr]
  double* part_pq() { return red.partials; }
  double* part_rz() { return red.partials + g_spmv; }
  const double* rz_src() { return fin ? tot + 1 : part_rz(); }
  unsigned rz_g() { return fin ? 1u : g_vec; }
  const double* pq_src() { return fin ? tot : part_pq(); }
  unsigned pq_g() { return fin ? 1u : g_spmv; }
  void finish_rz(hipStream_t s) {
    if (fin) k_finish_sums<<<2, 256, 0, s>>>(part_rz(), g_vec, tot + 1);
  }
  void finish_pq(hipStream_t s) {
    if (fin) k_finish_sums<<<1, 256, 0, s>>>(part_pq(), g_spmv, tot);
  }

  int setup(mfs_mat A_, bool precond, int chunk_) {
    A = A_;
    n = A->n();
    chunk = std::max(2, chunk_ + (chunk_ & 1));  // even: the ping-pong parity repeats
    const bool f64 = std::is_same<T, double>::value;
    if (!f64 && !A->val32_valid) MFS_TRY(mat_refresh_derived(A));
    val = f64 ? (const T*)A->d_val64 : (const T*)A->d_val32;
    dinv = precond ? (f64 ? (const T*)A->d_dinv64 : (const T*)A->d_dinv32) : nullptr;
    g_spmv = (unsigned)spmv_grid(A);
    // update-kernel grid: about 4 elements per thread (its loop keeps 4 in
    // flight) but at least one block per CU; fewer blocks also means fewer
    // partial sums for every block of the next kernel to add up
    const int64_t cus = std::max(A->ctx->cu_count, 64);
    g_vec = (unsigned)std::min<int64_t>(4 * cus, std::max<int64_t>(cus, (n + 1023) / 1024));
    fin = (uint64_t)g_spmv * g_vec > 2000000ull;
    if (const char* e = std::getenv("MFS_FINISH_SUMS")) fin = e[0] == '1';  // force either way (tests)
    return MFS_OK;
  }
  int64_t partial_capacity() const { return (int64_t)g_spmv + 2 * (int64_t)g_vec + 64; }

  int init(mfs_mat A_, bool precond, int chunk_) {
    MFS_TRY(setup(A_, precond, chunk_));
    for (T** b : {&x, &r, &z, &q, &pbuf[0], &pbuf[1]}) MFS_TRY(dmalloc((void**)b, sizeof(T) * n));
    MFS_TRY(dmalloc((void**)&st, sizeof(CgState)));
    MFS_HIP(hipMemset(st, 0, sizeof(CgState)));
    MFS_TRY(reduce_scratch_alloc(red, partial_capacity()));
    MFS_TRY(dmalloc((void**)&tot, sizeof(double) * 4));
    return capture();
  }

  // The same engine on buffers carved from `ar` (no allocation, no graph);
  // with ar.base == nullptr only the sizes are counted.  Never destroy()ed.
  int init_ext(mfs_mat A_, bool precond, int chunk_, Arena& ar) {
    MFS_TRY(setup(A_, precond, chunk_));
    for (T** b : {&x, &r, &z, &q, &pbuf[0], &pbuf[1]}) *b = ar.take<T>(n);
    st = ar.take<CgState>(1);
    red.capacity = partial_capacity();
    red.partials = ar.take<double>(red.capacity);
    red.counter = ar.take<unsigned>(16);
    tot = ar.take<double>(4);
    use_graph = false;
    sync_only = true;
    if (ar.base && ar.ok) {
      // the scratch keeps the previous call's contents: clear the state
      MFS_HIP(hipMemsetAsync(st, 0, sizeof(CgState), A->ctx->stream));
      MFS_HIP(hipMemsetAsync(red.counter, 0, 16 * sizeof(unsigned), A->ctx->stream));
    }
    return MFS_OK;
  }

  // K1 with 16-bit row-relative columns when the matrix has them
  void launch_k1(hipStream_t s, T* p_old, T* p_new) {
    const int64_t ns = A->sell.n_slices;
    if (A->d_col16) {
      if (A->bs == 1)
        k_cg_dir_spmv<T, 1, int16_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col16, val, z,
                                                            p_old, p_new, q, A->nb, ns, st, rz_src(), rz_g(),
                                                            part_pq(), A->xcd_map);
      else
        k_cg_dir_spmv<T, 3, int16_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col16, val, z,
                                                            p_old, p_new, q, A->nb, ns, st, rz_src(), rz_g(),
                                                            part_pq(), A->xcd_map);
    } else {
      if (A->bs == 1)
        k_cg_dir_spmv<T, 1, int32_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, val, z,
                                                            p_old, p_new, q, A->nb, ns, st, rz_src(), rz_g(),
                                                            part_pq(), A->xcd_map);
      else
        k_cg_dir_spmv<T, 3, int32_t><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, val, z,
                                                            p_old, p_new, q, A->nb, ns, st, rz_src(), rz_g(),
                                                            part_pq(), A->xcd_map);
    }
  }

  void launch_iteration(hipStream_t s, int parity) {
    T* p_old = pbuf[parity];
    T* p_new = pbuf[parity ^ 1];
    launch_k1(s, p_old, p_new);
    finish_pq(s);
    k_cg_update<T><<<g_vec, 256, 0, s>>>(x, p_new, r, z, q, dinv, n, st, pq_src(), pq_g(), part_rz(), A->xcd_map,
                                         g_spmv, A->bs);
    finish_rz(s);
  }

  int capture() {
    const char* e = std::getenv("MFS_NO_GRAPH");
    use_graph = !(e && e[0] == '1');
    if (!use_graph) return MFS_OK;
    hipStream_t s = A->ctx->stream;
    MFS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < chunk; k++) launch_iteration(s, k & 1);
    MFS_HIP(hipStreamEndCapture(s, &graph));
    MFS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return MFS_OK;
  }

  // Start from r (holding the initial residual) and x; bb = ||b||^2 of the
  // system being solved.  Runs until converged / maxit / breakdown.  The
  // last K1 to run records the final residual in the state.
  //
  // keep_ratio > 0 continues the previous run's search direction (residual
  // replacement instead of a restart): the previous run must have stopped
  // on convergence after at least one iteration, and r now holds the new
  // residual, equal to the old recursive one times keep_ratio.
  int last_iter = 0;
  double last_rz_prev = 0.0;
  bool can_keep = false;

  int run(double tol, double bb, int maxit, CgState* host_state, double keep_ratio = 0.0) {
    hipStream_t s = A->ctx->stream;
    const bool warm = keep_ratio > 0.0 && can_keep;
    if (warm) {
      k_cg_start_keep<T><<<g_vec, 256, 0, s>>>(z, pbuf[0], pbuf[last_iter & 1], r, dinv, (T)keep_ratio, n, part_rz());
      k_cg_start_state<<<1, 64, 0, s>>>(st, tol, bb, maxit, 1, last_rz_prev * keep_ratio * keep_ratio);
    } else {
      k_cg_start<T><<<g_vec, 256, 0, s>>>(z, pbuf[0], r, dinv, n, part_rz());
      k_cg_start_state<<<1, 64, 0, s>>>(st, tol, bb, maxit, 0, 0.0);
    }
    finish_rz(s);
    MFS_HIP(hipGetLastError());
    // Chunks of iterations; each ends with a copy of the state into one of
    // two pinned slots.  While the residual is far above the tolerance one
    // more chunk is queued before the host waits for the previous one, so
    // the GPU does not idle during the host's check; near convergence the
    // host checks every chunk (a chunk queued after convergence returns at
    // once in every kernel and changes nothing).
    CgState* hs = (CgState*)A->ctx->pinned;
    if (sync_only) {
      // one chunk at a time, the host waits on the stream
      for (;;) {
        for (int k = 0; k < chunk; k++) launch_iteration(s, k & 1);
        MFS_HIP(hipGetLastError());
        MFS_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        MFS_HIP(hipStreamSynchronize(s));
        if (hs->done != CG_RUN) break;
      }
      *host_state = *hs;
      host_state->iter2 = host_state->iter1;
      last_iter = host_state->iter1;
      last_rz_prev = host_state->rz2;
      can_keep = host_state->done == CG_CONVERGED && host_state->iter1 > 0 && std::isfinite(host_state->rz2) &&
                 host_state->rz2 > 0.0;
      return MFS_OK;
    }
    for (hipEvent_t* e : {&ev[0], &ev[1]})
      if (!*e) MFS_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    auto submit = [&](int slot) -> int {
      if (use_graph) {
        MFS_HIP(hipGraphLaunch(exec, s));
      } else {
        for (int k = 0; k < chunk; k++) launch_iteration(s, k & 1);
        MFS_HIP(hipGetLastError());
      }
      MFS_HIP(hipMemcpyAsync(hs + slot, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
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
      if (hs[head].done != CG_RUN) break;
      ahead = hs[head].rr > 100.0 * hs[head].tol2;
      head ^= 1;
      if (inflight == 0) {
        MFS_TRY(submit(head));
        inflight = 1;
      }
    }
    *host_state = hs[head];
    hs = host_state;
    if (inflight) MFS_HIP(hipStreamSynchronize(s));
    host_state->iter2 = hs->iter1;  // iterations completed when K1 stopped
    // the newest direction is in pbuf[iter & 1]; rz2 holds r.z of the
    // iteration before the last (the denominator of the next beta)
    last_iter = hs->iter1;
    last_rz_prev = hs->rz2;
    can_keep = hs->done == CG_CONVERGED && hs->iter1 > 0 && std::isfinite(hs->rz2) && hs->rz2 > 0.0;
    return MFS_OK;
  }

  void destroy() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e);
    ev[0] = ev[1] = nullptr;
    for (T* b : {x, r, z, q, pbuf[0], pbuf[1]}) dfree(b);
    dfree(st);
    dfree(tot);
    reduce_scratch_free(red);
  }
};

// ---------------------------------------------------------------------------
// GMRES(m) pieces (precision T): multi-dot h_k = v_k . w for k < kcount,
// w -= sum_k h_k v_k, norms.  V is column-blocked: V + k*n.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_multidot(const T* __restrict__ V, const T* __restrict__ w, int64_t n,
                                                  double* __restrict__ partials) {
  // block (x, k): partials[k*G + x] = this block's rows of v_k . w (G =
  // gridDim.x); every block reduces one value, and the blocks of all k run
  // at once (w is re-read from the cache by each k)
  const int k = blockIdx.y;
  const T* __restrict__ v = V + (int64_t)k * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  T s[4] = {0, 0, 0, 0};
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * stride;
      if (i < n) s[u] += v[i] * w[i];
    }
  }
  T t[1] = {(s[0] + s[1]) + (s[2] + s[3])};
  write_block_partials<T, 1>(t, partials + (size_t)k * gridDim.x);
}

// w -= sum_k h_k v_k (h in device memory, double); per-block partial ||w||^2.
template <typename T>
__global__ __launch_bounds__(256) void k_orth_update(const T* __restrict__ V, T* __restrict__ w, int64_t n,
                                                     int kcount, const double* __restrict__ h,
                                                     double* __restrict__ partials) {
  __shared__ T hs[32];
  if (threadIdx.x < kcount) hs[threadIdx.x] = (T)h[threadIdx.x];
  __syncthreads();
  T red[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T wi = w[i];
    for (int k = 0; k < kcount; k++) wi -= hs[k] * V[(int64_t)k * n + i];
    w[i] = wi;
    red[0] += wi * wi;
  }
  write_block_partials<T, 1>(red, partials);
}

template <typename T>
__global__ void k_scale_into(const T* __restrict__ w, T* __restrict__ v, const double* __restrict__ nrm, int64_t n) {
  const T inv = nrm[0] > 0.0 ? (T)(1.0 / sqrt(nrm[0])) : (T)0;  // a zero vector stays zero
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = w[i] * inv;
}
template <typename T>
__global__ void k_dscale(const T* __restrict__ v, const T* __restrict__ dinv, T* __restrict__ z, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    z[i] = dinv ? dinv[i] * v[i] : v[i];
}
// u = sum_k y_k v_k  (y on device, double)
template <typename T>
__global__ void k_combine(const T* __restrict__ V, const double* __restrict__ y, int kcount, T* __restrict__ u,
                          int64_t n) {
  __shared__ T ys[32];
  if (threadIdx.x < kcount) ys[threadIdx.x] = (T)y[threadIdx.x];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T s = 0;
    for (int k = 0; k < kcount; k++) s += ys[k] * V[(int64_t)k * n + i];
    u[i] = s;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Preconditioned MINRES for symmetric (possibly indefinite) A with SPD M =
// diag(A) (Elman, Silvester & Wathen, "Finite Elements and Fast Iterative
// Solvers", Algorithm 2.4).  The scalar recurrence (Lanczos norms, Givens
// rotations, eta) stays on the device: a one-thread kernel per step updates
// it, and the host reads the stop record every 8 steps; updates of steps
// launched past the stopping one leave x unchanged.
// ---------------------------------------------------------------------------
// MINRES recurrence scalars on the device (doubles): the Lanczos norms, the
// Givens rotation of the last two steps, eta, the update coefficients of
// the current step, and the stop record (-1 while running, else the index of
// the last step whose update applies; kMsBad = 1 when that step's rotation
// failed and its update must not apply either).
enum {
  kMsG, kMsGPrev, kMsC, kMsCPrev, kMsSPrev, kMsS, kMsEta, kMsEta0, kMsA1, kMsA2, kMsA3, kMsCEta, kMsInvG, kMsRtol,
  kMsStop, kMsBad, kMsCount
};

// one thread: rotation of step `step` from delta (dh[1]) and ||v_{j+1}||_M^2 (dh[2])
__global__ void k_minres_rot(const double* __restrict__ dh, double* __restrict__ ms, int step) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || ms[kMsStop] >= 0.0) return;
  const double delta = dh[1], g_next = sqrt(fmax(dh[2], 0.0));
  const double g = ms[kMsG], c = ms[kMsC], c_prev = ms[kMsCPrev], s_prev = ms[kMsSPrev], s_cur = ms[kMsS];
  const double a0 = c * delta - c_prev * s_cur * g;
  const double a1 = sqrt(a0 * a0 + g_next * g_next);
  if (!isfinite(delta) || !isfinite(g_next) || !(a1 > 0)) {
    ms[kMsBad] = 1.0;
    ms[kMsStop] = step;
    return;
  }
  const double c_next = a0 / a1, s_next = g_next / a1, eta = ms[kMsEta];
  ms[kMsA1] = a1;
  ms[kMsA2] = s_cur * delta + c_prev * c * g;
  ms[kMsA3] = s_prev * g;
  ms[kMsCEta] = c_next * eta;
  ms[kMsInvG] = g_next > 0 ? 1.0 / g_next : 0.0;
  const double eta_next = -s_next * eta;
  ms[kMsEta] = eta_next;
  ms[kMsCPrev] = c;
  ms[kMsC] = c_next;
  ms[kMsSPrev] = s_cur;
  ms[kMsS] = s_next;
  ms[kMsGPrev] = g;
  ms[kMsG] = g_next;
  if (fabs(eta_next) <= ms[kMsRtol] * ms[kMsEta0] || g_next == 0) ms[kMsStop] = step;
}

// v_new = q - (delta/g) v - (g/g_prev) v_old; z_new = dinv v_new;
// per-block partial of z_new . v_new.   delta from device memory.
template <typename T>
__global__ __launch_bounds__(256) void k_minres_lanczos(const T* __restrict__ q, const T* __restrict__ v,
                                                        const T* __restrict__ v_old, T* __restrict__ v_new,
                                                        T* __restrict__ z_new, const T* __restrict__ dinv,
                                                        const double* __restrict__ delta,
                                                        const double* __restrict__ ms, int64_t n,
                                                        double* __restrict__ partials) {
  const double g = ms[kMsG], g_prev = ms[kMsGPrev];
  const T c1 = (T)(delta[0] / g), c2 = g_prev > 0 ? (T)(g / g_prev) : (T)0;
  T red[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T vn = q[i] - c1 * v[i] - c2 * v_old[i];
    v_new[i] = vn;
    const T zn = dinv ? dinv[i] * vn : vn;
    z_new[i] = zn;
    red[0] += zn * vn;
  }
  write_block_partials<T, 1>(red, partials);
}
// w_new = (z - a3 w_old - a2 w) / a1; x += (c eta) w_new; z_next *= 1/g_next
template <typename T>
__global__ __launch_bounds__(256) void k_minres_update(const T* __restrict__ z, const T* __restrict__ w,
                                                       const T* __restrict__ w_old, T* __restrict__ w_new,
                                                       T* __restrict__ x, T* __restrict__ z_next,
                                                       const double* __restrict__ ms, int step, int64_t n) {
  // steps after the stopping one, and a step whose rotation failed, change nothing
  const double stop = ms[kMsStop];
  if (stop >= 0.0 && ((double)step > stop || ms[kMsBad] != 0.0)) return;
  const T ia1 = (T)(1.0 / ms[kMsA1]), ta2 = (T)ms[kMsA2], ta3 = (T)ms[kMsA3], tc = (T)ms[kMsCEta],
          tg = (T)ms[kMsInvG];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T wn = (z[i] - ta3 * w_old[i] - ta2 * w[i]) * ia1;
    w_new[i] = wn;
    x[i] += tc * wn;
    z_next[i] *= tg;
  }
}
template <typename T>
__global__ void k_scale_vec(T* __restrict__ a, double s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] *= (T)s;
}
template <typename T>
__global__ void k_minres_start(const T* __restrict__ r, T* __restrict__ v, T* __restrict__ z, const T* __restrict__ dinv,
                               T* __restrict__ x, T* __restrict__ v_old, T* __restrict__ w, T* __restrict__ w_old,
                               int64_t n, double* __restrict__ partials) {
  T red[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = r[i];
    v[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    x[i] = 0;
    v_old[i] = 0;
    w[i] = 0;
    w_old[i] = 0;
    red[0] += zi * ri;
  }
  write_block_partials<T, 1>(red, partials);
}

// ---------------------------------------------------------------------------
// BiCGStab with right Jacobi preconditioning (van der Vorst 1992; the
// reference's BiCG_Stab_Preconditioning, CSSD_Matrix.h:1058-1163, with its
// inner M-solve replaced by M = diag(A)):
//   p^ = M^-1 p, v = A p^, alpha = rho / (r0.v), s = r - alpha v,
//   s^ = M^-1 s, t = A s^, omega = (t.s) / (t.t), x += alpha p^ + omega s^,
//   r = s - omega t, rho' = r0.r, beta = (rho'/rho)(alpha/omega),
//   p = r + beta (p - omega v).
// Three kernels per iteration, no separate reduction or scalar kernels:
//   K_a: sums K_c's partials (r0.r, r.r) -> stop test, beta; gathers
//        p^ = r^ + beta w^ on the fly (r^ = M^-1 r, w^ = p^_old - omega v^
//        stored by K_c); v = A p^, v^ = M^-1 v; partial r0.v.
//   K_b: sums K_a's partials -> alpha; gathers s^ = r^ - alpha v^ on the
//        fly; t = A s^, s = r - alpha v; partials t.s, t.t.
//   K_c: sums K_b's partials -> omega; x += alpha p^ + omega s^,
//        r = s - omega t, r^ = M^-1 r, w^ = p^ - omega v^; partials r0.r, r.r.
// Every block of a consumer adds the producer's partials in the same order
// (identical scalars in all blocks); scalars another kernel needs are
// written by block 0 into fields that kernel does not write.  Scalar rows
// form the gathered operands as two products (A r^ + beta A w^, A r^ -
// alpha A v^), so their loads are issued before the partials are reduced.
// ---------------------------------------------------------------------------
struct BsState {
  double rho_a, rho_b;  // rho of this iteration (K_a), of the previous one (K_b)
  double alpha, omega, beta, tol2, rr;
  int done, iter, iter_c, maxit;
};
enum { BS_RUN = 0, BS_CONVERGED = 1, BS_BREAKDOWN = 2, BS_MAXIT = 3 };

// y = A (x1 + c x2) for one slice row; scalar rows as two products (a1, a2)
// combined by the caller, 3x3 blocks with the operand formed on the fly.
template <typename T, int BS, typename CI>
__device__ __forceinline__ void bs_rows_pair(const int64_t* __restrict__ slice_base,
                                             const int32_t* __restrict__ slice_width, const CI* __restrict__ col,
                                             const T* __restrict__ val, const T* __restrict__ x1,
                                             const T* __restrict__ x2, int64_t slice, int64_t n_slices, int lane,
                                             T (&a1)[BS], T (&a2)[BS]) {
  const bool in_range = slice < n_slices;
  sell_row_pair<T, BS>(in_range ? slice_base[slice] : 0, in_range ? slice_width[slice] : 0, col, val, x1, x2, slice,
                       lane, a1, a2);
}

// K_a
template <typename T, int BS, typename CI>
__global__ __launch_bounds__(256) void k_bs_dir_spmv(const int64_t* __restrict__ slice_base,
                                                     const int32_t* __restrict__ slice_width,
                                                     const CI* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ rh, const T* __restrict__ wh,
                                                     const T* __restrict__ r0, const T* __restrict__ dinv,
                                                     T* __restrict__ v, T* __restrict__ vh, int64_t nb,
                                                     int64_t n_slices, BsState* st, const double* __restrict__ part_c,
                                                     unsigned g_c, double* __restrict__ part_a, int xcd) {
  const int was_done = st->done;
  double pre[2][4];
  partials_issue<2, 4>(part_c, g_c, pre);
  if (was_done) return;
  const int64_t slice = (int64_t)xcd_block(xcd) * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const int64_t row = slice * 64 + lane;
  T a1[BS], a2[BS];
  if constexpr (BS == 1) bs_rows_pair<T, BS>(slice_base, slice_width, col, val, rh, wh, slice, n_slices, lane, a1, a2);
  double t[2];
  partials_finish<2, 4>(part_c, g_c, pre, t);
  const double rho = t[0], rr = t[1];
  const int it = st->iter_c;
  int done = BS_RUN;
  double beta = 0.0;
  if (it > 0) {
    const double alpha = st->alpha, omega = st->omega;
    if (!isfinite(alpha) || !isfinite(omega) || !isfinite(rho) || !isfinite(rr))
      done = BS_BREAKDOWN;
    else if (rr <= st->tol2)
      done = BS_CONVERGED;
    else if (it >= st->maxit)
      done = BS_MAXIT;
    else if (rho == 0.0 || omega == 0.0)
      done = BS_BREAKDOWN;
    beta = (rho / st->rho_b) * (alpha / omega);
  }
  // every thread of this block has read the stop flag (partials_finish
  // ends with a barrier) before block 0 rewrites it
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rho_a = rho;
    st->rr = rr;
    st->iter = it;
    st->beta = beta;
    st->done = done;
  }
  if (done != BS_RUN) return;
  const T b = (T)beta;
  if constexpr (BS != 1) {
    a2[0] = a2[1] = a2[2] = 0;
    if (slice < n_slices)
      sell_row_x<T, BS>(slice_base, slice_width, col, val, slice, lane, nb,
                        [&](int64_t i) { return rh[i] + b * wh[i]; }, a1);
    else
      a1[0] = a1[1] = a1[2] = 0;
  }
  T red[1] = {0};
  if (row < nb) {
#pragma unroll
    for (int a = 0; a < BS; a++) {
      const int64_t i = BS * row + a;
      const T vi = a1[a] + b * a2[a];
      v[i] = vi;
      vh[i] = dinv ? dinv[i] * vi : vi;
      red[0] += r0[i] * vi;
    }
  }
  block_partials<T, 1>(red, part_a);
}

// K_b
template <typename T, int BS, typename CI>
__global__ __launch_bounds__(256) void k_bs_s_spmv(const int64_t* __restrict__ slice_base,
                                                   const int32_t* __restrict__ slice_width,
                                                   const CI* __restrict__ col, const T* __restrict__ val,
                                                   const T* __restrict__ rh, const T* __restrict__ vh,
                                                   const T* __restrict__ r, const T* __restrict__ v,
                                                   T* __restrict__ s, T* __restrict__ t, int64_t nb,
                                                   int64_t n_slices, BsState* st, const double* __restrict__ part_a,
                                                   unsigned g_a, double* __restrict__ part_b, int xcd) {
  const int was_done = st->done;
  const double rho = st->rho_a;
  double pre[1][4];
  partials_issue<1, 4>(part_a, g_a, pre);
  if (was_done) return;
  const int64_t slice = (int64_t)xcd_block(xcd) * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const int64_t row = slice * 64 + lane;
  T a1[BS], a2[BS];
  if constexpr (BS == 1) bs_rows_pair<T, BS>(slice_base, slice_width, col, val, rh, vh, slice, n_slices, lane, a1, a2);
  double r0v[1];
  partials_finish<1, 4>(part_a, g_a, pre, r0v);
  const double alpha = rho / r0v[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha;
    st->rho_b = rho;
  }
  const T al = (T)alpha;
  if constexpr (BS != 1) {
    a2[0] = a2[1] = a2[2] = 0;
    if (slice < n_slices)
      sell_row_x<T, BS>(slice_base, slice_width, col, val, slice, lane, nb,
                        [&](int64_t i) { return rh[i] - al * vh[i]; }, a1);
    else
      a1[0] = a1[1] = a1[2] = 0;
  }
  T red[2] = {0, 0};
  if (row < nb) {
#pragma unroll
    for (int a = 0; a < BS; a++) {
      const int64_t i = BS * row + a;
      const T ti = a1[a] - al * a2[a];
      const T si = r[i] - al * v[i];
      s[i] = si;
      t[i] = ti;
      red[0] += ti * si;
      red[1] += ti * ti;
    }
  }
  block_partials<T, 2>(red, part_b);
}

// K_c
template <typename T>
__global__ __launch_bounds__(256) void k_bs_update(T* __restrict__ x, T* __restrict__ r, T* __restrict__ rh,
                                                   T* __restrict__ wh, const T* __restrict__ vh,
                                                   const T* __restrict__ s, const T* __restrict__ t,
                                                   const T* __restrict__ r0, const T* __restrict__ dinv, int64_t n,
                                                   BsState* st, const double* __restrict__ part_b, unsigned g_b,
                                                   double* __restrict__ part_c, int xcd, unsigned g1, int bs) {
  const int was_done = st->done;
  const T alpha = (T)st->alpha, beta = (T)st->beta;
  double pre[2][4];
  partials_issue<2, 4>(part_b, g_b, pre);
  if (was_done) return;
  // the rows that this XCD's blocks of k_bs_dir_spmv gather next; the loads
  // of the first 4 elements per thread are issued before omega is known
  int64_t e0, e1;
  update_range(n, xcd, g1, bs, 64, e0, e1);
  const int64_t i0 = e0 + threadIdx.x;
  T xv[4], rv[4], wv[4], vv[4], sv[4], tv[4], r0v[4], dv[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int64_t i = i0 + u * 256;
    const bool in = i < e1;
    xv[u] = in ? x[i] : (T)0;
    rv[u] = in ? rh[i] : (T)0;
    wv[u] = in ? wh[i] : (T)0;
    vv[u] = in ? vh[i] : (T)0;
    sv[u] = in ? s[i] : (T)0;
    tv[u] = in ? t[i] : (T)0;
    r0v[u] = in ? r0[i] : (T)0;
    dv[u] = dinv && in ? dinv[i] : (T)1;
  }
  double ts[2];
  partials_finish<2, 4>(part_b, g_b, pre, ts);
  const double omega = ts[0] / ts[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->omega = omega;
    st->iter_c = st->iter + 1;
  }
  const T om = (T)omega;
  T red[2] = {0, 0};
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int64_t i = i0 + u * 256;
    if (i >= e1) continue;
    const T ph = rv[u] + beta * wv[u];
    x[i] = xv[u] + (alpha * ph + om * (dv[u] * sv[u]));
    const T ri = sv[u] - om * tv[u];
    r[i] = ri;
    rh[i] = dv[u] * ri;
    wh[i] = ph - om * vv[u];
    red[0] += r0v[u] * ri;
    red[1] += ri * ri;
  }
  // further elements when a block's range exceeds 1,024
  for (int64_t i = i0 + 1024; i < e1; i += blockDim.x) {
    const T di = dinv ? dinv[i] : (T)1;
    const T ph = rh[i] + beta * wh[i];
    const T si = s[i];
    x[i] += alpha * ph + om * (di * si);
    const T ri = si - om * t[i];
    r[i] = ri;
    rh[i] = di * ri;
    wh[i] = ph - om * vh[i];
    red[0] += r0[i] * ri;
    red[1] += ri * ri;
  }
  block_partials<T, 2>(red, part_c);
}

// r0 = r (the right-hand side), r^ = M^-1 r, w^ = 0, x = 0; partials of
// r0.r and r.r (both r.r) in K_a's input layout
template <typename T>
__global__ __launch_bounds__(256) void k_bs_start(const T* __restrict__ r, T* __restrict__ r0, T* __restrict__ rh,
                                                  T* __restrict__ wh, T* __restrict__ x, const T* __restrict__ dinv,
                                                  int64_t n, double* __restrict__ part) {
  T red[2] = {0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = r[i];
    r0[i] = ri;
    rh[i] = dinv ? dinv[i] * ri : ri;
    wh[i] = 0;
    x[i] = 0;
    red[0] += ri * ri;
    red[1] += ri * ri;
  }
  block_partials<T, 2>(red, part);
}

__global__ void k_bs_init_state(BsState* st, const double* __restrict__ tot, double rtol, int maxit) {
  st->rho_a = st->rho_b = tot[0];
  st->tol2 = rtol * rtol * tot[0];
  st->rr = tot[0];
  st->alpha = st->omega = st->beta = 0.0;
  st->iter = st->iter_c = 0;
  st->maxit = maxit;
  st->done = tot[0] > 0.0 ? (isfinite(tot[0]) ? BS_RUN : BS_BREAKDOWN) : BS_CONVERGED;
}

// Inner BiCGStab engine in precision T: solves A d = rhs from d = 0 (rhs in r).
template <typename T>
struct BsEngine {
  mfs_mat A = nullptr;
  int64_t n = 0;
  T *x = nullptr, *r = nullptr, *r0 = nullptr, *rh = nullptr, *wh = nullptr, *v = nullptr, *vh = nullptr,
    *sv = nullptr, *t = nullptr;
  const T* val = nullptr;
  const T* dinv = nullptr;
  BsState* st = nullptr;
  // partials of K_a (1 value), K_b (2), K_c / start (2); tot: reduced
  // values when the partials are first added once by k_finish_sums
  double *part_a = nullptr, *part_b = nullptr, *part_c = nullptr, *tot = nullptr;
  unsigned g_spmv = 0, g_vec = 0;
  bool fin = false;
  int chunk = 8;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};  // end of the chunks in the two status slots

  const double* src_a() { return fin ? tot : part_a; }
  const double* src_b() { return fin ? tot + 1 : part_b; }
  const double* src_c() { return fin ? tot + 3 : part_c; }
  unsigned ga() { return fin ? 1u : g_spmv; }
  unsigned gc() { return fin ? 1u : g_vec; }

  int init(mfs_mat A_, bool precond) {
    A = A_;
    n = A->n();
    const bool f64 = std::is_same<T, double>::value;
    if (!f64 && !A->val32_valid) MFS_TRY(mat_refresh_derived(A));
    val = f64 ? (const T*)A->d_val64 : (const T*)A->d_val32;
    dinv = precond ? (f64 ? (const T*)A->d_dinv64 : (const T*)A->d_dinv32) : nullptr;
    for (T** b : {&x, &r, &r0, &rh, &wh, &v, &vh, &sv, &t}) MFS_TRY(dmalloc((void**)b, sizeof(T) * n));
    MFS_TRY(dmalloc((void**)&st, sizeof(BsState)));
    g_spmv = (unsigned)spmv_grid(A);
    const int64_t cus = std::max(A->ctx->cu_count, 64);
    g_vec = (unsigned)std::min<int64_t>(4 * cus, std::max<int64_t>(cus, (n + 1023) / 1024));
    MFS_TRY(dmalloc((void**)&part_a, sizeof(double) * g_spmv));
    MFS_TRY(dmalloc((void**)&part_b, sizeof(double) * 2 * g_spmv));
    MFS_TRY(dmalloc((void**)&part_c, sizeof(double) * 2 * g_vec));
    MFS_TRY(dmalloc((void**)&tot, sizeof(double) * 8));
    // every block of a consumer adds all partials of its producer; for
    // large grids a one-block-per-value kernel adds them once instead
    fin = (uint64_t)g_spmv * std::max(g_spmv, g_vec) > 2000000ull;
    if (const char* e = std::getenv("MFS_FINISH_SUMS")) fin = e[0] == '1';  // force either way (tests)
    // MFS_NO_GRAPH=1 launches the same kernels directly (kernel tracing)
    const char* e = std::getenv("MFS_NO_GRAPH");
    if (e && e[0] == '1') return MFS_OK;
    hipStream_t s = A->ctx->stream;
    MFS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < chunk; k++) launch_iteration(s);
    MFS_HIP(hipStreamEndCapture(s, &graph));
    MFS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return MFS_OK;
  }

  template <int BS, typename CI>
  void launch_spmvs(hipStream_t s, const CI* col, int which) {
    const int64_t ns = A->sell.n_slices;
    if (which == 0)
      k_bs_dir_spmv<T, BS, CI><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, col, val, rh, wh, r0, dinv,
                                                      v, vh, A->nb, ns, st, src_c(), gc(), part_a, A->xcd_map);
    else
      k_bs_s_spmv<T, BS, CI><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, col, val, rh, vh, r, v, sv,
                                                    t, A->nb, ns, st, src_a(), ga(), part_b, A->xcd_map);
  }
  void spmv(hipStream_t s, int which) {
    if (A->d_col16) {
      if (A->bs == 1)
        launch_spmvs<1>(s, A->d_col16, which);
      else
        launch_spmvs<3>(s, A->d_col16, which);
    } else {
      if (A->bs == 1)
        launch_spmvs<1>(s, A->d_col, which);
      else
        launch_spmvs<3>(s, A->d_col, which);
    }
  }

  void launch_iteration(hipStream_t s) {
    spmv(s, 0);
    if (fin) k_finish_sums<<<1, 256, 0, s>>>(part_a, g_spmv, tot);
    spmv(s, 1);
    if (fin) k_finish_sums<<<2, 256, 0, s>>>(part_b, g_spmv, tot + 1);
    k_bs_update<T><<<g_vec, 256, 0, s>>>(x, r, rh, wh, vh, sv, t, r0, dinv, n, st, src_b(), fin ? 1u : g_spmv,
                                         part_c, A->xcd_map, g_spmv, A->bs);
    if (fin) k_finish_sums<<<2, 256, 0, s>>>(part_c, g_vec, tot + 3);
  }

  // r holds the right-hand side; x returns the solution of A x = r.
  int run(double rtol, int maxit, BsState* out) {
    hipStream_t s = A->ctx->stream;
    k_bs_start<T><<<g_vec, 256, 0, s>>>(r, r0, rh, wh, x, dinv, n, part_c);
    k_finish_sums<<<1, 256, 0, s>>>(part_c, g_vec, tot + 6);
    k_bs_init_state<<<1, 1, 0, s>>>(st, tot + 6, rtol, maxit);
    if (fin) k_finish_sums<<<2, 256, 0, s>>>(part_c, g_vec, tot + 3);
    MFS_HIP(hipGetLastError());
    BsState* hs = (BsState*)A->ctx->pinned;
    MFS_HIP(hipMemcpyAsync(hs, st, sizeof(BsState), hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    if (hs->done) {
      *out = *hs;
      return MFS_OK;
    }
    // chunks of iterations, one queued ahead while the residual is far above
    // the tolerance (as in the PCG engine)
    for (hipEvent_t* e : {&ev[0], &ev[1]})
      if (!*e) MFS_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    auto submit = [&](int slot) -> int {
      if (exec)
        MFS_HIP(hipGraphLaunch(exec, s));
      else
        for (int k = 0; k < chunk; k++) launch_iteration(s);
      MFS_HIP(hipGetLastError());
      MFS_HIP(hipMemcpyAsync(hs + slot, st, sizeof(BsState), hipMemcpyDeviceToHost, s));
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
      if (hs[head].done) break;
      ahead = hs[head].rr > 100.0 * hs[head].tol2;
      head ^= 1;
      if (inflight == 0) {
        MFS_TRY(submit(head));
        inflight = 1;
      }
    }
    *out = hs[head];
    if (inflight) MFS_HIP(hipStreamSynchronize(s));
    return MFS_OK;
  }

  void destroy() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e);
    ev[0] = ev[1] = nullptr;
    for (T* b : {x, r, r0, rh, wh, v, vh, sv, t}) dfree(b);
    x = r = r0 = rh = wh = v = vh = sv = t = nullptr;
    dfree(st);
    for (double* b : {part_a, part_b, part_c, tot}) dfree(b);
    st = nullptr;
    part_a = part_b = part_c = tot = nullptr;
  }
};

// ---------------------------------------------------------------------------
// Solver object
// ---------------------------------------------------------------------------
struct mfs_solver_s {
  mfs_mat A = nullptr;
  mfs_solve_desc desc;
  int64_t n = 0;
  // FP64 outer / FP64 solve
  double* x64 = nullptr;
  double* b64 = nullptr;
  double* r64 = nullptr;
  ReduceScratch red64;
  double* dout = nullptr;  // small device outputs (norms)
  PcgEngine<double> cg64;
  PcgEngine<float> cg32;
  bool has64 = false, has32 = false;
  // GMRES
  // MINRES vectors: rhs / z_next, x, v[3], w[3], z, A z
  void* mr[10] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void* V = nullptr;   // (m+1) * n in the inner precision
  void* w = nullptr;
  void* z = nullptr;
  double* dh = nullptr;  // device h / y (64 doubles)
  // GMRES: one captured graph per Arnoldi step index j (FP64 and FP32 inner)
  hipGraphExec_t gstep[2][30] = {};
  ReduceScratch redg;
  BsEngine<double> bs64;
  BsEngine<float> bs32;
};

extern "C" void mfs_solve_desc_default(mfs_solve_desc* d) {
  if (!d) return;
  d->method = MFS_PCG_JACOBI;
  d->precision = MFS_PREC_F64;
  d->tol = 1e-8;
  d->maxit = 100000;
  d->restart = 30;
  d->inner_tol = 1e-4;
  d->check_every = 0;
}

extern "C" int mfs_solver_destroy(mfs_solver S) {
  if (!S) return MFS_OK;
  (void)hipStreamSynchronize(S->A->ctx->stream);
  mfs_mat A = S->A;
  if (S->has64) S->cg64.destroy();
  if (S->has32) S->cg32.destroy();
  S->bs64.destroy();
  S->bs32.destroy();
  dfree(S->x64);
  dfree(S->b64);
  dfree(S->r64);
  dfree(S->dout);
  for (void* v : S->mr) dfree(v);
  dfree(S->V);
  dfree(S->w);
  dfree(S->z);
  dfree(S->dh);
  for (auto& row : S->gstep)
    for (hipGraphExec_t& g : row)
      if (g) (void)hipGraphExecDestroy(g);
  reduce_scratch_free(S->red64);
  reduce_scratch_free(S->redg);
  delete S;
  mat_release(A);
  return MFS_OK;
}

extern "C" int mfs_solver_create(mfs_mat A, const mfs_solve_desc* desc, mfs_solver* out) {
  MFS_CHECK_ARG(A && desc && out, "mfs_solver_create: bad argument");
  MFS_CHECK_ARG(desc->method == MFS_PCG_JACOBI || desc->method == MFS_CG || desc->method == MFS_GMRES ||
                    desc->method == MFS_MINRES || desc->method == MFS_BICGSTAB,
                "mfs_solver_create: unknown method");
  MFS_CHECK_ARG(desc->precision == MFS_PREC_F64 || desc->precision == MFS_PREC_MIXED,
                "mfs_solver_create: unknown precision");
  MFS_CHECK_ARG(desc->tol > 0 && desc->maxit > 0, "mfs_solver_create: tol and maxit must be positive");
  MFS_CHECK_ARG(desc->method != MFS_GMRES || (desc->restart >= 1 && desc->restart <= 30),
                "mfs_solver_create: GMRES restart must be in 1..30");
  MFS_HIP(hipSetDevice(A->ctx->device));
  *out = nullptr;
  mfs_solver S = new mfs_solver_s();
  S->A = A;
  mat_retain(A);
  S->desc = *desc;
  S->n = A->n();
  const int64_t n = S->n;
  int chunk = desc->check_every > 0 ? desc->check_every : 16;
  int s = MFS_OK;
  auto tryit = [&](int v) {
    if (s == MFS_OK) s = v;
  };
  tryit(dmalloc((void**)&S->x64, 8 * n));
  tryit(dmalloc((void**)&S->b64, 8 * n));
  tryit(dmalloc((void**)&S->r64, 8 * n));
  tryit(dmalloc((void**)&S->dout, 8 * 64));
  if (s == MFS_OK) tryit(reduce_scratch_alloc(S->red64, 2 * spmv_grid(A) + 64));
  const bool precond = desc->method != MFS_CG;
  if (desc->method == MFS_BICGSTAB) {
    if (desc->precision == MFS_PREC_F64)
      tryit(S->bs64.init(A, true));
    else
      tryit(S->bs32.init(A, true));
  } else if (desc->method == MFS_MINRES) {
    const size_t sv = desc->precision == MFS_PREC_F64 ? 8 : 4;
    if (desc->precision == MFS_PREC_MIXED && !A->val32_valid) tryit(mat_refresh_derived(A));
    for (void*& v : S->mr) tryit(dmalloc(&v, sv * n));
    tryit(dmalloc((void**)&S->dh, 8 * 64));
    if (s == MFS_OK) tryit(reduce_scratch_alloc(S->redg, 2 * 1024 + 64));
  } else if (desc->method != MFS_GMRES) {
    if (s == MFS_OK && desc->precision == MFS_PREC_F64) {
      S->has64 = true;
      tryit(S->cg64.init(A, precond, chunk));
    }
    if (s == MFS_OK && desc->precision == MFS_PREC_MIXED) {
      S->has32 = true;
      tryit(S->cg32.init(A, precond, chunk));
    }
  } else {
    const size_t sv = desc->precision == MFS_PREC_F64 ? 8 : 4;
    if (desc->precision == MFS_PREC_MIXED && !A->val32_valid) tryit(mat_refresh_derived(A));
    tryit(dmalloc(&S->V, sv * n * (desc->restart + 1)));
    tryit(dmalloc(&S->w, sv * n));
    tryit(dmalloc(&S->z, sv * n));
    // 64 doubles of h / y, then the rotation state of k_gmres_rot (m <= 30)
    tryit(dmalloc((void**)&S->dh, 8 * (64 + 31 * 30 + 3 * 30 + 1 + 3 * 30)));
    if (s == MFS_OK) tryit(reduce_scratch_alloc(S->redg, 32 * 1024 + 64));
  }
  if (s != MFS_OK) {
    mfs_solver_destroy(S);
    return s;
  }
  *out = S;
  return MFS_OK;
}

// r64 = b - A x64; returns ||r||^2, ||b||^2 (host), one sync.
static int true_residual(mfs_solver S, double* rr, double* bb) {
  hipStream_t s = S->A->ctx->stream;
  MFS_TRY(spmv_residual_launch(S->A, MFS_F64, S->x64, S->r64, S->b64, S->red64.partials, S->red64.counter, S->dout, s));
  double* hp = (double*)S->A->ctx->pinned;
  MFS_HIP(hipMemcpyAsync(hp, S->dout, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  *rr = hp[0];
  *bb = hp[1];
  return MFS_OK;
}

// GMRES Arnoldi step j on the device (one thread): column j of H from the
// two CGS2 passes (dh[i] + dh[32 + i], i <= j) and ||w||^2 (dh[62]), the
// previous Givens rotations applied, the new rotation, g updated and
// est[j] = |g[j+1]|.  gm: H[(m+1) m] (row i, column j at i m + j), cs[m],
// sn[m], g[m+1], est[m], hn[m], bad[m] (1: zero or nonfinite pivot; later
// steps then change nothing).
__global__ __launch_bounds__(64) void k_gmres_rot(const double* __restrict__ dh, double* __restrict__ gm, int j,
                                                  int m) {
  // the column and the previous rotations are staged in LDS by the wave, so
  // the serial rotation chain of thread 0 reads no global memory
  __shared__ double col[32], c_s[32], s_s[32];
  double* H = gm;
  double* cs = H + (m + 1) * m;
  double* sn = cs + m;
  double* g = sn + m;
  double* est = g + m + 1;
  double* hnv = est + m;
  double* bad = hnv + m;
  const int tid = threadIdx.x;
  if (j > 0 && bad[j - 1] != 0.0) {
    if (tid == 0) bad[j] = 1.0;
    return;
  }
  if (tid <= j) col[tid] = dh[tid] + dh[32 + tid];
  if (tid < j) {
    c_s[tid] = cs[tid];
    s_s[tid] = sn[tid];
  }
  __syncthreads();
  if (tid != 0) return;
  const double hn = sqrt(fmax(dh[62], 0.0));
  double below = hn;  // entry j + 1 of the column
  for (int i = 0; i < j; i++) {
    const double a = col[i], c = col[i + 1];
    col[i] = c_s[i] * a + s_s[i] * c;
    col[i + 1] = -s_s[i] * a + c_s[i] * c;
  }
  const double a = col[j], c = below;
  const double d = sqrt(a * a + c * c);
  hnv[j] = hn;
  for (int i = 0; i < j; i++) H[i * m + j] = col[i];
  if (!(d > 0) || !isfinite(d)) {
    bad[j] = 1.0;
    return;
  }
  bad[j] = 0.0;
  const double cj = a / d, sj = c / d;
  cs[j] = cj;
  sn[j] = sj;
  H[j * m + j] = d;
  H[(j + 1) * m + j] = 0;
  const double gj = g[j];
  g[j + 1] = -sj * gj;
  g[j] = cj * gj;
  est[j] = fabs(sj * gj);
}

__global__ void k_gmres_rot_init(double* __restrict__ gm, int m, double beta) {
  double* g = gm + (m + 1) * m + 2 * m;
  for (int i = threadIdx.x; i <= m; i += blockDim.x) g[i] = i == 0 ? beta : 0.0;
}

template <typename T>
static int gmres_cycle(mfs_solver S, const T* val, const T* dinv, T* rhs, double beta, double tol_abs, int m,
                       int maxsteps, int* steps_done, double* est_abs, T* u_out) {
  // Solve A D^-1 y = rhs approximately with one GMRES(m) cycle from zero;
  // u_out = D^-1 V y.  rhs has norm beta.  Returns estimated residual norm.
  hipStream_t s = S->A->ctx->stream;
  mfs_mat A = S->A;
  const int64_t n = S->n;
  T* V = (T*)S->V;
  T* w = (T*)S->w;
  T* z = (T*)S->z;
  const unsigned gv = grid_for(n, 1024);
  const int prec = std::is_same<T, double>::value ? MFS_F64 : MFS_F32;
  // The Arnoldi steps and their Givens rotations run on the device; the
  // host reads the residual estimates every kSync steps (not every step) and
  // uses the first step whose estimate reaches tol_abs, so the steps run
  // past it are ignored and the result is that of a per-step check.
  constexpr int kSync = 6;
  double* gm = S->dh + 64;
  // one Arnoldi step with its rotation; replayed from a graph per j (the
  // launches of a step are otherwise the host-side bottleneck at ~12 per
  // step), directly with MFS_NO_GRAPH=1 (kernel tracing)
  const char* ng = std::getenv("MFS_NO_GRAPH");
  const bool use_graph = !(ng && ng[0] == '1');
  // the FP32 copy of the values is refreshed outside any capture
  if (prec == MFS_F32 && !A->val32_valid) MFS_TRY(mat_refresh_derived(A));
  const unsigned gmd = std::max(1u, gv / 4);  // multi-dot blocks per vector (4 rows per thread)
  auto arnoldi_step = [&](int jj) -> int {
    k_dscale<T><<<gv, 256, 0, s>>>(V + (int64_t)jj * n, dinv, z, n);
    MFS_TRY(spmv_launch(A, prec, z, w, s));
    // CGS2: two passes of h = V^T w, w -= V h
    for (int pass = 0; pass < 2; pass++) {
      k_multidot<T><<<dim3(gmd, jj + 1), 256, 0, s>>>(V, w, n, S->redg.partials);
      k_finish_sums<<<jj + 1, 256, 0, s>>>(S->redg.partials, gmd, S->dh + pass * 32);
      k_orth_update<T><<<gv, 256, 0, s>>>(V, w, n, jj + 1, S->dh + pass * 32, S->redg.partials);
      // ||w||^2 is needed after the second pass only
      if (pass == 1) k_finish_sums<<<1, 256, 0, s>>>(S->redg.partials, gv, S->dh + 62);
    }
    k_scale_into<T><<<gv, 256, 0, s>>>(w, V + (int64_t)(jj + 1) * n, S->dh + 62, n);
    k_gmres_rot<<<1, 64, 0, s>>>(S->dh, gm, jj, m);
    return MFS_OK;
  };
  double* hp = (double*)A->ctx->pinned;
  hp[0] = beta * beta;
  MFS_HIP(hipMemcpyAsync(S->dh + 40, hp, 8, hipMemcpyHostToDevice, s));
  k_scale_into<T><<<gv, 256, 0, s>>>(rhs, V, S->dh + 40, n);
  k_gmres_rot_init<<<1, 64, 0, s>>>(gm, m, beta);
  const size_t off_est = (size_t)(m + 1) * m + 3 * m + 1;  // est, hn, bad follow
  int j = 0, used = -1;
  double est = beta;
  while (j < m && *steps_done + j < maxsteps && used < 0) {
    const int j0 = j;
    const int nb = std::min(kSync, std::min(m - j, maxsteps - *steps_done - j));
    if (nb <= 0) break;
    for (int t = 0; t < nb; t++, j++) {
      if (use_graph) {
        hipGraphExec_t& ge = S->gstep[std::is_same<T, double>::value ? 0 : 1][j];
        if (!ge) {
          hipGraph_t g;
          MFS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          const int st = arnoldi_step(j);
          MFS_HIP(hipStreamEndCapture(s, &g));
          MFS_TRY(st);
          MFS_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
          (void)hipGraphDestroy(g);
        }
        MFS_HIP(hipGraphLaunch(ge, s));
      } else {
        MFS_TRY(arnoldi_step(j));
      }
    }
    MFS_HIP(hipGetLastError());
    MFS_HIP(hipMemcpyAsync(hp, gm + off_est, 3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    const double* est_h = hp;
    const double* hn_h = hp + m;
    const double* bad_h = hp + 2 * m;
    for (int t = j0; t < j; t++) {
      if (bad_h[t] != 0.0) return MFS_ERR_NONFINITE;
      est = est_h[t];
      if (est <= tol_abs || hn_h[t] == 0) {
        used = t + 1;
        break;
      }
    }
  }
  if (used < 0) used = j;
  *steps_done += used;
  j = used;
  // H (upper triangle of the first j columns) and g for the back substitution
  std::vector<double> H((size_t)(m + 1) * m), g(m + 1);
  MFS_HIP(hipMemcpyAsync(hp, gm, ((size_t)(m + 1) * m + 3 * m + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  std::copy(hp, hp + (size_t)(m + 1) * m, H.begin());
  std::copy(hp + (size_t)(m + 1) * m + 2 * m, hp + (size_t)(m + 1) * m + 3 * m + 1, g.begin());
  std::vector<double> y(j);
  for (int i = j - 1; i >= 0; i--) {
    double sacc = g[i];
    for (int k2 = i + 1; k2 < j; k2++) sacc -= H[i * m + k2] * y[k2];
    y[i] = sacc / H[i * m + i];
  }
  for (int i = 0; i < j; i++) hp[i] = y[i];
  MFS_HIP(hipMemcpyAsync(S->dh, hp, j * sizeof(double), hipMemcpyHostToDevice, s));
  k_combine<T><<<gv, 256, 0, s>>>(V, S->dh, j, z, n);
  k_dscale<T><<<gv, 256, 0, s>>>(z, dinv, u_out, n);
  MFS_HIP(hipGetLastError());
  *est_abs = est;
  return MFS_OK;
}

// One MINRES run on A d = rhs (rhs in mr[0]; d returned in mr[1]), to
// relative (M^-1-norm) tolerance rtol or maxsteps.  Returns the steps taken.
template <typename T>
static int minres_run(mfs_solver S, const T* dinv, double rtol, int maxsteps, int* steps, double* est_rel) {
  hipStream_t s = S->A->ctx->stream;
  mfs_mat A = S->A;
  const int64_t n = S->n;
  const int prec = std::is_same<T, double>::value ? MFS_F64 : MFS_F32;
  T* rhs = (T*)S->mr[0];
  T* x = (T*)S->mr[1];
  T* v[3] = {(T*)S->mr[2], (T*)S->mr[3], (T*)S->mr[4]};  // v_{j-1}, v_j, v_{j+1}
  T* w[3] = {(T*)S->mr[5], (T*)S->mr[6], (T*)S->mr[7]};  // w_{j-1}, w_j, w_{j+1}
  T* zbuf[2] = {(T*)S->mr[8], nullptr};
  T* Az = (T*)S->mr[9];
  for (void* b : S->mr)
    if (!b) {
      set_error("minres: work vector not allocated");
      return MFS_ERR_ARG;
    }
  if (!S->dh || !S->redg.partials) {
    set_error("minres: reduction scratch not allocated");
    return MFS_ERR_ARG;
  }
  const unsigned gv = grid_for(n, 1024);
  double* hp = (double*)A->ctx->pinned;
  // z buffers: z_j in zbuf[0]; z_{j+1} in the spare v slot is not possible,
  // so z_{j+1} goes to rhs (rhs is consumed by the start kernel).
  T* zj = zbuf[0];
  T* znext = rhs;
  k_minres_start<T><<<gv, 256, 0, s>>>(rhs, v[1], zj, dinv, x, v[0], w[1], w[0], n, S->redg.partials);
  k_finish_sums<<<1, 256, 0, s>>>(S->redg.partials, gv, S->dh);
  MFS_HIP(hipMemcpyAsync(hp, S->dh, 8, hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  const double g = std::sqrt(std::max(hp[0], 0.0));
  if (!(g > 0) || !std::isfinite(g)) {
    *est_rel = std::isfinite(g) ? 0.0 : 1.0;
    return std::isfinite(g) ? MFS_OK : MFS_ERR_NONFINITE;
  }
  // recurrence state on the device; the host checks it every kSync steps
  // (steps launched past the stopping one leave x unchanged)
  constexpr int kSync = 8;
  double* ms = S->dh + 8;
  {
    double init[kMsCount] = {};
    init[kMsG] = g;
    init[kMsC] = init[kMsCPrev] = 1.0;
    init[kMsEta] = init[kMsEta0] = g;
    init[kMsRtol] = rtol;
    init[kMsStop] = -1.0;
    for (int q = 0; q < kMsCount; q++) hp[q] = init[q];
    MFS_HIP(hipMemcpyAsync(ms, hp, sizeof(double) * kMsCount, hipMemcpyHostToDevice, s));
  }
  k_scale_vec<T><<<gv, 256, 0, s>>>(zj, 1.0 / g, n);
  int j = 0;
  double stop = -1.0, bad = 0.0;
  while (j < maxsteps && stop < 0.0) {
    const int nb = std::min(kSync, maxsteps - j);
    for (int t = 0; t < nb; t++, j++) {
      MFS_TRY(spmv_launch(A, prec, zj, Az, s));
      // delta = <A z_j, z_j>
      k_multidot<T><<<dim3(gv, 1), 256, 0, s>>>(zj, Az, n, S->redg.partials);
      k_finish_sums<<<1, 256, 0, s>>>(S->redg.partials, gv, S->dh + 1);
      k_minres_lanczos<T><<<gv, 256, 0, s>>>(Az, v[1], v[0], v[2], znext, dinv, S->dh + 1, ms, n,
                                             S->redg.partials);
      k_finish_sums<<<1, 256, 0, s>>>(S->redg.partials, gv, S->dh + 2);
      k_minres_rot<<<1, 64, 0, s>>>(S->dh, ms, j);
      k_minres_update<T><<<gv, 256, 0, s>>>(zj, w[1], w[0], w[2], x, znext, ms, j, n);
      // rotate the buffers: v_{j-1} <- v_j, v_j <- v_{j+1} (the same for w), z
      std::swap(v[0], v[1]);
      std::swap(v[1], v[2]);
      std::swap(w[0], w[1]);
      std::swap(w[1], w[2]);
      std::swap(zj, znext);
    }
    MFS_HIP(hipGetLastError());
    MFS_HIP(hipMemcpyAsync(hp, ms, sizeof(double) * kMsCount, hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    stop = hp[kMsStop];
    bad = hp[kMsBad];
  }
  if (bad != 0.0) return MFS_ERR_NONFINITE;
  if (stop >= 0.0) j = (int)stop + 1;
  const double eta = hp[kMsEta], eta0 = hp[kMsEta0];
  // the correction must end in mr[1]: x never moves
  *steps += j;
  *est_rel = std::fabs(eta) / eta0;
  return MFS_OK;
}

__global__ void k_add_f64(double* __restrict__ x, const double* __restrict__ u, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] += u[i];
}

extern "C" int mfs_solver_run(mfs_solver S, mfs_vec b, mfs_vec x, mfs_solve_report* rep) {
  MFS_CHECK_ARG(S && b && x && rep, "mfs_solver_run: bad argument");
  MFS_CHECK_ARG(b->n == S->n && x->n == S->n && b->prec == MFS_F64 && x->prec == MFS_F64,
                "mfs_solver_run: b and x must be FP64 vectors of the matrix size");
  mfs_mat A = S->A;
  hipStream_t s = A->ctx->stream;
  const int64_t n = S->n;
  const mfs_solve_desc& d = S->desc;
  std::memset(rep, 0, sizeof(*rep));
  MFS_HIP(hipStreamSynchronize(s));
  auto t0 = std::chrono::steady_clock::now();
  MFS_HIP(hipMemcpyAsync(S->b64, b->d, 8 * n, hipMemcpyDeviceToDevice, s));
  MFS_HIP(hipMemcpyAsync(S->x64, x->d, 8 * n, hipMemcpyDeviceToDevice, s));
  double rr, bb;
  MFS_TRY(true_residual(S, &rr, &bb));
  if (bb == 0.0) bb = 1.0;
  if (!std::isfinite(rr)) {
    rep->status = MFS_ERR_NONFINITE;
    rep->true_relres = rr;
    mfs::set_error("mfs_solver_run: non-finite initial residual");
    return MFS_ERR_NONFINITE;
  }
  double rel = std::sqrt(rr / bb);
  int iters = 0, outer = 0;
  int status = MFS_OK;
  const unsigned gv = grid_for(n, 1024);
  const bool is_cg = d.method == MFS_PCG_JACOBI || d.method == MFS_CG;
  // plain FP64 (P)CG on b - A x0: r = r64 already; x correction starts at x0
  // FP64 PCG to tol.  When the recursion reports convergence but the true
  // residual is still above tol (the rounding floor of long runs on
  // ill-conditioned or indefinite operators, where the recursive and true
  // residuals part by a few per cent), the iteration continues from the
  // true residual with the kept search direction (residual replacement, as
  // in the mixed path), for at most 8 steps.
  auto run_f64 = [&](auto& E) -> int {
    MFS_HIP(hipMemcpyAsync(E.r, S->r64, 8 * n, hipMemcpyDeviceToDevice, s));
    MFS_HIP(hipMemcpyAsync(E.x, S->x64, 8 * n, hipMemcpyDeviceToDevice, s));
    E.can_keep = false;
    double keep_ratio = 0.0;
    for (int step = 0; step < 8; step++) {
      CgState hs;
      MFS_TRY(E.run(d.tol, bb, d.maxit - iters, &hs, keep_ratio));
      iters += hs.iter1;
      outer++;
      MFS_HIP(hipMemcpyAsync(S->x64, E.x, 8 * n, hipMemcpyDeviceToDevice, s));
      const double predicted = std::sqrt(hs.rr / bb);
      rep->recursive_relres = predicted;
      MFS_TRY(true_residual(S, &rr, &bb));
      rel = std::sqrt(rr / bb);
      if (hs.done == CG_BREAKDOWN || !std::isfinite(rel)) {
        status = MFS_ERR_NONFINITE;
        break;
      }
      if (rel <= d.tol || hs.done != CG_CONVERGED || iters >= d.maxit) break;
      // continue from the true residual (same units); restart instead when
      // the recursion has drifted far from it
      MFS_HIP(hipMemcpyAsync(E.r, S->r64, 8 * n, hipMemcpyDeviceToDevice, s));
      keep_ratio = (E.can_keep && rel <= 2.0 * predicted) ? 1.0 : 0.0;
    }
    if (status == MFS_OK && rel > d.tol) status = MFS_ERR_NOT_CONVERGED;
    return MFS_OK;
  };
  // FP64 outer refinement around FP32 (P)CG.  When the FP32 recursion
  // stayed consistent with the true residual (the FP64 residual after the
  // step is within 2x of what the recursion predicted), the next step
  // continues the FP32 iteration instead of restarting it: the recursive
  // residual is replaced by the true FP64 one and the search directions are
  // kept (rescaled to the new units), which preserves the Krylov
  // information.  When the two have drifted apart the kept directions would
  // no longer be conjugate, so the step restarts.  A continued step that
  // takes more than twice the first step's iterations is cut short and the
  // next one restarts.
  auto run_mixed = [&](auto& E) -> int {
    E.can_keep = false;
    double keep_ratio = 0.0;
    int first_inner = 0;
    while (rel > d.tol && iters < d.maxit) {
      const double rn = std::sqrt(rr);
      k_demote_scaled<<<gv, 256, 0, s>>>(S->r64, 1.0 / rn, E.r, E.x, n);
      // inner tolerance: no tighter than needed to reach tol in this step
      const double itol = std::min(0.5, std::max(d.inner_tol, 0.5 * d.tol / rel));
      CgState hs;
      const int cap = keep_ratio > 0.0 ? std::min(d.maxit - iters, 2 * first_inner + 50) : d.maxit - iters;
      MFS_TRY(E.run(itol, 1.0, cap, &hs, keep_ratio));
      if (outer == 0) first_inner = hs.iter1;
      iters += hs.iter1;
      outer++;
      k_promote_add<<<gv, 256, 0, s>>>(S->x64, E.x, nullptr, rn, n);
      const double prev = rel;
      MFS_TRY(true_residual(S, &rr, &bb));
      rel = std::sqrt(rr / bb);
      rep->recursive_relres = std::sqrt(hs.rr) * prev;
      const double predicted = std::sqrt(hs.rr) * prev;
      keep_ratio = (E.can_keep && rel > 0.0 && rel <= 2.0 * predicted) ? rn / std::sqrt(rr) : 0.0;
      if (!std::isfinite(rel) || hs.done == CG_BREAKDOWN) {
        status = MFS_ERR_NONFINITE;
        break;
      }
      if (rel > 0.99 * prev && hs.iter1 <= 1) {  // no progress possible
        status = MFS_ERR_NOT_CONVERGED;
        break;
      }
    }
    if (status == MFS_OK && rel > d.tol) status = MFS_ERR_NOT_CONVERGED;
    return MFS_OK;
  };
  if (is_cg && d.precision == MFS_PREC_F64) {
    MFS_TRY(run_f64(S->cg64));
  } else if (is_cg) {
    MFS_TRY(run_mixed(S->cg32));
  } else if (d.method == MFS_BICGSTAB) {
    // BiCGStab runs, each from the current FP64 residual: FP64 to tol, or
    // FP32 to the inner tolerance with FP64 refinement around it
    while (rel > d.tol && iters < d.maxit) {
      const double rn = std::sqrt(rr);
      BsState hs;
      if (d.precision == MFS_PREC_F64) {
        MFS_HIP(hipMemcpyAsync(S->bs64.r, S->r64, 8 * n, hipMemcpyDeviceToDevice, s));
        MFS_TRY(S->bs64.run(0.5 * d.tol * std::sqrt(bb) / rn, d.maxit - iters, &hs));
        k_add_f64<<<gv, 256, 0, s>>>(S->x64, S->bs64.x, n);
      } else {
        float* dummy = S->bs32.v;
        k_demote_scaled<<<gv, 256, 0, s>>>(S->r64, 1.0 / rn, S->bs32.r, dummy, n);
        const double itol = std::min(0.5, std::max(d.inner_tol, 0.5 * d.tol / rel));
        MFS_TRY(S->bs32.run(itol, d.maxit - iters, &hs));
        k_promote_add<<<gv, 256, 0, s>>>(S->x64, S->bs32.x, nullptr, rn, n);
      }
      MFS_HIP(hipGetLastError());
      iters += hs.iter;
      outer++;
      const double prev = rel;
      MFS_TRY(true_residual(S, &rr, &bb));
      rel = std::sqrt(rr / bb);
      // the FP64 run solved for the unscaled residual, the FP32 run for r / ||r||
      rep->recursive_relres = d.precision == MFS_PREC_F64 ? std::sqrt(hs.rr / bb) : std::sqrt(hs.rr) * prev;
      if (!std::isfinite(rel)) {
        status = MFS_ERR_NONFINITE;
        break;
      }
      if (hs.done == BS_BREAKDOWN && !(rel < 0.99 * prev)) {
        status = MFS_ERR_NONFINITE;
        break;
      }
      if (rel > 0.99 * prev && hs.iter <= 1) {
        status = MFS_ERR_NOT_CONVERGED;
        break;
      }
    }
    if (status == MFS_OK && rel > d.tol) status = MFS_ERR_NOT_CONVERGED;
  } else if (d.method == MFS_MINRES) {
    // MINRES runs, each from the current FP64 residual (refinement steps)
    while (rel > d.tol && iters < d.maxit) {
      const double rn = std::sqrt(rr);
      double est = 1;
      int steps = 0, st2;
      if (d.precision == MFS_PREC_F64) {
        MFS_HIP(hipMemcpyAsync(S->mr[0], S->r64, 8 * n, hipMemcpyDeviceToDevice, s));
        st2 = minres_run<double>(S, A->d_dinv64, 0.5 * d.tol * std::sqrt(bb) / rn, d.maxit - iters, &steps, &est);
        if (st2 == MFS_OK) k_add_f64<<<gv, 256, 0, s>>>(S->x64, (const double*)S->mr[1], n);
      } else {
        float* dummy = (float*)S->mr[2];
        k_demote_scaled<<<gv, 256, 0, s>>>(S->r64, 1.0 / rn, (float*)S->mr[0], dummy, n);
        const double itol = std::min(0.5, std::max(d.inner_tol, 0.5 * d.tol / rel));
        st2 = minres_run<float>(S, A->d_dinv32, itol, d.maxit - iters, &steps, &est);
        if (st2 == MFS_OK) k_promote_add<<<gv, 256, 0, s>>>(S->x64, (const float*)S->mr[1], nullptr, rn, n);
      }
      if (st2 != MFS_OK) {
        status = st2;
        break;
      }
      MFS_HIP(hipGetLastError());
      iters += steps;
      outer++;
      const double prev = rel;
      MFS_TRY(true_residual(S, &rr, &bb));
      rel = std::sqrt(rr / bb);
      rep->recursive_relres = est * prev;
      if (!std::isfinite(rel)) {
        status = MFS_ERR_NONFINITE;
        break;
      }
      if (rel > 0.99 * prev && steps <= 1) {
        status = MFS_ERR_NOT_CONVERGED;
        break;
      }
    }
    if (status == MFS_OK && rel > d.tol) status = MFS_ERR_NOT_CONVERGED;
  } else {
    // GMRES(m): every restart recomputes the true residual in FP64
    const int m = d.restart;
    while (rel > d.tol && iters < d.maxit) {
      const double rn = std::sqrt(rr);
      double est;
      int steps = iters;
      if (d.precision == MFS_PREC_F64) {
        // rhs = r64 (FP64), correction into r64's spare: use w as output
        double* u = (double*)S->z;
        MFS_HIP(hipMemcpyAsync(S->w, S->r64, 8 * n, hipMemcpyDeviceToDevice, s));
        // gmres_cycle reads rhs from a buffer distinct from w/z: stage in V's last column
        double* rhs = (double*)S->V + (int64_t)m * n;
        MFS_HIP(hipMemcpyAsync(rhs, S->r64, 8 * n, hipMemcpyDeviceToDevice, s));
        // output buffer must not alias z (used inside): use r64 after staging
        int st2 = gmres_cycle<double>(S, A->d_val64, A->d_dinv64, rhs, rn, d.tol * std::sqrt(bb) * 0.5, m,
                                      d.maxit, &steps, &est, S->r64);
        if (st2 != MFS_OK) {
          status = st2;
          break;
        }
        (void)u;
        k_add_f64<<<gv, 256, 0, s>>>(S->x64, S->r64, n);
      } else {
        float* rhs = (float*)S->V + (int64_t)m * n;
        float* dummy = (float*)S->w;
        k_demote_scaled<<<gv, 256, 0, s>>>(S->r64, 1.0 / rn, rhs, dummy, n);
        const double itol = std::min(0.5, std::max(d.inner_tol, 0.5 * d.tol / rel));
        float* u32 = (float*)S->w;
        // u32 aliases w which gmres_cycle uses: write the result into rhs's
        // slot instead (rhs is consumed at the start of the cycle).
        int st2 = gmres_cycle<float>(S, A->d_val32, A->d_dinv32, rhs, 1.0, itol, m, d.maxit, &steps, &est, rhs);
        (void)u32;
        if (st2 != MFS_OK) {
          status = st2;
          break;
        }
        k_promote_add<<<gv, 256, 0, s>>>(S->x64, rhs, nullptr, rn, n);
      }
      MFS_HIP(hipGetLastError());
      iters = steps;
      outer++;
      const double prev = rel;
      MFS_TRY(true_residual(S, &rr, &bb));
      rel = std::sqrt(rr / bb);
      rep->recursive_relres = d.precision == MFS_PREC_F64 ? est / std::sqrt(bb) : est * prev;
      if (!std::isfinite(rel)) {
        status = MFS_ERR_NONFINITE;
        break;
      }
    }
    if (status == MFS_OK && rel > d.tol) status = MFS_ERR_NOT_CONVERGED;
  }
  MFS_HIP(hipMemcpyAsync(x->d, S->x64, 8 * n, hipMemcpyDeviceToDevice, s));
  MFS_HIP(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  rep->iterations = iters;
  rep->outer_steps = outer;
  rep->true_relres = rel;
  rep->seconds = std::chrono::duration<double>(t1 - t0).count();
  rep->status = status;
  if (status == MFS_ERR_NOT_CONVERGED) mfs::set_error("mfs_solver_run: tolerance not reached within maxit");
  if (status == MFS_ERR_NONFINITE) mfs::set_error("mfs_solver_run: non-finite value or breakdown in the iteration");
  return status;
}

namespace {
__global__ void k_fill(float* a, double* b, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (a) a[i] = (float)v;
    if (b) b[i] = v;
  }
}

template <typename T>
int time_cg_kernel(PcgEngine<T>& E, int kernel, int reps, double* avg_ms, double* bytes) {
  mfs_mat A = E.A;
  hipStream_t s = A->ctx->stream;
  const int64_t n = E.n;
  const bool f64 = std::is_same<T, double>::value;
  const unsigned gv = grid_for(n, 1024);
  // r = 1, x = 0, then a fresh start with tol 0 so no launch stops early
  k_fill<<<gv, 256, 0, s>>>(f64 ? nullptr : (float*)E.r, f64 ? (double*)E.r : nullptr, 1.0, n);
  k_fill<<<gv, 256, 0, s>>>(f64 ? nullptr : (float*)E.x, f64 ? (double*)E.x : nullptr, 0.0, n);
  k_cg_start<T><<<E.g_vec, 256, 0, s>>>(E.z, E.pbuf[0], E.r, E.dinv, n, E.part_rz());
  k_cg_start_state<<<1, 64, 0, s>>>(E.st, 0.0, 1.0, 1 << 30, 0, 0.0);
  E.finish_rz(s);
  // one K1 so that K2 has partials to read
  E.launch_iteration(s, 0);
  MFS_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  MFS_HIP(hipEventCreate(&e0));
  MFS_HIP(hipEventCreate(&e1));
  MFS_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) {
    if (kernel == 0) {
      E.launch_k1(s, E.pbuf[0], E.pbuf[1]);
    } else {
      k_cg_update<T><<<E.g_vec, 256, 0, s>>>(E.x, E.pbuf[1], E.r, E.z, E.q, E.dinv, n, E.st, E.pq_src(), E.pq_g(),
                                             E.part_rz(), E.A->xcd_map, E.g_spmv, E.A->bs);
    }
  }
  MFS_HIP(hipEventRecord(e1, s));
  MFS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the probe must not have stopped (every launch did the full work)
  CgState hs;
  MFS_HIP(hipMemcpy(&hs, E.st, sizeof(hs), hipMemcpyDeviceToHost));
  if (hs.done != CG_RUN) {
    set_error("mfs_solver_kernel_time: the timed iteration stopped early");
    return MFS_ERR_NONFINITE;
  }
  *avg_ms = ms / reps;
  const double sv = f64 ? 8.0 : 4.0;
  // K1: matrix + z and p gathered (+ own rows) + p_new and q written;
  // K2: x, p, r, q (+ dinv) read, x, r, z written
  *bytes = kernel == 0 ? spmv_bytes(A, f64 ? MFS_F64 : MFS_F32) + 2.0 * (double)n * sv
                       : (double)n * sv * (E.dinv ? 8.0 : 7.0);
  return MFS_OK;
}
}  // namespace

extern "C" int mfs_solver_kernel_time(mfs_solver S, int kernel, int reps, double* avg_ms, double* bytes) {
  MFS_CHECK_ARG(S && avg_ms && bytes && reps > 0 && (kernel == 0 || kernel == 1),
                "mfs_solver_kernel_time: bad argument");
  const mfs_solve_desc& d = S->desc;
  MFS_CHECK_ARG(d.method == MFS_PCG_JACOBI || d.method == MFS_CG,
                "mfs_solver_kernel_time: only the CG methods have these kernels");
  MFS_HIP(hipSetDevice(S->A->ctx->device));
  if (d.precision == MFS_PREC_MIXED) return time_cg_kernel(S->cg32, kernel, reps, avg_ms, bytes);
  return time_cg_kernel(S->cg64, kernel, reps, avg_ms, bytes);
}

extern "C" int mfs_solve(mfs_mat A, const mfs_solve_desc* desc, mfs_vec b, mfs_vec x, mfs_solve_report* report) {
  mfs_solver S;
  MFS_TRY(mfs_solver_create(A, desc, &S));
  int st = mfs_solver_run(S, b, x, report);
  mfs_solver_destroy(S);
  return st;
}

// ---------------------------------------------------------------------------
// mfs_solve without device allocation: the solver's work vectors, state and
// reduction partials are carved from the context's fixed scratch (a device
// array of the library, 64 MiB per device), kernels are launched directly
// (no graphs) and the host waits for each iteration chunk on the stream.
// Same kernels and iteration order as mfs_solve, so the same iterates.
// ---------------------------------------------------------------------------
static int noalloc_layout(mfs_solver_s& S, Arena& ar) {
  const int64_t n = S.n;
  S.x64 = ar.take<double>(n);
  S.b64 = ar.take<double>(n);
  S.r64 = ar.take<double>(n);
  S.dout = ar.take<double>(64);
  S.red64.capacity = 2 * spmv_grid(S.A) + 64;
  S.red64.partials = ar.take<double>(S.red64.capacity);
  S.red64.counter = ar.take<unsigned>(16);
  const bool precond = S.desc.method != MFS_CG;
  const int chunk = S.desc.check_every > 0 ? S.desc.check_every : 16;
  if (S.desc.precision == MFS_PREC_F64) {
    S.has64 = true;
    MFS_TRY(S.cg64.init_ext(S.A, precond, chunk, ar));
  } else {
    S.has32 = true;
    MFS_TRY(S.cg32.init_ext(S.A, precond, chunk, ar));
  }
  return MFS_OK;
}

extern "C" int mfs_solve_noalloc(mfs_mat A, const mfs_solve_desc* desc, mfs_vec b, mfs_vec x,
                                 mfs_solve_report* report) {
  MFS_CHECK_ARG(A && desc && b && x && report, "mfs_solve_noalloc: bad argument");
  MFS_CHECK_ARG(desc->method == MFS_PCG_JACOBI || desc->method == MFS_CG,
                "mfs_solve_noalloc: method must be MFS_PCG_JACOBI or MFS_CG");
  MFS_CHECK_ARG(desc->precision == MFS_PREC_F64 || desc->precision == MFS_PREC_MIXED,
                "mfs_solve_noalloc: unknown precision");
  MFS_CHECK_ARG(desc->tol > 0 && desc->maxit > 0, "mfs_solve_noalloc: tol and maxit must be positive");
  mfs_ctx ctx = A->ctx;
  MFS_CHECK_ARG(ctx->noalloc_scratch && ctx->noalloc_scratch_bytes > 0, "mfs_solve_noalloc: no scratch on this device");
  MFS_HIP(hipSetDevice(ctx->device));
  // size check first (measuring pass, nothing touched)
  {
    mfs_solver_s M;
    M.A = A;
    M.desc = *desc;
    M.n = A->n();
    Arena meas;
    meas.cap = ~size_t(0) >> 1;
    MFS_TRY(noalloc_layout(M, meas));
    if (meas.off > ctx->noalloc_scratch_bytes) {
      set_error("mfs_solve_noalloc: this system needs " + std::to_string(meas.off) +
                " bytes of scratch, the fixed scratch holds " + std::to_string(ctx->noalloc_scratch_bytes) +
                " (use mfs_solve for larger systems)");
      return MFS_ERR_ARG;
    }
  }
  mfs_solver_s S;
  S.A = A;
  S.desc = *desc;
  S.n = A->n();
  Arena ar;
  ar.base = (char*)ctx->noalloc_scratch;
  ar.cap = ctx->noalloc_scratch_bytes;
  MFS_TRY(noalloc_layout(S, ar));
  if (!ar.ok) {
    set_error("mfs_solve_noalloc: scratch layout overflow");
    return MFS_ERR_ARG;
  }
  MFS_HIP(hipMemsetAsync(S.red64.counter, 0, 16 * sizeof(unsigned), ctx->stream));
  // S owns nothing: it is not passed to mfs_solver_destroy
  return mfs_solver_run(&S, b, x, report);
}

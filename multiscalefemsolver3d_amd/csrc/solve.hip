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
#include <cstring>
#include <vector>

#include "mfs_device.hpp"

using namespace mfs;

namespace {

struct CgState {
  double rz, pq, alpha, beta, rr, bb, tol2, rz0;
  int done, iter, maxit, pad;
};
// done codes
enum { CG_RUN = 0, CG_CONVERGED = 1, CG_BREAKDOWN = 2, CG_MAXIT = 3 };

inline unsigned grid_for(int64_t n, int64_t cap) {
  int64_t g = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// q = A p, grid p.q; last block: alpha = rz / pq.
template <typename T, int BS>
__global__ __launch_bounds__(256) void k_cg_spmv(const int64_t* __restrict__ slice_base,
                                                 const int32_t* __restrict__ slice_width,
                                                 const int32_t* __restrict__ col, const T* __restrict__ val,
                                                 const T* __restrict__ p, T* __restrict__ q, int64_t nb,
                                                 int64_t n_slices, CgState* st, double* partials,
                                                 unsigned* counter) {
  if (st->done) return;
  const int64_t slice = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const int64_t r = slice * 64 + lane;
  T acc[BS];
  if (slice < n_slices) {
    sell_row<T, BS>(slice_base, slice_width, col, val, p, slice, lane, acc);
  } else {
#pragma unroll
    for (int a = 0; a < BS; a++) acc[a] = 0;
  }
  T red[1] = {0};
  if (r < nb) {
#pragma unroll
    for (int a = 0; a < BS; a++) {
      q[BS * r + a] = acc[a];
      red[0] += acc[a] * p[BS * r + a];
    }
  }
  double tot[1];
  if (grid_sum<T, 1, 256>(red, partials, counter, tot) && threadIdx.x == 0) {
    st->pq = tot[0];
    if (!(tot[0] != 0.0) || !isfinite(tot[0]))
      st->done = CG_BREAKDOWN;
    else
      st->alpha = st->rz / tot[0];
  }
}

// x += alpha p; r -= alpha q; grid r.z and r.r with z = dinv r.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(T* __restrict__ x, const T* __restrict__ p, T* __restrict__ r,
                                                   const T* __restrict__ q, const T* __restrict__ dinv, int64_t n,
                                                   CgState* st, double* partials, unsigned* counter) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  T red[2] = {0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    red[0] += ri * zi;
    red[1] += ri * ri;
  }
  double tot[2];
  if (grid_sum<T, 2, 256>(red, partials, counter, tot) && threadIdx.x == 0) {
    st->iter += 1;
    st->rr = tot[1];
    if (!isfinite(tot[0]) || !isfinite(tot[1])) {
      st->done = CG_BREAKDOWN;
    } else if (tot[1] <= st->tol2) {
      st->done = CG_CONVERGED;
    } else if (st->iter >= st->maxit) {
      st->done = CG_MAXIT;
    } else {
      st->beta = tot[0] / st->rz;
      st->rz = tot[0];
    }
  }
}

// p = dinv r + beta p
template <typename T>
__global__ __launch_bounds__(256) void k_cg_direction(T* __restrict__ p, const T* __restrict__ r,
                                                      const T* __restrict__ dinv, int64_t n, const CgState* st) {
  if (st->done) return;
  const T beta = (T)st->beta;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (dinv ? dinv[i] * r[i] : r[i]) + beta * p[i];
}

// Start: p = dinv r, grid r.z, r.r -> rz, rr; tol2 = tol^2 * bb (bb given),
// converged already if rr <= tol2.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_start(T* __restrict__ p, const T* __restrict__ r,
                                                  const T* __restrict__ dinv, int64_t n, CgState* st, double tol,
                                                  double bb, int maxit, double* partials, unsigned* counter) {
  T red[2] = {0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = r[i];
    const T zi = dinv ? dinv[i] * ri : ri;
    p[i] = zi;
    red[0] += ri * zi;
    red[1] += ri * ri;
  }
  double tot[2];
  if (grid_sum<T, 2, 256>(red, partials, counter, tot) && threadIdx.x == 0) {
    st->rz = tot[0];
    st->rz0 = tot[0];
    st->rr = tot[1];
    st->bb = bb;
    st->tol2 = tol * tol * bb;
    st->iter = 0;
    st->maxit = maxit;
    st->alpha = st->beta = 0.0;
    if (!isfinite(tot[0]) || !isfinite(tot[1]))
      st->done = CG_BREAKDOWN;
    else if (tot[1] <= st->tol2)
      st->done = CG_CONVERGED;
    else if (maxit <= 0)
      st->done = CG_MAXIT;
    else
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
template <typename T>
struct PcgEngine {
  mfs_mat A = nullptr;
  int64_t n = 0;
  T *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
  const T* dinv = nullptr;
  const T* val = nullptr;
  CgState* st = nullptr;
  ReduceScratch red;
  unsigned g_spmv = 0, g_vec = 0;
  int chunk = 16;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;

  int init(mfs_mat A_, bool precond, int chunk_) {
    A = A_;
    n = A->n();
    chunk = chunk_;
    const bool f64 = std::is_same<T, double>::value;
    if (!f64 && !A->val32_valid) MFS_TRY(mat_refresh_derived(A));
    val = f64 ? (const T*)A->d_val64 : (const T*)A->d_val32;
    dinv = precond ? (f64 ? (const T*)A->d_dinv64 : (const T*)A->d_dinv32) : nullptr;
    MFS_TRY(dmalloc((void**)&x, sizeof(T) * n));
    MFS_TRY(dmalloc((void**)&r, sizeof(T) * n));
    MFS_TRY(dmalloc((void**)&p, sizeof(T) * n));
    MFS_TRY(dmalloc((void**)&q, sizeof(T) * n));
    MFS_TRY(dmalloc((void**)&st, sizeof(CgState)));
    MFS_HIP(hipMemset(st, 0, sizeof(CgState)));
    g_spmv = (unsigned)spmv_grid(A);
    g_vec = grid_for(n, 4 * std::max(A->ctx->cu_count, 64));
    MFS_TRY(reduce_scratch_alloc(red, 2 * (int64_t)std::max(g_spmv, g_vec) + 64));
    return capture();
  }

  int launch_iteration(hipStream_t s) {
    const int64_t ns = A->sell.n_slices;
    if (A->bs == 1)
      k_cg_spmv<T, 1><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, val, p, q, A->nb, ns, st,
                                             red.partials, red.counter);
    else
      k_cg_spmv<T, 3><<<g_spmv, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, val, p, q, A->nb, ns, st,
                                             red.partials, red.counter);
    k_cg_update<T><<<g_vec, 256, 0, s>>>(x, p, r, q, dinv, n, st, red.partials, red.counter);
    k_cg_direction<T><<<g_vec, 256, 0, s>>>(p, r, dinv, n, st);
    return MFS_OK;
  }

  int capture() {
    hipStream_t s = A->ctx->stream;
    MFS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < chunk; k++) launch_iteration(s);
    MFS_HIP(hipStreamEndCapture(s, &graph));
    MFS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return MFS_OK;
  }

  // Start from r (already holding the initial residual) and x; bb = ||b||^2
  // of the system being solved.  Runs until converged / maxit / breakdown.
  int run(double tol, double bb, int maxit, CgState* host_state) {
    hipStream_t s = A->ctx->stream;
    k_cg_start<T><<<g_vec, 256, 0, s>>>(p, r, dinv, n, st, tol, bb, maxit, red.partials, red.counter);
    MFS_HIP(hipGetLastError());
    CgState* hs = (CgState*)A->ctx->pinned;
    MFS_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    while (hs->done == CG_RUN) {
      MFS_HIP(hipGraphLaunch(exec, s));
      MFS_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
      MFS_HIP(hipStreamSynchronize(s));
    }
    *host_state = *hs;
    return MFS_OK;
  }

  void destroy() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    dfree(x);
    dfree(r);
    dfree(p);
    dfree(q);
    dfree(st);
    reduce_scratch_free(red);
  }
};

// ---------------------------------------------------------------------------
// GMRES(m) pieces (precision T): multi-dot h_k = v_k . w for k < kcount,
// w -= sum_k h_k v_k, norms.  V is column-blocked: V + k*n.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_multidot(const T* __restrict__ V, const T* __restrict__ w, int64_t n,
                                                  int kcount, double* __restrict__ partials,
                                                  unsigned* __restrict__ counter, double* __restrict__ h) {
  __shared__ double lds[32][4];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < kcount; k++) {
    T s = 0;
    const T* v = V + (int64_t)k * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
      s += v[i] * w[i];
    s = wave_sum(s);
    if (lane == 0) lds[k][wave] = (double)s;
  }
  __syncthreads();
  if (threadIdx.x < kcount)
    partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
        lds[threadIdx.x][0] + lds[threadIdx.x][1] + lds[threadIdx.x][2] + lds[threadIdx.x][3];
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int k = wave; k < kcount; k += 4) {
    double s = 0;
    for (unsigned b = lane; b < gridDim.x; b += 64)
      s += __hip_atomic_load(&partials[(size_t)k * gridDim.x + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = wave_sum(s);
    if (lane == 0) h[k] = s;
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// w -= sum_k h_k v_k (h in device memory, double); accumulate h_acc += h;
// then grid ||w||^2 into out[0].
template <typename T>
__global__ __launch_bounds__(256) void k_orth_update(const T* __restrict__ V, T* __restrict__ w, int64_t n,
                                                     int kcount, const double* __restrict__ h,
                                                     double* partials, unsigned* counter, double* out) {
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
  double tot[1];
  if (grid_sum<T, 1, 256>(red, partials, counter, tot) && threadIdx.x == 0) out[0] = tot[0];
}

template <typename T>
__global__ void k_scale_into(const T* __restrict__ w, T* __restrict__ v, const double* __restrict__ nrm, int64_t n) {
  const T inv = (T)(1.0 / sqrt(nrm[0]));
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
  void* V = nullptr;   // (m+1) * n in the inner precision
  void* w = nullptr;
  void* z = nullptr;
  double* dh = nullptr;  // device h / y (64 doubles)
  ReduceScratch redg;
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
  dfree(S->x64);
  dfree(S->b64);
  dfree(S->r64);
  dfree(S->dout);
  dfree(S->V);
  dfree(S->w);
  dfree(S->z);
  dfree(S->dh);
  reduce_scratch_free(S->red64);
  reduce_scratch_free(S->redg);
  delete S;
  mat_release(A);
  return MFS_OK;
}

extern "C" int mfs_solver_create(mfs_mat A, const mfs_solve_desc* desc, mfs_solver* out) {
  MFS_CHECK_ARG(A && desc && out, "mfs_solver_create: bad argument");
  MFS_CHECK_ARG(desc->method == MFS_PCG_JACOBI || desc->method == MFS_CG || desc->method == MFS_GMRES,
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
  if (desc->method != MFS_GMRES) {
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
    tryit(dmalloc((void**)&S->dh, 8 * 64));
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
  std::vector<double> H((m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0);
  g[0] = beta;
  // v0 = rhs / beta
  double* hp = (double*)A->ctx->pinned;
  hp[0] = beta * beta;
  MFS_HIP(hipMemcpyAsync(S->dh + 40, hp, 8, hipMemcpyHostToDevice, s));
  k_scale_into<T><<<gv, 256, 0, s>>>(rhs, V, S->dh + 40, n);
  int j = 0;
  double est = beta;
  for (; j < m && *steps_done < maxsteps; j++) {
    k_dscale<T><<<gv, 256, 0, s>>>(V + (int64_t)j * n, dinv, z, n);
    MFS_TRY(spmv_launch(A, prec, z, w, s));
    // CGS2: two passes of h = V^T w, w -= V h
    for (int pass = 0; pass < 2; pass++) {
      k_multidot<T><<<gv, 256, 0, s>>>(V, w, n, j + 1, S->redg.partials, S->redg.counter, S->dh + pass * 32);
      k_orth_update<T><<<gv, 256, 0, s>>>(V, w, n, j + 1, S->dh + pass * 32, S->redg.partials, S->redg.counter,
                                          S->dh + 62 + pass * 0);
    }
    MFS_HIP(hipGetLastError());
    MFS_HIP(hipMemcpyAsync(hp, S->dh, 64 * sizeof(double), hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    const double hn = std::sqrt(std::max(hp[62], 0.0));
    for (int i = 0; i <= j; i++) H[i * m + j] = hp[i] + hp[32 + i];
    H[(j + 1) * m + j] = hn;
    if (hn > 0) k_scale_into<T><<<gv, 256, 0, s>>>(w, V + (int64_t)(j + 1) * n, S->dh + 62, n);
    for (int i = 0; i < j; i++) {
      double a = H[i * m + j], c = H[(i + 1) * m + j];
      H[i * m + j] = cs[i] * a + sn[i] * c;
      H[(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
    }
    double a = H[j * m + j], c = H[(j + 1) * m + j];
    double d = std::sqrt(a * a + c * c);
    if (!(d > 0) || !std::isfinite(d)) return MFS_ERR_NONFINITE;
    cs[j] = a / d;
    sn[j] = c / d;
    H[j * m + j] = d;
    H[(j + 1) * m + j] = 0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    (*steps_done)++;
    est = std::fabs(g[j + 1]);
    if (est <= tol_abs || hn == 0) {
      j++;
      break;
    }
  }
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
  if (d.method != MFS_GMRES && d.precision == MFS_PREC_F64) {
    // plain FP64 (P)CG on b - A x0: r = r64 already; x correction starts at x0
    PcgEngine<double>& E = S->cg64;
    MFS_HIP(hipMemcpyAsync(E.r, S->r64, 8 * n, hipMemcpyDeviceToDevice, s));
    MFS_HIP(hipMemcpyAsync(E.x, S->x64, 8 * n, hipMemcpyDeviceToDevice, s));
    CgState hs;
    MFS_TRY(E.run(d.tol, bb, d.maxit, &hs));
    iters = hs.iter;
    outer = 1;
    MFS_HIP(hipMemcpyAsync(S->x64, E.x, 8 * n, hipMemcpyDeviceToDevice, s));
    rep->recursive_relres = std::sqrt(hs.rr / bb);
    MFS_TRY(true_residual(S, &rr, &bb));
    rel = std::sqrt(rr / bb);
    if (hs.done == CG_BREAKDOWN) status = MFS_ERR_NONFINITE;
    else if (rel > d.tol) status = MFS_ERR_NOT_CONVERGED;
  } else if (d.method != MFS_GMRES) {
    // FP64 outer refinement around FP32 (P)CG
    PcgEngine<float>& E = S->cg32;
    while (rel > d.tol && iters < d.maxit) {
      const double rn = std::sqrt(rr);
      k_demote_scaled<<<gv, 256, 0, s>>>(S->r64, 1.0 / rn, E.r, E.x, n);
      // inner tolerance: no tighter than needed to reach tol in this step
      const double itol = std::min(0.5, std::max(d.inner_tol, 0.5 * d.tol / rel));
      CgState hs;
      MFS_TRY(E.run(itol, 1.0, d.maxit - iters, &hs));
      iters += hs.iter;
      outer++;
      k_promote_add<<<gv, 256, 0, s>>>(S->x64, E.x, nullptr, rn, n);
      const double prev = rel;
      MFS_TRY(true_residual(S, &rr, &bb));
      rel = std::sqrt(rr / bb);
      rep->recursive_relres = std::sqrt(hs.rr) * prev;
      if (!std::isfinite(rel) || hs.done == CG_BREAKDOWN) {
        status = MFS_ERR_NONFINITE;
        break;
      }
      if (rel > 0.99 * prev && hs.iter <= 1) {  // no progress possible
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

extern "C" int mfs_solve(mfs_mat A, const mfs_solve_desc* desc, mfs_vec b, mfs_vec x, mfs_solve_report* report) {
  mfs_solver S;
  MFS_TRY(mfs_solver_create(A, desc, &S));
  int st = mfs_solver_run(S, b, x, report);
  mfs_solver_destroy(S);
  return st;
}

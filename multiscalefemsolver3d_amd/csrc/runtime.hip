// Context, device vectors, sparse matrices (SELL-64 layout), SpMV and vector
// kernels of the FEM solve path for gfx950.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "mfs_device.hpp"

namespace mfs {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

int dmalloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    *p = nullptr;
    return e == hipErrorOutOfMemory ? MFS_ERR_OOM : MFS_ERR_HIP;
  }
  return MFS_OK;
}
void dfree(void* p) {
  if (p) (void)hipFree(p);
}

int reduce_scratch_alloc(ReduceScratch& r, int64_t capacity) {
  r.capacity = capacity;
  MFS_TRY(dmalloc((void**)&r.partials, sizeof(double) * capacity));
  MFS_TRY(dmalloc((void**)&r.counter, 64));
  MFS_HIP(hipMemset(r.counter, 0, 64));
  return MFS_OK;
}
void reduce_scratch_free(ReduceScratch& r) {
  dfree(r.partials);
  dfree(r.counter);
  r.partials = nullptr;
  r.counter = nullptr;
}

}  // namespace mfs

using namespace mfs;

extern "C" const char* mfs_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* mfs_version(void) { return "mfs 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Fixed scratch of mfs_solve_noalloc: one file-scope device array, held by
// every device the library runs on until the process ends.  Host code never
// uses the array's name as an address: mfs_ctx_create resolves this device's
// copy with hipGetSymbolAddress once per device and the context carries that
// pointer.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device `device` must be current
int resolve_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("mfs_ctx_create: device index beyond the scratch table");
    return MFS_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return MFS_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return MFS_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
extern "C" int mfs_ctx_create(int device, mfs_ctx* out) {
  MFS_CHECK_ARG(out, "mfs_ctx_create: out is NULL");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    set_error("mfs_ctx_create: no HIP device visible");
    return MFS_ERR_NO_DEVICE;
  }
  MFS_CHECK_ARG(device >= 0 && device < count, "mfs_ctx_create: device index out of range");
  MFS_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MFS_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("mfs_ctx_create: device is ") + prop.gcnArchName + ", this build targets gfx950");
    return MFS_ERR_NO_DEVICE;
  }
  mfs_ctx c = new mfs_ctx_s();
  c->device = device;
  c->cu_count = prop.multiProcessorCount;
  std::snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return MFS_ERR_HIP;
  }
  c->pinned_bytes = 1 << 16;
  e = hipHostMalloc(&c->pinned, c->pinned_bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
    return MFS_ERR_OOM;
  }
  const int rs = resolve_scratch(device, &c->noalloc_scratch);
  if (rs != MFS_OK) {
    (void)hipHostFree(c->pinned);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rs;
  }
  c->noalloc_scratch_bytes = kNoAllocScratchBytes;
  *out = c;
  return MFS_OK;
}

namespace mfs {
void ctx_retain(mfs_ctx c) {
  if (c) c->refs.fetch_add(1);
}
void ctx_release(mfs_ctx c) {
  if (!c || c->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamDestroy(c->stream);
  (void)hipHostFree(c->pinned);
  delete c;
}
}  // namespace mfs

// Drops the caller's reference; the stream and pinned buffer are released
// once every object created on the context has been destroyed too.
extern "C" int mfs_ctx_destroy(mfs_ctx ctx) {
  if (!ctx) return MFS_OK;
  ctx_release(ctx);
  return MFS_OK;
}

extern "C" int mfs_ctx_sync(mfs_ctx ctx) {
  MFS_CHECK_ARG(ctx, "mfs_ctx_sync: NULL context");
  MFS_HIP(hipStreamSynchronize(ctx->stream));
  return MFS_OK;
}

extern "C" int mfs_ctx_info(mfs_ctx ctx, char* name, int name_len, int* cu_count) {
  MFS_CHECK_ARG(ctx, "mfs_ctx_info: NULL context");
  if (name && name_len > 0) {
    std::strncpy(name, ctx->name, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = ctx->cu_count;
  return MFS_OK;
}

// ---------------------------------------------------------------------------
// Vectors
// ---------------------------------------------------------------------------
__global__ void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
__global__ void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}
template <typename T>
__global__ void k_fill(T* __restrict__ a, T v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

static inline unsigned grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

static int vec_check(mfs_vec v, const char* who) {
  if (!v || !v->ctx || !v->d) {
    set_error(std::string(who) + ": invalid vector");
    return MFS_ERR_ARG;
  }
  return MFS_OK;
}

extern "C" int mfs_vec_create(mfs_ctx ctx, int64_t n, int prec, mfs_vec* out) {
  MFS_CHECK_ARG(ctx && out && n >= 0 && (prec == MFS_F64 || prec == MFS_F32), "mfs_vec_create: bad argument");
  MFS_HIP(hipSetDevice(ctx->device));
  mfs_vec v = new mfs_vec_s();
  v->ctx = ctx;
  v->n = n;
  v->prec = prec;
  int s = dmalloc(&v->d, (size_t)n * (prec == MFS_F64 ? 8 : 4));
  if (s != MFS_OK) {
    delete v;
    return s;
  }
  ctx_retain(ctx);
  MFS_HIP(hipMemsetAsync(v->d, 0, (size_t)n * (prec == MFS_F64 ? 8 : 4), ctx->stream));
  *out = v;
  return MFS_OK;
}

extern "C" int mfs_vec_destroy(mfs_vec v) {
  if (!v) return MFS_OK;
  if (v->ctx) (void)hipStreamSynchronize(v->ctx->stream);
  dfree(v->d);
  ctx_release(v->ctx);
  delete v;
  return MFS_OK;
}

extern "C" int mfs_vec_size(mfs_vec v, int64_t* n, int* prec) {
  MFS_TRY(vec_check(v, "mfs_vec_size"));
  if (n) *n = v->n;
  if (prec) *prec = v->prec;
  return MFS_OK;
}

extern "C" int mfs_vec_upload(mfs_vec v, const double* host) {
  MFS_TRY(vec_check(v, "mfs_vec_upload"));
  MFS_CHECK_ARG(host || v->n == 0, "mfs_vec_upload: NULL host pointer");
  hipStream_t s = v->ctx->stream;
  if (v->prec == MFS_F64) {
    MFS_HIP(hipMemcpyAsync(v->d, host, v->n * 8, hipMemcpyHostToDevice, s));
  } else {
    double* tmp;
    MFS_TRY(dmalloc((void**)&tmp, v->n * 8));
    MFS_HIP(hipMemcpyAsync(tmp, host, v->n * 8, hipMemcpyHostToDevice, s));
    k_f64_to_f32<<<grid_for(v->n), 256, 0, s>>>(tmp, (float*)v->d, v->n);
    MFS_HIP(hipStreamSynchronize(s));
    dfree(tmp);
  }
  MFS_HIP(hipStreamSynchronize(s));
  return MFS_OK;
}

extern "C" int mfs_vec_download(mfs_vec v, double* host) {
  MFS_TRY(vec_check(v, "mfs_vec_download"));
  MFS_CHECK_ARG(host || v->n == 0, "mfs_vec_download: NULL host pointer");
  hipStream_t s = v->ctx->stream;
  if (v->prec == MFS_F64) {
    MFS_HIP(hipMemcpyAsync(host, v->d, v->n * 8, hipMemcpyDeviceToHost, s));
  } else {
    double* tmp;
    MFS_TRY(dmalloc((void**)&tmp, v->n * 8));
    k_f32_to_f64<<<grid_for(v->n), 256, 0, s>>>((const float*)v->d, tmp, v->n);
    MFS_HIP(hipMemcpyAsync(host, tmp, v->n * 8, hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    dfree(tmp);
  }
  MFS_HIP(hipStreamSynchronize(s));
  return MFS_OK;
}

extern "C" int mfs_vec_fill(mfs_vec v, double value) {
  MFS_TRY(vec_check(v, "mfs_vec_fill"));
  hipStream_t s = v->ctx->stream;
  if (v->prec == MFS_F64)
    k_fill<double><<<grid_for(v->n), 256, 0, s>>>((double*)v->d, value, v->n);
  else
    k_fill<float><<<grid_for(v->n), 256, 0, s>>>((float*)v->d, (float)value, v->n);
  MFS_HIP(hipGetLastError());
  return MFS_OK;
}

extern "C" int mfs_vec_copy(mfs_vec dst, mfs_vec src) {
  MFS_TRY(vec_check(dst, "mfs_vec_copy"));
  MFS_TRY(vec_check(src, "mfs_vec_copy"));
  MFS_CHECK_ARG(dst->n == src->n, "mfs_vec_copy: size mismatch");
  hipStream_t s = dst->ctx->stream;
  if (dst->prec == src->prec) {
    MFS_HIP(hipMemcpyAsync(dst->d, src->d, src->n * (src->prec == MFS_F64 ? 8 : 4), hipMemcpyDeviceToDevice, s));
  } else if (src->prec == MFS_F64) {
    k_f64_to_f32<<<grid_for(src->n), 256, 0, s>>>((const double*)src->d, (float*)dst->d, src->n);
  } else {
    k_f32_to_f64<<<grid_for(src->n), 256, 0, s>>>((const float*)src->d, (double*)dst->d, src->n);
  }
  MFS_HIP(hipGetLastError());
  return MFS_OK;
}

// ---------------------------------------------------------------------------
// Matrices
// ---------------------------------------------------------------------------
// Scatter CSR-ordered values (bs*bs per slot, row-major block) into SELL.
template <int BS>
__global__ void k_csr_to_sell(const double* __restrict__ v, const int64_t* __restrict__ pos, int64_t nnzb, int C,
                              double* __restrict__ sell) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnzb; k += (int64_t)gridDim.x * blockDim.x) {
    int64_t p = pos[k];
    int64_t base = (p / C) * C * (BS * BS) + (p % C);  // entry p = slicebase + j*C + lane
#pragma unroll
    for (int e = 0; e < BS * BS; e++) sell[base + (int64_t)e * C] = v[(int64_t)BS * BS * k + e];
  }
}
template <int BS>
__global__ void k_sell_to_csr(const double* __restrict__ sell, const int64_t* __restrict__ pos, int64_t nnzb, int C,
                              double* __restrict__ v) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnzb; k += (int64_t)gridDim.x * blockDim.x) {
    int64_t p = pos[k];
    int64_t base = (p / C) * C * (BS * BS) + (p % C);
#pragma unroll
    for (int e = 0; e < BS * BS; e++) v[(int64_t)BS * BS * k + e] = sell[base + (int64_t)e * C];
  }
}
// FP32 copy of the values + inverse diagonal (FP64 and FP32).
template <int BS>
__global__ void k_derive(const double* __restrict__ v64, float* __restrict__ v32, int64_t nval,
                         const int64_t* __restrict__ diag_pos, int64_t nb, int C, double* __restrict__ dinv64,
                         float* __restrict__ dinv32) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nval; i += stride) v32[i] = (float)v64[i];
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nb; r += stride) {
    int64_t p = diag_pos[r];
    int64_t base = (p / C) * C * (BS * BS) + (p % C);
#pragma unroll
    for (int a = 0; a < BS; a++) {
      double d = v64[base + (int64_t)(a * BS + a) * C];
      double inv = d != 0.0 ? 1.0 / d : 1.0;
      dinv64[BS * r + a] = inv;
      dinv32[BS * r + a] = (float)inv;
    }
  }
}

namespace mfs {

int mat_create_from_pattern(mfs_ctx ctx, int bs, Pattern&& P, mfs_mat* out) {
  MFS_CHECK_ARG(P.nnz() < (int64_t)1 << 31, "matrix: more than 2^31 block nonzeros");
  MFS_HIP(hipSetDevice(ctx->device));
  mfs_mat A = new mfs_mat_s();
  A->ctx = ctx;
  ctx_retain(ctx);
  A->bs = bs;
  A->nb = P.n;
  A->nnzb = P.nnz();
  A->pat = std::move(P);
  build_sell(A->pat, 64, A->sell);
  A->sell_entries = A->sell.slice_base[A->sell.n_slices];
  std::vector<int64_t> diag_pos(A->nb);
  for (int64_t r = 0; r < A->nb; r++) {
    diag_pos[r] = -1;
    for (int64_t k = A->pat.rowptr[r]; k < A->pat.rowptr[r + 1]; k++)
      if (A->pat.col[k] == r) diag_pos[r] = A->sell.pos[k];
    if (diag_pos[r] < 0) {
      mat_release(A);
      set_error("matrix: row " + std::to_string(r) + " has no diagonal entry (CSSD rows always hold one)");
      return MFS_ERR_ARG;
    }
  }
  const int64_t nval = (int64_t)bs * bs * A->sell_entries;
  int s = MFS_OK;
  auto alloc = [&](void** p, size_t b) {
    if (s == MFS_OK) s = dmalloc(p, b);
  };
  alloc((void**)&A->d_slice_base, sizeof(int64_t) * (A->sell.n_slices + 1));
  alloc((void**)&A->d_slice_width, sizeof(int32_t) * A->sell.n_slices);
  alloc((void**)&A->d_col, sizeof(int32_t) * A->sell_entries);
  alloc((void**)&A->d_pos, sizeof(int64_t) * A->nnzb);
  alloc((void**)&A->d_diag_pos, sizeof(int64_t) * A->nb);
  alloc((void**)&A->d_val64, sizeof(double) * nval);
  alloc((void**)&A->d_val32, sizeof(float) * nval);
  alloc((void**)&A->d_dinv64, sizeof(double) * bs * A->nb);
  alloc((void**)&A->d_dinv32, sizeof(float) * bs * A->nb);
  if (s != MFS_OK) {
    mfs_mat_destroy(A);
    return s;
  }
  hipStream_t st = ctx->stream;
  MFS_HIP(hipMemcpyAsync(A->d_slice_base, A->sell.slice_base.data(), sizeof(int64_t) * (A->sell.n_slices + 1),
                         hipMemcpyHostToDevice, st));
  MFS_HIP(hipMemcpyAsync(A->d_slice_width, A->sell.slice_width.data(), sizeof(int32_t) * A->sell.n_slices,
                         hipMemcpyHostToDevice, st));
  MFS_HIP(hipMemcpyAsync(A->d_col, A->sell.col.data(), sizeof(int32_t) * A->sell_entries, hipMemcpyHostToDevice, st));
  MFS_HIP(hipMemcpyAsync(A->d_pos, A->sell.pos.data(), sizeof(int64_t) * A->nnzb, hipMemcpyHostToDevice, st));
  MFS_HIP(hipMemcpyAsync(A->d_diag_pos, diag_pos.data(), sizeof(int64_t) * A->nb, hipMemcpyHostToDevice, st));
  MFS_HIP(hipMemsetAsync(A->d_val64, 0, sizeof(double) * nval, st));
  // contiguous slice ranges per XCD in the SELL kernels (xcd_block) while
  // the FP32 values fit in the 256 MB Infinity Cache: the operand gathers of
  // neighbouring rows then hit one L2 (64^3 Laplace PCG 5 % faster).  A
  // larger matrix streams from HBM and keeps the dispatch order (128^3
  // elastic SpMV 3 % faster).  MFS_XCD_MAP=0/1 forces either (A/B runs).
  A->xcd_map = (double)A->sell_entries * A->bs * A->bs * sizeof(float) <= 256.0 * 1024 * 1024;
  if (const char* e = std::getenv("MFS_XCD_MAP")) A->xcd_map = e[0] == '1';
  // 16-bit row-relative columns when every offset fits (structured meshes
  // in natural order: offsets up to (n+1)^2 + n + 2 for n^3 cells)
  if (!std::getenv("MFS_NO_COL16")) {
    std::vector<int16_t> c16(A->sell_entries);
    bool fits = true;
    for (int64_t q = 0; q < A->sell.n_slices && fits; q++)
      for (int32_t j = 0; j < A->sell.slice_width[q] && fits; j++)
        for (int l = 0; l < A->sell.C; l++) {
          const int64_t at = A->sell.slice_base[q] + (int64_t)j * A->sell.C + l;
          const int64_t d = (int64_t)A->sell.col[at] - (q * A->sell.C + l);
          if (d < -32768 || d > 32767) {
            fits = false;
            break;
          }
          c16[at] = (int16_t)d;
        }
    if (fits && dmalloc((void**)&A->d_col16, sizeof(int16_t) * A->sell_entries) == MFS_OK) {
      MFS_HIP(hipMemcpyAsync(A->d_col16, c16.data(), sizeof(int16_t) * A->sell_entries, hipMemcpyHostToDevice, st));
    }
    MFS_HIP(hipStreamSynchronize(st));
  }
  MFS_HIP(hipStreamSynchronize(st));
  // the host SELL column copy is not needed after upload
  std::vector<int32_t>().swap(A->sell.col);
  *out = A;
  return MFS_OK;
}

int mat_refresh_derived(mfs_mat A) {
  hipStream_t s = A->ctx->stream;
  const int64_t nval = (int64_t)A->bs * A->bs * A->sell_entries;
  unsigned g = grid_for(std::max<int64_t>(nval, A->nb));
  if (A->bs == 1)
    k_derive<1><<<g, 256, 0, s>>>(A->d_val64, A->d_val32, nval, A->d_diag_pos, A->nb, A->sell.C, A->d_dinv64, A->d_dinv32);
  else
    k_derive<3><<<g, 256, 0, s>>>(A->d_val64, A->d_val32, nval, A->d_diag_pos, A->nb, A->sell.C, A->d_dinv64, A->d_dinv32);
  MFS_HIP(hipGetLastError());
  A->val32_valid = true;
  return MFS_OK;
}

double spmv_bytes(mfs_mat A, int prec) {
  const double sv = prec == MFS_F64 ? 8.0 : 4.0;
  return (double)A->nnz() * sv + (double)A->nnzb * 4.0 + (double)(A->nb + 1) * 4.0 + 2.0 * (double)A->n() * sv;
}

}  // namespace mfs

static int mat_create_common(mfs_ctx ctx, int bs, int64_t nb, const int32_t* rowptr, const int32_t* col,
                             const double* val, mfs_mat* out) {
  MFS_CHECK_ARG(ctx && out && nb > 0 && rowptr && col && val, "mfs_mat_create: bad argument");
  *out = nullptr;
  MFS_CHECK_ARG(rowptr[0] == 0, "mfs_mat_create: rowptr[0] != 0");
  for (int64_t r = 0; r < nb; r++) {
    MFS_CHECK_ARG(rowptr[r + 1] > rowptr[r], "mfs_mat_create: empty row");
    for (int32_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
      MFS_CHECK_ARG(col[k] >= 0 && col[k] < nb, "mfs_mat_create: column out of range");
      MFS_CHECK_ARG(k == rowptr[r] || col[k] > col[k - 1], "mfs_mat_create: columns not strictly ascending");
    }
  }
  Pattern P;
  P.n = nb;
  P.rowptr.assign(rowptr, rowptr + nb + 1);
  P.col.assign(col, col + rowptr[nb]);
  mfs_mat A;
  MFS_TRY(mat_create_from_pattern(ctx, bs, std::move(P), &A));
  int s = mfs_mat_update_values(A, val);
  if (s != MFS_OK) {
    mfs_mat_destroy(A);
    return s;
  }
  *out = A;
  return MFS_OK;
}

extern "C" int mfs_mat_create_csr(mfs_ctx ctx, int64_t n, const int32_t* rowptr, const int32_t* col,
                                  const double* val, mfs_mat* out) {
  return mat_create_common(ctx, 1, n, rowptr, col, val, out);
}

extern "C" int mfs_mat_create_bsr3(mfs_ctx ctx, int64_t nb, const int32_t* browptr, const int32_t* bcol,
                                   const double* val, mfs_mat* out) {
  return mat_create_common(ctx, 3, nb, browptr, bcol, val, out);
}

extern "C" int mfs_mat_update_values(mfs_mat A, const double* val) {
  MFS_CHECK_ARG(A && val, "mfs_mat_update_values: bad argument");
  hipStream_t s = A->ctx->stream;
  const int64_t nv = (int64_t)A->bs * A->bs * A->nnzb;
  double* tmp;
  MFS_TRY(dmalloc((void**)&tmp, sizeof(double) * nv));
  MFS_HIP(hipMemcpyAsync(tmp, val, sizeof(double) * nv, hipMemcpyHostToDevice, s));
  if (A->bs == 1)
    k_csr_to_sell<1><<<grid_for(A->nnzb), 256, 0, s>>>(tmp, A->d_pos, A->nnzb, A->sell.C, A->d_val64);
  else
    k_csr_to_sell<3><<<grid_for(A->nnzb), 256, 0, s>>>(tmp, A->d_pos, A->nnzb, A->sell.C, A->d_val64);
  MFS_HIP(hipGetLastError());
  MFS_TRY(mat_refresh_derived(A));
  MFS_HIP(hipStreamSynchronize(s));
  dfree(tmp);
  return MFS_OK;
}

namespace mfs {
void mat_retain(mfs_mat A) {
  if (A) A->refs.fetch_add(1);
}
void mat_release(mfs_mat A) {
  if (!A || A->refs.fetch_sub(1) != 1) return;
  if (A->ctx) (void)hipStreamSynchronize(A->ctx->stream);
  dfree(A->d_slice_base);
  dfree(A->d_slice_width);
  dfree(A->d_col);
  dfree(A->d_col16);
  dfree(A->d_pos);
  dfree(A->d_diag_pos);
  dfree(A->d_val64);
  dfree(A->d_val32);
  dfree(A->d_dinv64);
  dfree(A->d_dinv32);
  mfs_ctx c = A->ctx;
  delete A;
  ctx_release(c);
}
}  // namespace mfs

// Drops the caller's reference (solvers built on A keep it alive).
extern "C" int mfs_mat_destroy(mfs_mat A) {
  if (!A) return MFS_OK;
  mat_release(A);
  return MFS_OK;
}

extern "C" int mfs_mat_info(mfs_mat A, int64_t* n, int64_t* nnz, int* format, int64_t* nb, int64_t* nnzb) {
  MFS_CHECK_ARG(A, "mfs_mat_info: NULL matrix");
  if (n) *n = A->n();
  if (nnz) *nnz = A->nnz();
  if (format) *format = A->bs == 1 ? MFS_FMT_CSR : MFS_FMT_BSR3;
  if (nb) *nb = A->nb;
  if (nnzb) *nnzb = A->nnzb;
  return MFS_OK;
}

extern "C" int mfs_mat_export_csr(mfs_mat A, int32_t* rowptr, int32_t* col, double* val) {
  MFS_CHECK_ARG(A, "mfs_mat_export_csr: NULL matrix");
  const int bs = A->bs;
  const int64_t nv = (int64_t)bs * bs * A->nnzb;
  std::vector<double> bv;
  if (val) {
    bv.resize(nv);
    double* tmp;
    MFS_TRY(dmalloc((void**)&tmp, sizeof(double) * nv));
    hipStream_t s = A->ctx->stream;
    if (bs == 1)
      k_sell_to_csr<1><<<grid_for(A->nnzb), 256, 0, s>>>(A->d_val64, A->d_pos, A->nnzb, A->sell.C, tmp);
    else
      k_sell_to_csr<3><<<grid_for(A->nnzb), 256, 0, s>>>(A->d_val64, A->d_pos, A->nnzb, A->sell.C, tmp);
    MFS_HIP(hipMemcpyAsync(bv.data(), tmp, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    dfree(tmp);
  }
  // Scalar expansion: row 3i+a lists 3c+b for every block column c (ascending)
  // and b = 0..2, which is exactly down ++ diag ++ up of Math.h:1963-2054.
  const Pattern& P = A->pat;
  int64_t p = 0;
  if (rowptr) rowptr[0] = 0;
  for (int64_t i = 0; i < A->nb; i++)
    for (int a = 0; a < bs; a++) {
      for (int64_t k = P.rowptr[i]; k < P.rowptr[i + 1]; k++)
        for (int b = 0; b < bs; b++) {
          if (col) col[p] = (int32_t)(bs * P.col[k] + b);
          if (val) val[p] = bv[(size_t)bs * bs * k + a * bs + b];
          p++;
        }
      if (rowptr) rowptr[bs * i + a + 1] = (int32_t)p;
    }
  return MFS_OK;
}

// ---------------------------------------------------------------------------
// SpMV, SELL-64, one thread per block row.  MODE 0: y = A x.
// MODE 1: y = A x and grid dot x.y (CG denominator p.Ap).
// MODE 2: y = b - A x and grid ||y||^2, ||b||^2 (true residual).
// ---------------------------------------------------------------------------
template <typename T, int BS, int MODE, typename CI>
__global__ __launch_bounds__(256) void k_sell_spmv(const int64_t* __restrict__ slice_base,
                                                   const int32_t* __restrict__ slice_width,
                                                   const CI* __restrict__ col, const T* __restrict__ val,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   const T* __restrict__ b, int64_t nb, int64_t n_slices,
                                                   double* partials, unsigned* counter, double* out, int xcd) {
  constexpr int C = 64;
  const int64_t slice = (int64_t)xcd_block(xcd) * 4 + threadIdx.x / C;
  const int lane = threadIdx.x & (C - 1);
  T acc[BS];
  const int64_t r = slice * C + lane;
  if (slice < n_slices) {
    sell_row<T, BS, CI>(slice_base, slice_width, col, val, x, slice, lane, nb, acc);
  } else {
#pragma unroll
    for (int a = 0; a < BS; a++) acc[a] = 0;
  }
  T red[2] = {0, 0};
  if (r < nb) {
#pragma unroll
    for (int a = 0; a < BS; a++) {
      const int64_t i = BS * r + a;
      if constexpr (MODE == 0) {
        y[i] = acc[a];
      } else if constexpr (MODE == 1) {
        y[i] = acc[a];
        red[0] += acc[a] * x[i];
      } else {
        const T bi = b[i];
        const T ri = bi - acc[a];
        y[i] = ri;
        red[0] += ri * ri;
        red[1] += bi * bi;
      }
    }
  }
  if constexpr (MODE != 0) write_block_partials<T, 2>(red, partials);
}

// Block k adds up value k's G partials part[k*G + b] in a fixed order and
// writes out[k] (grid = number of values, so the values are summed in
// parallel).
__global__ __launch_bounds__(256) void k_finish_sums(const double* __restrict__ part, unsigned G,
                                                     double* __restrict__ out) {
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k = blockIdx.x;
  double s8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (unsigned b0 = threadIdx.x; b0 < G; b0 += 8 * blockDim.x) {
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const unsigned b = b0 + u * blockDim.x;
      if (b < G) s8[u] += part[(size_t)k * G + b];
    }
  }
  double s = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
  s = wave_sum(s);
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[k] = (red[0] + red[1]) + (red[2] + red[3]);
}

namespace mfs {
template <typename T, int BS, int MODE>
static void sell_spmv_launch(mfs_mat A, const T* val, const void* x, void* y, const void* b, double* partials,
                             unsigned* counter, double* out, unsigned grid, hipStream_t s) {
  const int64_t ns = A->sell.n_slices;
  if (A->d_col16)
    k_sell_spmv<T, BS, MODE, int16_t><<<grid, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col16, val,
                                                           (const T*)x, (T*)y, (const T*)b, A->nb, ns, partials,
                                                           counter, out, A->xcd_map);
  else
    k_sell_spmv<T, BS, MODE, int32_t><<<grid, 256, 0, s>>>(A->d_slice_base, A->d_slice_width, A->d_col, val,
                                                           (const T*)x, (T*)y, (const T*)b, A->nb, ns, partials,
                                                           counter, out, A->xcd_map);
}

template <int MODE>
static int spmv_mode_launch(mfs_mat A, int prec, const void* x, void* y, const void* b, double* partials,
                            unsigned* counter, double* out, hipStream_t s) {
  const int64_t ns = A->sell.n_slices;
  const unsigned grid = (unsigned)((ns + 3) / 4);
  if (prec == MFS_F64) {
    if (A->bs == 1)
      sell_spmv_launch<double, 1, MODE>(A, A->d_val64, x, y, b, partials, counter, out, grid, s);
    else
      sell_spmv_launch<double, 3, MODE>(A, A->d_val64, x, y, b, partials, counter, out, grid, s);
  } else {
    if (!A->val32_valid) MFS_TRY(mat_refresh_derived(A));
    if (A->bs == 1)
      sell_spmv_launch<float, 1, MODE>(A, A->d_val32, x, y, b, partials, counter, out, grid, s);
    else
      sell_spmv_launch<float, 3, MODE>(A, A->d_val32, x, y, b, partials, counter, out, grid, s);
  }
  MFS_HIP(hipGetLastError());
  if (MODE != 0) k_finish_sums<<<2, 256, 0, s>>>(partials, grid, out);
  return MFS_OK;
}

int spmv_launch(mfs_mat A, int prec, const void* x, void* y, hipStream_t s) {
  return spmv_mode_launch<0>(A, prec, x, y, nullptr, nullptr, nullptr, nullptr, s);
}
int spmv_dot_launch(mfs_mat A, int prec, const void* x, void* y, double* partials, unsigned* counter, double* out,
                    hipStream_t s) {
  return spmv_mode_launch<1>(A, prec, x, y, nullptr, partials, counter, out, s);
}
int spmv_residual_launch(mfs_mat A, int prec, const void* x, void* r, const void* b, double* partials,
                         unsigned* counter, double* out, hipStream_t s) {
  return spmv_mode_launch<2>(A, prec, x, r, b, partials, counter, out, s);
}
int64_t spmv_grid(mfs_mat A) { return (A->sell.n_slices + 3) / 4; }
}  // namespace mfs

extern "C" int mfs_spmv(mfs_mat A, int prec, mfs_vec x, mfs_vec y) {
  MFS_CHECK_ARG(A && x && y && (prec == MFS_F64 || prec == MFS_F32), "mfs_spmv: bad argument");
  MFS_CHECK_ARG(x->n == A->n() && y->n == A->n(), "mfs_spmv: vector size mismatch");
  MFS_CHECK_ARG(x->prec == prec && y->prec == prec, "mfs_spmv: vector precision must match prec");
  MFS_CHECK_ARG(x->d != y->d, "mfs_spmv: x and y must differ");
  return spmv_launch(A, prec, x->d, y->d, A->ctx->stream);
}

extern "C" int mfs_spmv_time(mfs_mat A, int prec, mfs_vec x, mfs_vec y, int reps, double* avg_ms) {
  MFS_CHECK_ARG(reps > 0 && avg_ms, "mfs_spmv_time: bad argument");
  MFS_TRY(mfs_spmv(A, prec, x, y));  // warm-up, also validates
  hipStream_t s = A->ctx->stream;
  hipEvent_t e0, e1;
  MFS_HIP(hipEventCreate(&e0));
  MFS_HIP(hipEventCreate(&e1));
  MFS_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) MFS_TRY(spmv_launch(A, prec, x->d, y->d, s));
  MFS_HIP(hipEventRecord(e1, s));
  MFS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = ms / reps;
  return MFS_OK;
}

// ---------------------------------------------------------------------------
// dot / axpby
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_dot(const T* __restrict__ x, const T* __restrict__ y, int64_t n,
                                             double* partials, unsigned* counter, double* out) {
  T v[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[0] += x[i] * y[i];
  write_block_partials<T, 1>(v, partials);
}

template <typename T>
__global__ void k_axpby(T a, const T* __restrict__ x, T b, T* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

extern "C" int mfs_dot(mfs_vec x, mfs_vec y, double* result) {
  MFS_TRY(vec_check(x, "mfs_dot"));
  MFS_TRY(vec_check(y, "mfs_dot"));
  MFS_CHECK_ARG(result && x->n == y->n && x->prec == y->prec, "mfs_dot: size/precision mismatch");
  hipStream_t s = x->ctx->stream;
  const unsigned g = std::min<unsigned>(grid_for(x->n), 1024);
  ReduceScratch rs;
  MFS_TRY(reduce_scratch_alloc(rs, 1024 + 8));
  double* out = rs.partials + 1024;
  if (x->prec == MFS_F64)
    k_dot<double><<<g, 256, 0, s>>>((const double*)x->d, (const double*)y->d, x->n, rs.partials, rs.counter, out);
  else
    k_dot<float><<<g, 256, 0, s>>>((const float*)x->d, (const float*)y->d, x->n, rs.partials, rs.counter, out);
  k_finish_sums<<<1, 256, 0, s>>>(rs.partials, g, out);
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipMemcpyAsync(result, out, sizeof(double), hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  reduce_scratch_free(rs);
  return MFS_OK;
}

extern "C" int mfs_axpby(double a, mfs_vec x, double b, mfs_vec y) {
  MFS_TRY(vec_check(x, "mfs_axpby"));
  MFS_TRY(vec_check(y, "mfs_axpby"));
  MFS_CHECK_ARG(x->n == y->n && x->prec == y->prec, "mfs_axpby: size/precision mismatch");
  hipStream_t s = x->ctx->stream;
  if (x->prec == MFS_F64)
    k_axpby<double><<<grid_for(x->n), 256, 0, s>>>(a, (const double*)x->d, b, (double*)y->d, x->n);
  else
    k_axpby<float><<<grid_for(x->n), 256, 0, s>>>((float)a, (const float*)x->d, (float)b, (float*)y->d, x->n);
  MFS_HIP(hipGetLastError());
  return MFS_OK;
}

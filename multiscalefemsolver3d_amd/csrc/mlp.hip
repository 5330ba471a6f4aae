// Learned initial guess: a ReLU MLP trained and evaluated with hand-written
// BF16 MFMA GEMMs (FP32 accumulation) on gfx950.  No reference counterpart:
// the network predicts each MsFEM cell's local bases from its modulus field;
// the zero-guess solve stays the fallback (the batch solver rejects a guess
// whose residual exceeds the zero guess's).
//
// GEMM: C[M][N] = sum_k A(m,k) B(k,n), A/B bf16 with arbitrary strides,
// 128x128 output tile per 256-thread workgroup, 4 waves of 64x64, each a 4x4
// grid of v_mfma_f32_16x16x32_bf16, K staged through LDS in steps of 32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mfs_device.hpp"

using namespace mfs;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_RELU = 2 };

// A 128 x 32 (rows x k) bf16 operand tile is moved global -> registers ->
// LDS dst[row][k] (row stride 40 halves) as 512 chunks of 8 halves, two per
// thread.  The operand element (row, k) is at src[row*s_row + k*s_k].  When k
// is contiguous (s_k == 1) a chunk is 8 consecutive k of one row, else 8
// consecutive rows of one k; a chunk that is in bounds and 16-byte aligned
// moves with one 16-byte load, otherwise element by element (zero padded).
struct TileRegs {
  uint4 v[2];
};

__device__ __forceinline__ uint4 pack8(const uint16_t (&e)[8]) {
  uint4 r;
  r.x = (uint32_t)e[0] | ((uint32_t)e[1] << 16);
  r.y = (uint32_t)e[2] | ((uint32_t)e[3] << 16);
  r.z = (uint32_t)e[4] | ((uint32_t)e[5] << 16);
  r.w = (uint32_t)e[6] | ((uint32_t)e[7] << 16);
  return r;
}

__device__ __forceinline__ void tile_load(TileRegs& t, const uint16_t* __restrict__ src, int64_t s_row, int64_t s_k,
                                          int row0, int k0, int R, int K) {
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int c = threadIdx.x + q * 256;
    if (s_k == 1) {
      const int row = c >> 2, kk = (c & 3) * 8;
      const int gr = row0 + row, gk = k0 + kk;
      const uint16_t* ptr = src + gr * s_row + gk;
      if (gr < R && gk + 8 <= K && ((uintptr_t)ptr & 15) == 0) {
        t.v[q] = *reinterpret_cast<const uint4*>(ptr);
      } else {
        uint16_t e8[8];
#pragma unroll
        for (int e = 0; e < 8; e++) e8[e] = (gr < R && gk + e < K) ? ptr[e] : (uint16_t)0;
        t.v[q] = pack8(e8);
      }
    } else {
      const int kk = c >> 4, row = (c & 15) * 8;
      const int gr = row0 + row, gk = k0 + kk;
      const uint16_t* ptr = src + gr * s_row + gk * s_k;
      if (gk < K && gr + 8 <= R && s_row == 1 && ((uintptr_t)ptr & 15) == 0) {
        t.v[q] = *reinterpret_cast<const uint4*>(ptr);
      } else {
        uint16_t e8[8];
#pragma unroll
        for (int e = 0; e < 8; e++) e8[e] = (gk < K && gr + e < R) ? src[(gr + e) * s_row + gk * s_k] : (uint16_t)0;
        t.v[q] = pack8(e8);
      }
    }
  }
}

__device__ __forceinline__ void tile_store(uint16_t (*dst)[40], const TileRegs& t, bool kcontig) {
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int c = threadIdx.x + q * 256;
    if (kcontig) {
      *reinterpret_cast<uint4*>(&dst[c >> 2][(c & 3) * 8]) = t.v[q];
    } else {
      const int kk = c >> 4, row = (c & 15) * 8;
      const uint32_t w[4] = {t.v[q].x, t.v[q].y, t.v[q].z, t.v[q].w};
#pragma unroll
      for (int e = 0; e < 8; e++) dst[row + e][kk] = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
    }
  }
}

// C[M][N] = sum_k A(m,k) B(k,n): 128x128 tile per 256-thread workgroup, each
// wave a 64x64 quadrant as 4x4 v_mfma_f32_16x16x32_bf16 tiles.  The next K
// step's operands are loaded into registers while the MFMAs run on the
// current LDS buffer, then stored to the other buffer: one barrier per K
// step.  Split-K: blockIdx.z covers k in [z*Kc, min(K, (z+1)*Kc)); with
// ws != nullptr the raw partial tile goes to ws[z][M][N] (summed by
// k_splitk_reduce in split order, so results do not depend on scheduling),
// else the epilogue writes C.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, int Kc, const uint16_t* __restrict__ A,
                                                   int64_t sam, int64_t sak, const uint16_t* __restrict__ B,
                                                   int64_t sbk, int64_t sbn, float* __restrict__ C, int64_t ldc,
                                                   const float* __restrict__ bias, uint16_t* __restrict__ Cbf,
                                                   float* __restrict__ ws) {
  __shared__ __align__(16) uint16_t As[2][128][40];
  __shared__ __align__(16) uint16_t Bs[2][128][40];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * 128, n0 = blockIdx.x * 128;
  const bool a_kc = sak == 1, b_kc = sbk == 1;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0, 0, 0, 0};
  const int kb = blockIdx.z * Kc, ke = min(K, kb + Kc);
  TileRegs ra, rb;
  if (kb < ke) {
    tile_load(ra, A, sam, sak, m0, kb, M, ke);
    tile_load(rb, B, sbn, sbk, n0, kb, N, ke);
    tile_store(As[0], ra, a_kc);
    tile_store(Bs[0], rb, b_kc);
  }
  __syncthreads();
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += 32) {
    const bool more = k0 + 32 < ke;
    if (more) {
      tile_load(ra, A, sam, sak, m0, k0 + 32, M, ke);
      tile_load(rb, B, sbn, sbk, n0, k0 + 32, N, ke);
    }
    bf16x8 af[4], bfr[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
      af[i] = *reinterpret_cast<const bf16x8*>(&As[buf][wm * 64 + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
    for (int j = 0; j < 4; j++)
      bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[buf][wn * 64 + j * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    if (more) {
      tile_store(As[buf ^ 1], ra, a_kc);
      tile_store(Bs[buf ^ 1], rb, b_kc);
    }
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + e;
        const int col = n0 + wn * 64 + j * 16 + (lane & 15);
        if (row >= M || col >= N) continue;
        float v = acc[i][j][e];
        if (ws) {
          ws[((int64_t)blockIdx.z * M + row) * N + col] = v;
          continue;
        }
        if (EPI != EPI_NONE) v += bias[col];
        if (EPI == EPI_BIAS_RELU) v = v > 0.f ? v : 0.f;
        C[(int64_t)row * ldc + col] = v;
        if (Cbf) Cbf[(int64_t)row * ldc + col] = f2bf(v);
      }
}

// Large GEMMs with both operands contiguous in k (the forward pass X W^T and
// the benchmark): 256x256 output tile per 512-thread workgroup, each of the
// 8 waves a 128x64 block as 4x2 v_mfma_f32_32x32x16_bf16 tiles, K staged
// through LDS in steps of 64 with two buffers (four MFMA sub-steps per
// barrier, so the fragment reads of one sub-step overlap the MFMAs of the
// previous one).  A wave reads (128 + 64) x 16 operand values from LDS per
// 128 x 64 x 16 of MFMA work, a third more arithmetic per LDS byte than the
// 64x64 quadrants of k_gemm_bf16, and two waves per SIMD overlap one's LDS
// reads with the other's MFMAs.  Interior tiles load without bounds checks;
// elsewhere operand rows beyond M / N and k beyond the split's end are read
// as zeros.  Split-K and epilogue as k_gemm_bf16.  Measured alternatives at
// 8192^3: steps of 32 (884 TFLOP/s), three LDS stages (879), fragments read
// one sub-step ahead (+0.5 %), 4 waves of 128x128 with AGPR accumulators
// (836; 940 with the next sub-step's reads interleaved between the MFMAs),
// s_setprio around the MFMAs (-3 %), waves 4-7 one sub-step behind waves
// 0-3 so that the two waves of a SIMD alternate reads and MFMAs (+1 %).  The main loop is LDS-bound: per step
// the 8 waves read 192 KiB and write 64 KiB of LDS, as many cycles at 128 B
// per clock as the step's MFMAs at peak, and the barrier keeps the waves of
// all SIMDs in the same phase.
constexpr int kBigThreads = 512;
constexpr int kBigK = 64;  // k values per LDS stage (2 stages of 2 x 256 x 72 bf16 = 144 KB)
// KS k values per LDS stage; rows padded to KS + 8 so that the 16-byte
// fragment reads of 8 consecutive rows fall in distinct banks
template <int KS, int NT = kBigThreads>
__device__ __forceinline__ void tile256_load(uint4 (&v)[KS * 32 / NT], const uint16_t* __restrict__ src,
                                             int64_t s_row, int row0, int k0, int R, int K) {
  constexpr int CPR = KS / 8;  // 16-byte chunks per row
#pragma unroll
  for (int q = 0; q < KS * 32 / NT; q++) {
    const int c = threadIdx.x + q * NT;
    const int row = c / CPR, kk = (c % CPR) * 8;
    const int gr = row0 + row, gk = k0 + kk;
    const uint16_t* ptr = src + (int64_t)gr * s_row + gk;
    if (gr < R && gk + 8 <= K && ((uintptr_t)ptr & 15) == 0) {
      v[q] = *reinterpret_cast<const uint4*>(ptr);
    } else {
      uint16_t e8[8];
#pragma unroll
      for (int e = 0; e < 8; e++) e8[e] = (gr < R && gk + e < K) ? ptr[e] : (uint16_t)0;
      v[q] = pack8(e8);
    }
  }
}
// Interior tiles (all 256 rows in range, the whole K step inside the split,
// 16-byte aligned rows): no bounds checks
template <int KS, int NT = kBigThreads>
__device__ __forceinline__ void tile256_load_interior(uint4 (&v)[KS * 32 / NT], const uint16_t* __restrict__ src,
                                                      int64_t s_row, int row0, int k0) {
  constexpr int CPR = KS / 8;
#pragma unroll
  for (int q = 0; q < KS * 32 / NT; q++) {
    const int c = threadIdx.x + q * NT;
    v[q] = *reinterpret_cast<const uint4*>(src + (int64_t)(row0 + c / CPR) * s_row + k0 + (c % CPR) * 8);
  }
}
template <int KS, int NT = kBigThreads>
__device__ __forceinline__ void tile256_store(uint16_t (*dst)[KS + 8], const uint4 (&v)[KS * 32 / NT]) {
  constexpr int CPR = KS / 8;
#pragma unroll
  for (int q = 0; q < KS * 32 / NT; q++) {
    const int c = threadIdx.x + q * NT;
    *reinterpret_cast<uint4*>(&dst[c / CPR][(c % CPR) * 8]) = v[q];
  }
}

template <int EPI, int KS>
__global__ __launch_bounds__(kBigThreads) void k_gemm_bf16_big(int M, int N, int K, int Kc,
                                                               const uint16_t* __restrict__ A, int64_t sam,
                                                               const uint16_t* __restrict__ B, int64_t sbn,
                                                               float* __restrict__ C, int64_t ldc,
                                                               const float* __restrict__ bias,
                                                               uint16_t* __restrict__ Cbf, float* __restrict__ ws) {
  __shared__ __align__(16) uint16_t As[2][256][KS + 8];
  __shared__ __align__(16) uint16_t Bs[2][256][KS + 8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 2, wn = w & 3;  // 2 x 4 waves of 128 x 64
  const int m0 = blockIdx.y * 256, n0 = blockIdx.x * 256;
  f32x16 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[i][j][e] = 0.f;
  const int kb = blockIdx.z * Kc, ke = min(K, kb + Kc);
  uint4 ra[KS / 16], rb[KS / 16];
  // workgroup-uniform: every row of both operand tiles in range and aligned
  const bool interior = m0 + 256 <= M && n0 + 256 <= N && (sam & 7) == 0 && (sbn & 7) == 0 &&
                        (kb & 7) == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
  if (kb < ke) {
    if (interior && kb + KS <= ke) {
      tile256_load_interior<KS>(ra, A, sam, m0, kb);
      tile256_load_interior<KS>(rb, B, sbn, n0, kb);
    } else {
      tile256_load<KS>(ra, A, sam, m0, kb, M, ke);
      tile256_load<KS>(rb, B, sbn, n0, kb, N, ke);
    }
    tile256_store<KS>(As[0], ra);
    tile256_store<KS>(Bs[0], rb);
  }
  __syncthreads();
  int buf = 0;
  const int fr = lane & 31, fk = 8 * (lane >> 5);
  for (int k0 = kb; k0 < ke; k0 += KS) {
    const bool more = k0 + KS < ke;
    if (more) {
      if (interior && k0 + 2 * KS <= ke) {
        tile256_load_interior<KS>(ra, A, sam, m0, k0 + KS);
        tile256_load_interior<KS>(rb, B, sbn, n0, k0 + KS);
      } else {
        tile256_load<KS>(ra, A, sam, m0, k0 + KS, M, ke);
        tile256_load<KS>(rb, B, sbn, n0, k0 + KS, N, ke);
      }
    }
#pragma unroll
    for (int ks = 0; ks < KS; ks += 16) {
      bf16x8 af[4], bfr[2];
#pragma unroll
      for (int i = 0; i < 4; i++)
        af[i] = *reinterpret_cast<const bf16x8*>(&As[buf][wm * 128 + i * 32 + fr][ks + fk]);
#pragma unroll
      for (int j = 0; j < 2; j++)
        bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[buf][wn * 64 + j * 32 + fr][ks + fk]);
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    if (more) {
      tile256_store<KS>(As[buf ^ 1], ra);
      tile256_store<KS>(Bs[buf ^ 1], rb);
    }
    __syncthreads();
    buf ^= 1;
  }
  // 32x32 accumulator layout: value e of a lane is row 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int row = m0 + wm * 128 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        const int col = n0 + wn * 64 + j * 32 + (lane & 31);
        if (row >= M || col >= N) continue;
        float v = acc[i][j][e];
        if (ws) {
          ws[((int64_t)blockIdx.z * M + row) * N + col] = v;
          continue;
        }
        if (EPI != EPI_NONE) v += bias[col];
        if (EPI == EPI_BIAS_RELU) v = v > 0.f ? v : 0.f;
        C[(int64_t)row * ldc + col] = v;
        if (Cbf) Cbf[(int64_t)row * ldc + col] = f2bf(v);
      }
}

// C = sum over the splits of ws[z] (in split order), then the epilogue
template <int EPI>
__global__ void k_splitk_reduce(const float* __restrict__ ws, int splits, int M, int N, float* __restrict__ C,
                                int64_t ldc, const float* __restrict__ bias, uint16_t* __restrict__ Cbf) {
  const int64_t MN = (int64_t)M * N;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < MN; q += (int64_t)gridDim.x * blockDim.x) {
    float v = 0.f;
    for (int z = 0; z < splits; z++) v += ws[z * MN + q];
    const int64_t row = q / N, col = q % N;
    if (EPI != EPI_NONE) v += bias[col];
    if (EPI == EPI_BIAS_RELU) v = v > 0.f ? v : 0.f;
    C[row * ldc + col] = v;
    if (Cbf) Cbf[row * ldc + col] = f2bf(v);
  }
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}
// gather rows idx[0..B) of X [*, D] into dst bf16 [B][D]
__global__ void k_gather_rows_bf16(const float* __restrict__ X, const int* __restrict__ idx, int B, int64_t D,
                                   uint16_t* __restrict__ dst) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < (int64_t)B * D;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = q / D, c = q % D;
    dst[q] = f2bf(X[(int64_t)idx[r] * D + c]);
  }
}
// dY = (Y - T[idx]) * scale (FP32 and bf16), partial loss sums
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             const int* __restrict__ idx, int B, int64_t D, float scale,
                                             float* __restrict__ dY, uint16_t* __restrict__ dYbf,
                                             double* __restrict__ loss_partial) {
  double s = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < (int64_t)B * D;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = q / D, c = q % D;
    const float t = idx ? T[(int64_t)idx[r] * D + c] : T[q];
    const float e = Y[q] - t;
    s += (double)e * e;
    const float g = e * scale;
    dY[q] = g;
    dYbf[q] = f2bf(g);
  }
  s = wave_sum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
// column sums: db[c] = sum_r dY[r][c]
__global__ void k_colsum(const float* __restrict__ dY, int B, int D, float* __restrict__ db) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < D; c += gridDim.x * blockDim.x) {
    float s = 0;
    for (int r = 0; r < B; r++) s += dY[(int64_t)r * D + c];
    db[c] = s;
  }
}
// dX *= (act > 0), refresh bf16 copy
__global__ void k_relu_back(float* __restrict__ dX, const float* __restrict__ act, uint16_t* __restrict__ dXbf,
                            int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = act[i] > 0.f ? dX[i] : 0.f;
    dX[i] = v;
    dXbf[i] = f2bf(v);
  }
}
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       uint16_t* __restrict__ wbf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wbf) wbf[i] = f2bf(wi);
  }
}

inline unsigned g1d(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(65536, (n + 255) / 256)); }

}  // namespace

struct mfs_mlp_s {
  mfs_ctx ctx = nullptr;
  int L = 0;
  std::vector<int> dims;
  double lr = 1e-3, b1 = 0.9, b2 = 0.999, eps = 1e-8;
  int64_t step = 0;
  std::vector<float*> W, b, mW, vW, mb, vb, gW, gb;
  std::vector<uint16_t*> Wbf;
  // activations for the current batch capacity
  int cap = 0;
  std::vector<float*> act;       // FP32 [cap][dims[l]] (l = 1..L)
  std::vector<uint16_t*> actbf;  // bf16 [cap][dims[l]] (l = 0..L-1)
  std::vector<float*> dact;      // FP32 grads [cap][dims[l]] (l = 1..L)
  std::vector<uint16_t*> dactbf;
  float* ws = nullptr;  // split-K partial tiles
  size_t ws_bytes = 0;
  int* d_idx = nullptr;
  double* d_loss = nullptr;
};

static int mlp_free(mfs_mlp M) {
  if (!M) return MFS_OK;
  if (M->ctx) (void)hipStreamSynchronize(M->ctx->stream);
  for (auto* v : {&M->W, &M->b, &M->mW, &M->vW, &M->mb, &M->vb, &M->gW, &M->gb, &M->act, &M->dact})
    for (float* p : *v) dfree(p);
  for (auto* v : {&M->Wbf, &M->actbf, &M->dactbf})
    for (uint16_t* p : *v) dfree(p);
  dfree(M->d_idx);
  dfree(M->d_loss);
  dfree(M->ws);
  mfs_ctx c = M->ctx;
  delete M;
  ctx_release(c);
  return MFS_OK;
}

extern "C" int mfs_mlp_destroy(mfs_mlp M) { return mlp_free(M); }

static size_t gemm_ws_bytes(mfs_mlp M, int B);

static int mlp_capacity(mfs_mlp M, int batch) {
  if (batch <= M->cap) return MFS_OK;
  for (auto* v : {&M->act, &M->dact})
    for (float*& p : *v) {
      dfree(p);
      p = nullptr;
    }
  for (auto* v : {&M->actbf, &M->dactbf})
    for (uint16_t*& p : *v) {
      dfree(p);
      p = nullptr;
    }
  dfree(M->d_idx);
  dfree(M->d_loss);
  const int L = M->L;
  M->act.assign(L + 1, nullptr);
  M->dact.assign(L + 1, nullptr);
  M->actbf.assign(L + 1, nullptr);
  M->dactbf.assign(L + 1, nullptr);
  for (int l = 0; l <= L; l++) {
    const size_t e = (size_t)batch * M->dims[l];
    MFS_TRY(dmalloc((void**)&M->act[l], 4 * e));
    MFS_TRY(dmalloc((void**)&M->dact[l], 4 * e));
    MFS_TRY(dmalloc((void**)&M->actbf[l], 2 * e));
    MFS_TRY(dmalloc((void**)&M->dactbf[l], 2 * e));
  }
  MFS_TRY(dmalloc((void**)&M->d_idx, sizeof(int) * batch));
  MFS_TRY(dmalloc((void**)&M->d_loss, sizeof(double) * 65536));
  const size_t wsb = gemm_ws_bytes(M, batch);
  if (wsb > M->ws_bytes) {
    dfree(M->ws);
    M->ws = nullptr;
    M->ws_bytes = 0;
    MFS_TRY(dmalloc((void**)&M->ws, wsb));
    M->ws_bytes = wsb;
  }
  M->cap = batch;
  return MFS_OK;
}

extern "C" int mfs_mlp_create(mfs_ctx ctx, const mfs_mlp_desc* d, mfs_mlp* out) {
  MFS_CHECK_ARG(ctx && d && out && d->n_layers >= 1 && d->n_layers <= 16 && d->dims, "mfs_mlp_create: bad argument");
  for (int l = 0; l <= d->n_layers; l++) MFS_CHECK_ARG(d->dims[l] > 0, "mfs_mlp_create: layer size must be > 0");
  MFS_HIP(hipSetDevice(ctx->device));
  *out = nullptr;
  mfs_mlp M = new mfs_mlp_s();
  M->ctx = ctx;
  ctx_retain(ctx);
  M->L = d->n_layers;
  M->dims.assign(d->dims, d->dims + d->n_layers + 1);
  M->lr = d->lr > 0 ? d->lr : 1e-3;
  M->b1 = d->beta1 > 0 ? d->beta1 : 0.9;
  M->b2 = d->beta2 > 0 ? d->beta2 : 0.999;
  M->eps = d->eps > 0 ? d->eps : 1e-8;
  const int L = M->L;
  for (auto* v : {&M->W, &M->b, &M->mW, &M->vW, &M->mb, &M->vb, &M->gW, &M->gb}) v->assign(L, nullptr);
  M->Wbf.assign(L, nullptr);
  hipStream_t s = ctx->stream;
  // He-uniform init from a counter-based generator; last layer scaled by 0.1
  for (int l = 0; l < L; l++) {
    const int64_t nin = M->dims[l], nout = M->dims[l + 1], nw = nin * nout;
    int st = MFS_OK;
    for (float** p : {&M->W[l], &M->mW[l], &M->vW[l], &M->gW[l]})
      if (st == MFS_OK) st = dmalloc((void**)p, 4 * nw);
    for (float** p : {&M->b[l], &M->mb[l], &M->vb[l], &M->gb[l]})
      if (st == MFS_OK) st = dmalloc((void**)p, 4 * nout);
    if (st == MFS_OK) st = dmalloc((void**)&M->Wbf[l], 2 * nw);
    if (st != MFS_OK) {
      mlp_free(M);
      return st;
    }
    std::vector<float> w(nw);
    std::vector<uint16_t> wb(nw);
    const double lim = std::sqrt(6.0 / nin) * (l == L - 1 ? 0.1 : 1.0);
    for (int64_t i = 0; i < nw; i++) {
      uint64_t z = d->seed * 0x9E3779B97F4A7C15ull + (uint64_t)l * 0xD1B54A32D192ED03ull + (uint64_t)i;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
      w[i] = (float)((2 * u - 1) * lim);
      wb[i] = f2bf(w[i]);
    }
    MFS_HIP(hipMemcpyAsync(M->W[l], w.data(), 4 * nw, hipMemcpyHostToDevice, s));
    MFS_HIP(hipMemcpyAsync(M->Wbf[l], wb.data(), 2 * nw, hipMemcpyHostToDevice, s));
    MFS_HIP(hipMemsetAsync(M->b[l], 0, 4 * nout, s));
    MFS_HIP(hipMemsetAsync(M->mW[l], 0, 4 * nw, s));
    MFS_HIP(hipMemsetAsync(M->vW[l], 0, 4 * nw, s));
    MFS_HIP(hipMemsetAsync(M->mb[l], 0, 4 * nout, s));
    MFS_HIP(hipMemsetAsync(M->vb[l], 0, 4 * nout, s));
    MFS_HIP(hipStreamSynchronize(s));
  }
  *out = M;
  return MFS_OK;
}

extern "C" int mfs_mlp_get_params(mfs_mlp M, int l, float* W, float* b) {
  MFS_CHECK_ARG(M && l >= 0 && l < M->L, "mfs_mlp_get_params: bad argument");
  const int64_t nw = (int64_t)M->dims[l] * M->dims[l + 1];
  hipStream_t s = M->ctx->stream;
  if (W) MFS_HIP(hipMemcpyAsync(W, M->W[l], 4 * nw, hipMemcpyDeviceToHost, s));
  if (b) MFS_HIP(hipMemcpyAsync(b, M->b[l], 4 * M->dims[l + 1], hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  return MFS_OK;
}

extern "C" int mfs_mlp_set_params(mfs_mlp M, int l, const float* W, const float* b) {
  MFS_CHECK_ARG(M && l >= 0 && l < M->L, "mfs_mlp_set_params: bad argument");
  const int64_t nw = (int64_t)M->dims[l] * M->dims[l + 1];
  hipStream_t s = M->ctx->stream;
  if (W) {
    MFS_HIP(hipMemcpyAsync(M->W[l], W, 4 * nw, hipMemcpyHostToDevice, s));
    k_to_bf16<<<g1d(nw), 256, 0, s>>>(M->W[l], M->Wbf[l], nw);
  }
  if (b) MFS_HIP(hipMemcpyAsync(M->b[l], b, 4 * M->dims[l + 1], hipMemcpyHostToDevice, s));
  MFS_HIP(hipStreamSynchronize(s));
  return MFS_OK;
}

// Both operands contiguous in k and at least one 256x256 tile: the
// 256x256 kernel.
static bool gemm_big(int M, int N, int64_t sak, int64_t sbk) { return sak == 1 && sbk == 1 && M >= 256 && N >= 256; }

// Split-K factor: enough workgroups to cover the device when the output has
// few output tiles (tile x tile) and K is long; every split keeps at least
// 256 of K.
static int gemm_splits(int M, int N, int K, int* Kc, int tile = 128) {
  const int64_t tiles = (int64_t)((M + tile - 1) / tile) * ((N + tile - 1) / tile);
  int sp = (int)std::min<int64_t>(std::max<int64_t>(1, 512 / tiles), std::max(1, K / 256));
  sp = std::min(sp, 64);
  int kc = (K + sp - 1) / sp;
  kc = (kc + 31) / 32 * 32;
  sp = (K + kc - 1) / kc;
  *Kc = kc;
  return sp;
}

template <int EPI>
static void gemm_epi(hipStream_t s, float* ws, size_t ws_bytes, int M, int N, int K, const uint16_t* A, int64_t sam,
                     int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn, float* C, int64_t ldc,
                     const float* bias, uint16_t* Cbf) {
  const bool big = gemm_big(M, N, sak, sbk);
  const int tile = big ? 256 : 128;
  int Kc;
  int sp = gemm_splits(M, N, K, &Kc, tile);
  if (sp > 1 && (!ws || (size_t)sp * M * N * sizeof(float) > ws_bytes)) {
    sp = 1;  // no room for the partial tiles: one pass over K
    Kc = K;
  }
  dim3 grid((N + tile - 1) / tile, (M + tile - 1) / tile, sp);
  if (big) {
    k_gemm_bf16_big<EPI, kBigK><<<grid, kBigThreads, 0, s>>>(M, N, K, Kc, A, sam, B, sbn, C, ldc, bias, Cbf,
                                                               sp == 1 ? nullptr : ws);
    if (sp > 1) {
      const int64_t mn = (int64_t)M * N;
      k_splitk_reduce<EPI><<<(unsigned)std::min<int64_t>(4096, (mn + 255) / 256), 256, 0, s>>>(ws, sp, M, N, C, ldc,
                                                                                               bias, Cbf);
    }
  } else if (sp == 1) {
    k_gemm_bf16<EPI><<<grid, 256, 0, s>>>(M, N, K, Kc, A, sam, sak, B, sbk, sbn, C, ldc, bias, Cbf, nullptr);
  } else {
    k_gemm_bf16<EPI_NONE><<<grid, 256, 0, s>>>(M, N, K, Kc, A, sam, sak, B, sbk, sbn, C, ldc, bias, Cbf, ws);
    const int64_t mn = (int64_t)M * N;
    k_splitk_reduce<EPI><<<(unsigned)std::min<int64_t>(4096, (mn + 255) / 256), 256, 0, s>>>(ws, sp, M, N, C, ldc,
                                                                                             bias, Cbf);
  }
}

static void gemm(mfs_mlp Mp, int epi, int M, int N, int K, const uint16_t* A, int64_t sam, int64_t sak,
                 const uint16_t* B, int64_t sbk, int64_t sbn, float* C, int64_t ldc, const float* bias, uint16_t* Cbf) {
  hipStream_t s = Mp->ctx->stream;
  if (epi == EPI_NONE)
    gemm_epi<EPI_NONE>(s, Mp->ws, Mp->ws_bytes, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Cbf);
  else if (epi == EPI_BIAS)
    gemm_epi<EPI_BIAS>(s, Mp->ws, Mp->ws_bytes, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Cbf);
  else
    gemm_epi<EPI_BIAS_RELU>(s, Mp->ws, Mp->ws_bytes, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Cbf);
}

// workspace for the split-K GEMMs of a batch size (forward and backward)
static size_t gemm_ws_bytes(mfs_mlp M, int B) {
  size_t need = 0;
  auto one = [&](int m, int n, int k) {
    int kc;
    const int sp = gemm_splits(m, n, k, &kc);
    if (sp > 1) need = std::max(need, (size_t)sp * m * n * sizeof(float));
  };
  for (int l = 0; l < M->L; l++) {
    const int nin = M->dims[l], nout = M->dims[l + 1];
    one(B, nout, nin);
    one(nout, nin, B);
    if (l > 0) one(B, nin, nout);
  }
  return need;
}

// forward on actbf[0] (already filled, B rows): act[L] holds the output
// out: where the last layer goes ([B][dims[L]], row stride dims[L]); the
// activation buffer when null
static void mlp_forward_dev(mfs_mlp M, int B, float* out = nullptr) {
  for (int l = 0; l < M->L; l++) {
    const int nin = M->dims[l], nout = M->dims[l + 1];
    const bool last = l == M->L - 1;
    // Y[B][out] = X[B][in] W[out][in]^T + b
    gemm(M, last ? EPI_BIAS : EPI_BIAS_RELU, B, nout, nin, M->actbf[l], nin, 1, M->Wbf[l], 1, nin,
         last && out ? out : M->act[l + 1], nout, M->b[l], last ? nullptr : M->actbf[l + 1]);
  }
}

// backward from dact[L]/dactbf[L] (filled): gradients into gW/gb
static void mlp_backward_dev(mfs_mlp M, int B) {
  hipStream_t s = M->ctx->stream;
  for (int l = M->L - 1; l >= 0; l--) {
    const int nin = M->dims[l], nout = M->dims[l + 1];
    // gW[out][in] = sum_m dY[m][out] X[m][in]
    gemm(M, EPI_NONE, nout, nin, B, M->dactbf[l + 1], 1, nout, M->actbf[l], nin, 1, M->gW[l], nin, nullptr, nullptr);
    k_colsum<<<g1d(nout), 256, 0, s>>>(M->dact[l + 1], B, nout, M->gb[l]);
    if (l > 0) {
      // dX[m][in] = sum_o dY[m][o] W[o][in], then ReLU mask
      gemm(M, EPI_NONE, B, nin, nout, M->dactbf[l + 1], nout, 1, M->Wbf[l], nin, 1, M->dact[l], nin, nullptr, nullptr);
      k_relu_back<<<g1d((int64_t)B * nin), 256, 0, s>>>(M->dact[l], M->act[l], M->dactbf[l], (int64_t)B * nin);
    }
  }
}

static void mlp_adam(mfs_mlp M) {
  hipStream_t s = M->ctx->stream;
  M->step++;
  const float c1 = (float)(1.0 - std::pow(M->b1, (double)M->step));
  const float c2 = (float)(1.0 - std::pow(M->b2, (double)M->step));
  for (int l = 0; l < M->L; l++) {
    const int64_t nw = (int64_t)M->dims[l] * M->dims[l + 1];
    k_adam<<<g1d(nw), 256, 0, s>>>(M->W[l], M->mW[l], M->vW[l], M->gW[l], nw, (float)M->lr, (float)M->b1,
                                   (float)M->b2, (float)M->eps, c1, c2, M->Wbf[l]);
    k_adam<<<g1d(M->dims[l + 1]), 256, 0, s>>>(M->b[l], M->mb[l], M->vb[l], M->gb[l], M->dims[l + 1], (float)M->lr,
                                                (float)M->b1, (float)M->b2, (float)M->eps, c1, c2, nullptr);
  }
}

extern "C" int mfs_mlp_forward(mfs_mlp M, int batch, const float* X, float* Y) {
  MFS_CHECK_ARG(M && batch > 0 && X && Y, "mfs_mlp_forward: bad argument");
  MFS_TRY(mlp_capacity(M, batch));
  hipStream_t s = M->ctx->stream;
  const int64_t n0 = (int64_t)batch * M->dims[0];
  MFS_HIP(hipMemcpyAsync(M->act[0], X, 4 * n0, hipMemcpyHostToDevice, s));
  k_to_bf16<<<g1d(n0), 256, 0, s>>>(M->act[0], M->actbf[0], n0);
  mlp_forward_dev(M, batch);
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipMemcpyAsync(Y, M->act[M->L], 4 * (int64_t)batch * M->dims[M->L], hipMemcpyDeviceToHost, s));
  MFS_HIP(hipStreamSynchronize(s));
  return MFS_OK;
}

static int mlp_loss_dev(mfs_mlp M, int B, const float* T_dev, const int* idx_dev, double* loss) {
  hipStream_t s = M->ctx->stream;
  const int64_t D = M->dims[M->L];
  const int64_t n = (int64_t)B * D;
  const unsigned g = std::min<unsigned>(g1d(n), 65536);
  k_mse<<<g, 256, 0, s>>>(M->act[M->L], T_dev, idx_dev, B, D, (float)(1.0 / (double)n), M->dact[M->L],
                          M->dactbf[M->L], M->d_loss);
  MFS_HIP(hipGetLastError());
  if (loss) {
    std::vector<double> lp(g);
    MFS_HIP(hipMemcpyAsync(lp.data(), M->d_loss, sizeof(double) * g, hipMemcpyDeviceToHost, s));
    MFS_HIP(hipStreamSynchronize(s));
    double t = 0;
    for (double v : lp) t += v;
    *loss = 0.5 * t / (double)n;
  }
  return MFS_OK;
}

extern "C" int mfs_mlp_loss_grad(mfs_mlp M, int batch, const float* X, const float* T, double* loss, float* const* gW,
                                 float* const* gb) {
  MFS_CHECK_ARG(M && batch > 0 && X && T && loss, "mfs_mlp_loss_grad: bad argument");
  MFS_TRY(mlp_capacity(M, batch));
  hipStream_t s = M->ctx->stream;
  const int64_t n0 = (int64_t)batch * M->dims[0], nL = (int64_t)batch * M->dims[M->L];
  float* Td;
  MFS_TRY(dmalloc((void**)&Td, 4 * nL));
  MFS_HIP(hipMemcpyAsync(M->act[0], X, 4 * n0, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(Td, T, 4 * nL, hipMemcpyHostToDevice, s));
  k_to_bf16<<<g1d(n0), 256, 0, s>>>(M->act[0], M->actbf[0], n0);
  mlp_forward_dev(M, batch);
  MFS_TRY(mlp_loss_dev(M, batch, Td, nullptr, loss));
  mlp_backward_dev(M, batch);
  MFS_HIP(hipGetLastError());
  for (int l = 0; l < M->L; l++) {
    if (gW && gW[l])
      MFS_HIP(hipMemcpyAsync(gW[l], M->gW[l], 4 * (int64_t)M->dims[l] * M->dims[l + 1], hipMemcpyDeviceToHost, s));
    if (gb && gb[l]) MFS_HIP(hipMemcpyAsync(gb[l], M->gb[l], 4 * M->dims[l + 1], hipMemcpyDeviceToHost, s));
  }
  MFS_HIP(hipStreamSynchronize(s));
  dfree(Td);
  return MFS_OK;
}

// Train on device-resident features X [n][dims0] and targets T [n][dimsL].
static int mlp_train_dev(mfs_mlp M, int64_t n, const float* X, const float* T, int batch, int epochs, uint64_t seed,
                         double* final_loss) {
  MFS_TRY(mlp_capacity(M, batch));
  hipStream_t s = M->ctx->stream;
  std::vector<int> perm(n);
  double last = 0;
  for (int ep = 0; ep < epochs; ep++) {
    for (int64_t i = 0; i < n; i++) perm[i] = (int)i;
    // Fisher-Yates with a splitmix64 stream
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)ep;
    for (int64_t i = n - 1; i > 0; i--) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t y = z;
      y = (y ^ (y >> 30)) * 0xBF58476D1CE4E5B9ull;
      y = (y ^ (y >> 27)) * 0x94D049BB133111EBull;
      y ^= y >> 31;
      std::swap(perm[i], perm[y % (uint64_t)(i + 1)]);
    }
    double ep_loss = 0;
    int nb = 0;
    for (int64_t b0 = 0; b0 + batch <= n; b0 += batch) {
      MFS_HIP(hipMemcpyAsync(M->d_idx, perm.data() + b0, sizeof(int) * batch, hipMemcpyHostToDevice, s));
      k_gather_rows_bf16<<<g1d((int64_t)batch * M->dims[0]), 256, 0, s>>>(X, M->d_idx, batch, M->dims[0],
                                                                          M->actbf[0]);
      mlp_forward_dev(M, batch);
      const bool want = (b0 + 2 * batch > n);  // read the loss of the last batch of the epoch
      double l = 0;
      MFS_TRY(mlp_loss_dev(M, batch, T, M->d_idx, want ? &l : nullptr));
      mlp_backward_dev(M, batch);
      mlp_adam(M);
      MFS_HIP(hipGetLastError());
      if (want) {
        ep_loss += l;
        nb++;
      }
      // the index upload reads perm on the host: keep the host copy stable
      MFS_HIP(hipStreamSynchronize(s));
    }
    last = nb ? ep_loss / nb : 0;
  }
  if (final_loss) *final_loss = last;
  return MFS_OK;
}

extern "C" int mfs_mlp_train(mfs_mlp M, int64_t n_samples, const float* X, const float* T, int batch, int epochs,
                             double* final_loss) {
  MFS_CHECK_ARG(M && X && T && batch > 0 && n_samples >= batch && epochs >= 0, "mfs_mlp_train: bad argument");
  float *Xd, *Td;
  const int64_t nx = n_samples * M->dims[0], nt = n_samples * M->dims[M->L];
  MFS_TRY(dmalloc((void**)&Xd, 4 * nx));
  MFS_TRY(dmalloc((void**)&Td, 4 * nt));
  hipStream_t s = M->ctx->stream;
  MFS_HIP(hipMemcpyAsync(Xd, X, 4 * nx, hipMemcpyHostToDevice, s));
  MFS_HIP(hipMemcpyAsync(Td, T, 4 * nt, hipMemcpyHostToDevice, s));
  int st = mlp_train_dev(M, n_samples, Xd, Td, batch, epochs, 12345, final_loss);
  MFS_HIP(hipStreamSynchronize(s));
  dfree(Xd);
  dfree(Td);
  return st;
}

extern "C" int mfs_mlp_train_vec(mfs_mlp M, int64_t n_samples, mfs_vec X, mfs_vec T, int batch, int epochs,
                                 uint64_t seed, double* final_loss) {
  MFS_CHECK_ARG(M && X && T && batch > 0 && n_samples >= batch && epochs >= 0, "mfs_mlp_train_vec: bad argument");
  MFS_CHECK_ARG(X->prec == MFS_F32 && T->prec == MFS_F32, "mfs_mlp_train_vec: FP32 vectors expected");
  MFS_CHECK_ARG(X->n >= n_samples * M->dims[0] && T->n >= n_samples * M->dims[M->L],
                "mfs_mlp_train_vec: vectors too short");
  return mlp_train_dev(M, n_samples, (const float*)X->d, (const float*)T->d, batch, epochs, seed, final_loss);
}

extern "C" int mfs_mlp_forward_vec(mfs_mlp M, int64_t n_samples, mfs_vec X, mfs_vec Y, int batch) {
  MFS_CHECK_ARG(M && X && Y && n_samples > 0 && batch > 0, "mfs_mlp_forward_vec: bad argument");
  MFS_CHECK_ARG(X->prec == MFS_F32 && Y->prec == MFS_F32, "mfs_mlp_forward_vec: FP32 vectors expected");
  MFS_CHECK_ARG(X->n >= n_samples * M->dims[0] && Y->n >= n_samples * M->dims[M->L],
                "mfs_mlp_forward_vec: vectors too short");
  MFS_TRY(mlp_capacity(M, batch));
  hipStream_t s = M->ctx->stream;
  const int64_t D0 = M->dims[0], DL = M->dims[M->L];
  for (int64_t b0 = 0; b0 < n_samples; b0 += batch) {
    const int B = (int)std::min<int64_t>(batch, n_samples - b0);
    k_to_bf16<<<g1d((int64_t)B * D0), 256, 0, s>>>((const float*)X->d + b0 * D0, M->actbf[0], (int64_t)B * D0);
    mlp_forward_dev(M, B, (float*)Y->d + b0 * DL);  // the last layer straight into Y
  }
  MFS_HIP(hipGetLastError());
  MFS_HIP(hipStreamSynchronize(s));
  return MFS_OK;
}

extern "C" int mfs_gemm_bf16_time(mfs_ctx ctx, int M, int N, int K, int reps, double* avg_ms, double* tflops) {
  MFS_CHECK_ARG(ctx && M > 0 && N > 0 && K > 0 && reps > 0 && avg_ms, "mfs_gemm_bf16_time: bad argument");
  MFS_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  uint16_t *A = nullptr, *B = nullptr;
  float *C = nullptr, *ws = nullptr;
  int st = dmalloc((void**)&A, 2 * (size_t)M * K);
  if (st == MFS_OK) st = dmalloc((void**)&B, 2 * (size_t)N * K);
  if (st == MFS_OK) st = dmalloc((void**)&C, 4 * (size_t)M * N);
  int kc;
  const int sp = gemm_splits(M, N, K, &kc);
  const size_t wsb = sp > 1 ? (size_t)sp * M * N * sizeof(float) : 0;
  if (st == MFS_OK && wsb) st = dmalloc((void**)&ws, wsb);
  if (st != MFS_OK) {
    dfree(A);
    dfree(B);
    dfree(C);
    return st;
  }
  // operands: X W^T layout of the forward pass (A row-major [M][K], B [N][K])
  MFS_HIP(hipMemsetAsync(A, 0x3c, 2 * (size_t)M * K, s));
  MFS_HIP(hipMemsetAsync(B, 0x3c, 2 * (size_t)N * K, s));
  gemm_epi<EPI_NONE>(s, ws, wsb, M, N, K, A, K, 1, B, 1, K, C, N, nullptr, nullptr);
  hipEvent_t e0, e1;
  MFS_HIP(hipEventCreate(&e0));
  MFS_HIP(hipEventCreate(&e1));
  MFS_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) gemm_epi<EPI_NONE>(s, ws, wsb, M, N, K, A, K, 1, B, 1, K, C, N, nullptr, nullptr);
  MFS_HIP(hipEventRecord(e1, s));
  MFS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(A);
  dfree(B);
  dfree(C);
  dfree(ws);
  *avg_ms = ms / reps;
  if (tflops) *tflops = 2.0 * M * N * (double)K / (*avg_ms * 1e-3) / 1e12;
  return MFS_OK;
}

// Eight independent multiply-add chains per thread, `iters` steps each: the
// VALU issue rate of plain (unpacked) FP32 and of FP64 arithmetic, as the
// solver kernels are compiled.  The sum is stored so nothing is removed.
template <typename T>
__global__ __launch_bounds__(256) void k_fma_rate(T* __restrict__ out, int iters, T a, T b) {
  T v[8];
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = (T)(threadIdx.x + c);
  for (int i = 0; i < iters; i++)
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = fma(v[c], a, b);
  T sum = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) sum += v[c];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

extern "C" int mfs_valu_fma_rate(mfs_ctx ctx, int dtype, double* tflops) {
  MFS_CHECK_ARG(ctx && tflops && (dtype == MFS_F32 || dtype == MFS_F64), "mfs_valu_fma_rate: bad argument");
  MFS_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  constexpr int kBlocks = 256 * 32, kThreads = 256, kIters = 8192;
  void* out = nullptr;
  int st = dmalloc(&out, (size_t)kBlocks * kThreads * 8);
  if (st != MFS_OK) return st;
  auto launch = [&]() {
    if (dtype == MFS_F32)
      k_fma_rate<float><<<kBlocks, kThreads, 0, s>>>((float*)out, kIters, 0.9999999f, 1e-7f);
    else
      k_fma_rate<double><<<kBlocks, kThreads, 0, s>>>((double*)out, kIters, 0.9999999, 1e-7);
  };
  launch();  // warm-up
  hipEvent_t e0, e1;
  MFS_HIP(hipEventCreate(&e0));
  MFS_HIP(hipEventCreate(&e1));
  constexpr int kReps = 5;
  MFS_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < kReps; r++) launch();
  MFS_HIP(hipEventRecord(e1, s));
  MFS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(out);
  const double flops = 2.0 * 8 * (double)kIters * kBlocks * kThreads;
  *tflops = flops / (ms / kReps * 1e-3) / 1e12;
  return MFS_OK;
}

extern "C" int mfs_mlp_time_step(mfs_mlp M, int batch, int reps, double* avg_ms, double* tflops) {
  MFS_CHECK_ARG(M && batch > 0 && reps > 0 && avg_ms, "mfs_mlp_time_step: bad argument");
  MFS_TRY(mlp_capacity(M, batch));
  hipStream_t s = M->ctx->stream;
  // synthetic inputs/targets already in the buffers (zeros are fine for timing)
  MFS_HIP(hipMemsetAsync(M->actbf[0], 0, 2 * (size_t)batch * M->dims[0], s));
  float* Td;
  MFS_TRY(dmalloc((void**)&Td, 4 * (size_t)batch * M->dims[M->L]));
  MFS_HIP(hipMemsetAsync(Td, 0, 4 * (size_t)batch * M->dims[M->L], s));
  hipEvent_t e0, e1;
  MFS_HIP(hipEventCreate(&e0));
  MFS_HIP(hipEventCreate(&e1));
  MFS_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) {
    mlp_forward_dev(M, batch);
    MFS_TRY(mlp_loss_dev(M, batch, Td, nullptr, nullptr));
    mlp_backward_dev(M, batch);
  }
  MFS_HIP(hipEventRecord(e1, s));
  MFS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(Td);
  *avg_ms = ms / reps;
  double flops = 0;
  for (int l = 0; l < M->L; l++) flops += 2.0 * batch * M->dims[l] * M->dims[l + 1] * (l == 0 ? 2 : 3);
  if (tflops) *tflops = flops / (*avg_ms * 1e-3) / 1e12;
  return MFS_OK;
}

// cnn.hip -- BF16 MFMA kernels of the event-triage CNN (forward, backward,
// optimiser).
//
// Every product of the network (3x3 convolutions as implicit GEMMs,
// their weight and input gradients, the two linear layers) runs through one
// GEMM kernel, C[m][n] = sum_k A(m,k) B(k,n), with BF16 operands, FP32
// accumulation on v_mfma_f32_16x16x16_bf16, and operand loaders that read
// dense strided matrices or gather im2col windows straight from NHWC
// activations (no im2col buffer):
//
//   tile 64 (m) x 64 (n) x 32 (k), 256 threads = 4 waves in 2 x 2, each wave
//   a 32 x 32 block = 2 x 2 MFMA 16x16 tiles.  A is staged as As[m][k] and B
//   as Bt[n][k] in LDS (row pitch 36 halves: 8-byte aligned lane reads, no
//   bank conflicts between the 16-lane groups), so each lane's MFMA operand
//   is one 8-byte LDS read of 4 consecutive k.
//   MFMA 16x16x16 bf16 lane layout: A row = lane%16, k = 4*(lane/16)+j;
//   B column = lane%16, same k; D column = lane%16, rows 4*(lane/16)+j.
//   Split-K: grid.z slices the reduction; slices write FP32 partials that
//   epilogue_kernel sums in slice order (deterministic), adds the bias,
//   applies ReLU or a ReLU mask and stores FP32 and/or BF16.
//
// Element-wise kernels: ReLU + 2x2 max pool (argmax and mask byte), its
// backward, per-channel bias gradients, softmax cross-entropy, Adam, NCHW ->
// NHWC input transpose.  BF16 conversion is round-to-nearest-even.
#include <hip/hip_runtime.h>

#include "cnn_internal.h"

namespace mdsm {
namespace cnn {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, PITCH = BK + 4, NT = 256;

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ------------------------------------------------------------ loaders
// Each loader fills 8 consecutive k of one operand row: fetchA(m, k, ...)
// gives A(m, k..k+7), fetchB(n, k, ...) gives B(k..k+7, n); k is a multiple
// of 8 and elements at or past kend are 0.  Index arithmetic is 32-bit
// (every operand here has fewer than 2^31 elements) and a run of 8 elements
// that is contiguous and 16-byte aligned in memory is one 16-byte load.
__device__ __forceinline__ bool aligned16(const uint16_t *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}
__device__ __forceinline__ void load8(const uint16_t *p, uint16_t v[8]) {
  const uint4 q = *reinterpret_cast<const uint4 *>(p);
  v[0] = (uint16_t)q.x; v[1] = (uint16_t)(q.x >> 16);
  v[2] = (uint16_t)q.y; v[3] = (uint16_t)(q.y >> 16);
  v[4] = (uint16_t)q.z; v[5] = (uint16_t)(q.z >> 16);
  v[6] = (uint16_t)q.w; v[7] = (uint16_t)(q.w >> 16);
}
__device__ __forceinline__ void zero8(uint16_t v[8]) {
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = 0;
}

struct Dense {  // (r, c) at p[r*sr + c*sc]
  const uint16_t *p;
  int64_t sr, sc;
  __device__ __forceinline__ void fetchA(int64_t m, int64_t k, int64_t kend, uint16_t v[8]) const {
    const uint16_t *q = p + m * sr + k * sc;
    if (sc == 1 && k + 8 <= kend && aligned16(q)) {
      load8(q, v);
      return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = k + e < kend ? q[e * sc] : (uint16_t)0;
  }
  __device__ __forceinline__ void fetchB(int64_t n, int64_t k, int64_t kend, uint16_t v[8]) const {
    const uint16_t *q = p + k * sr + n * sc;
    if (sr == 1 && k + 8 <= kend && aligned16(q)) {
      load8(q, v);
      return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = k + e < kend ? q[e * sr] : (uint16_t)0;
  }
};

// im2col view of NHWC x[B][H][W][C]: (pixel, k) with k = (ky*3 + kx)*C + c
// reads x at (h + ky - 1, w + kx - 1); flipped, at (h - ky + 1, w - kx + 1).
template <bool FLIP>
struct Im2col {
  const uint16_t *p;
  int H, W, C;
  __device__ __forceinline__ uint16_t elem(int b, int h, int w, int k) const {
    const int kk = k / C, c = k - kk * C;
    const int ky = kk / 3, kx = kk - 3 * ky;
    const int hh = FLIP ? h - ky + 1 : h + ky - 1, ww = FLIP ? w - kx + 1 : w + kx - 1;
    if (hh < 0 || hh >= H || ww < 0 || ww >= W) return 0;
    return p[(int64_t)((b * H + hh) * W + ww) * C + c];
  }
  // one pixel, 8 consecutive k: with C a multiple of 8 they share (ky, kx)
  // and are 8 consecutive channels, one aligned 16-byte load
  __device__ __forceinline__ void fetchA(int64_t m, int64_t k, int64_t kend, uint16_t v[8]) const {
    const int pix = (int)m, kk0 = (int)k;
    const int w = pix % W, bh = pix / W, h = bh % H, b = bh / H;
    if ((C & 7) == 0 && k + 8 <= kend) {
      const int kk = kk0 / C, c = kk0 - kk * C;
      const int ky = kk / 3, kx = kk - 3 * ky;
      const int hh = FLIP ? h - ky + 1 : h + ky - 1, ww = FLIP ? w - kx + 1 : w + kx - 1;
      if (hh < 0 || hh >= H || ww < 0 || ww >= W) {
        zero8(v);
      } else {
        load8(p + (int64_t)((b * H + hh) * W + ww) * C + c, v);
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = k + e < kend ? elem(b, h, w, kk0 + e) : (uint16_t)0;
  }
  // 8 consecutive pixels, one k: stepped through (w, h, b) without division
  __device__ __forceinline__ void fetchB(int64_t n, int64_t k, int64_t kend, uint16_t v[8]) const {
    int pix = (int)k;
    int w = pix % W, bh = pix / W, h = bh % H, b = bh / H;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = k + e < kend ? elem(b, h, w, (int)n) : (uint16_t)0;
      if (++w == W) {
        w = 0;
        if (++h == H) {
          h = 0;
          ++b;
        }
      }
    }
  }
};

// weights w[co][(ky*3+kx)*Ci + ci] seen as (k = (ky*3+kx)*Co + co, n = ci)
struct WFlip {
  const uint16_t *p;
  int Ci, Co;
  __device__ __forceinline__ void fetchB(int64_t n, int64_t k, int64_t kend, uint16_t v[8]) const {
    int kk = (int)k / Co, co = (int)k - kk * Co;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = k + e < kend ? p[(int64_t)co * 9 * Ci + kk * Ci + (int)n] : (uint16_t)0;
      if (++co == Co) {
        co = 0;
        ++kk;
      }
    }
  }
};

template <typename LA, typename LB>
__global__ __launch_bounds__(NT) void gemm_kernel(LA A, LB B, int64_t M, int64_t N, int64_t K,
                                                  int64_t kchunk, float *__restrict__ C,
                                                  int64_t ldc) {
  __shared__ __align__(16) uint16_t As[BM * PITCH];
  __shared__ __align__(16) uint16_t Bt[BN * PITCH];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = kbeg + kchunk < K ? kbeg + kchunk : K;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lr = t >> 2, lk = (t & 3) * 8;  // loader: row (m or n) and first k of 8
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    {
      const int64_t m = m0 + lr;
      uint16_t v[8];
      if (m < M) A.fetchA(m, k0 + lk, kend, v);
      else zero8(v);
      uint16_t *dst = As + lr * PITCH + lk;
#pragma unroll
      for (int e = 0; e < 8; ++e) dst[e] = v[e];
    }
    {
      const int64_t n = n0 + lr;
      uint16_t v[8];
      if (n < N) B.fetchB(n, k0 + lk, kend, v);
      else zero8(v);
      uint16_t *dst = Bt + lr * PITCH + lk;
#pragma unroll
      for (int e = 0; e < 8; ++e) dst[e] = v[e];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      bf16x4 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *reinterpret_cast<const bf16x4 *>(As + (wm * 32 + i * 16 + (lane & 15)) * PITCH + ks +
                                                 4 * (lane >> 4));
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *reinterpret_cast<const bf16x4 *>(Bt + (wn * 32 + j * 16 + (lane & 15)) * PITCH + ks +
                                                 4 * (lane >> 4));
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  float *Cz = C + (int64_t)blockIdx.z * M * ldc;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int64_t n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m < M && n < N) Cz[m * ldc + n] = acc[i][j][r];
      }
}

// sum of split partials [splits][M][N] + bias[n]; out_f32 = value; then
// mask (mask[m*N+n] > 0) and/or ReLU; out_bf16 = value
__global__ void epilogue_kernel(const float *__restrict__ part, int splits, int64_t M, int64_t N,
                                const float *__restrict__ bias, float *__restrict__ out_f32,
                                const float *__restrict__ mask, int relu,
                                uint16_t *__restrict__ out_bf16) {
  const int64_t MN = M * N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < MN;
       i += (int64_t)gridDim.x * blockDim.x) {
    float v = part[i];
    for (int z = 1; z < splits; ++z) v += part[(int64_t)z * MN + i];
    if (bias) v += bias[i % N];
    if (out_f32) out_f32[i] = v;
    if (mask && !(mask[i] > 0.f)) v = 0.f;
    if (relu) v = v > 0.f ? v : 0.f;
    if (out_bf16) out_bf16[i] = to_bf16(v);
  }
}

// z[B][H][W][C] + bias -> ReLU -> 2x2 max pool -> a[B][H/2][W/2][C] (BF16),
// idx = argmax (0..3, first maximum) | 4 when the maximum is positive
__global__ void relu_pool_kernel(const float *__restrict__ z, const float *__restrict__ bias,
                                 int64_t B, int H, int W, int C, uint16_t *__restrict__ a,
                                 uint8_t *__restrict__ idx) {
  const int Ho = H / 2, Wo = W / 2;
  const int64_t n = B * Ho * Wo * C;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    int64_t r = i / C;
    const int wo = (int)(r % Wo);
    r /= Wo;
    const int ho = (int)(r % Ho);
    const int64_t b = r / Ho;
    float best = -1.f;
    int arg = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float v = z[((b * H + 2 * ho + (k >> 1)) * W + 2 * wo + (k & 1)) * C + c] + bias[c];
      v = v > 0.f ? v : 0.f;
      if (v > best) {
        best = v;
        arg = k;
      }
    }
    a[i] = to_bf16(best);
    idx[i] = (uint8_t)(arg | (best > 0.f ? 4 : 0));
  }
}

// da[B][H/2][W/2][C] (FP32) -> dz[B][H][W][C] (BF16): the gradient goes to
// the argmax when it was positive, zeros elsewhere
__global__ void pool_bwd_kernel(const float *__restrict__ da, const uint8_t *__restrict__ idx,
                                int64_t B, int H, int W, int C, uint16_t *__restrict__ dz) {
  const int Ho = H / 2, Wo = W / 2;
  const int64_t n = B * Ho * Wo * C;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    int64_t r = i / C;
    const int wo = (int)(r % Wo);
    r /= Wo;
    const int ho = (int)(r % Ho);
    const int64_t b = r / Ho;
    const uint8_t q = idx[i];
    const uint16_t g = (q & 4) ? to_bf16(da[i]) : (uint16_t)0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      dz[((b * H + 2 * ho + (k >> 1)) * W + 2 * wo + (k & 1)) * C + c] = (k == (q & 3)) ? g : (uint16_t)0;
  }
}

// column sums of a BF16 matrix x[rows][C] (C <= 256): per-block partials
__global__ void colsum_kernel(const uint16_t *__restrict__ x, int64_t rows, int C,
                              int64_t rows_per_block, float *__restrict__ part) {
  __shared__ float s[NT];
  const int groups = NT / C;  // thread (g, c): rows g, g + groups, ...
  const int g = threadIdx.x / C, c = threadIdx.x - g * C;
  float acc = 0.f;
  if (g < groups) {
    const int64_t r0 = blockIdx.x * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
    for (int64_t r = r0 + g; r < r1; r += groups) acc += from_bf16(x[r * C + c]);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < C) {
    float v = 0.f;
    for (int k = 0; k < groups; ++k) v += s[k * C + threadIdx.x];
    part[(int64_t)blockIdx.x * C + threadIdx.x] = v;
  }
}

// one workgroup per channel: thread t sums parts t, t + NT, ... in order,
// then a fixed binary tree (deterministic for a given nparts)
__global__ __launch_bounds__(NT) void sum_parts_kernel(const float *__restrict__ part, int nparts,
                                                       int C, float *__restrict__ out) {
  __shared__ float s[NT];
  const int c = blockIdx.x;
  float v = 0.f;
  for (int p = threadIdx.x; p < nparts; p += NT) v += part[(int64_t)p * C + c];
  s[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int h = NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = s[0];
}

// mean softmax cross-entropy over B rows of 2 logits; dl = (p - onehot)/B (BF16)
__global__ void softmax_ce_kernel(const float *__restrict__ logits, const int32_t *__restrict__ y,
                                  int B, uint16_t *__restrict__ dl, float *__restrict__ loss) {
  __shared__ float s[NT];
  float acc = 0.f;
  for (int b = threadIdx.x; b < B; b += NT) {
    const float l0 = logits[2 * b], l1 = logits[2 * b + 1];
    const float m = fmaxf(l0, l1);
    const float e0 = expf(l0 - m), e1 = expf(l1 - m), sum = e0 + e1;
    const int lab = y[b] != 0;
    acc += -((lab ? l1 : l0) - m - logf(sum));
    dl[2 * b] = to_bf16((e0 / sum - (lab == 0 ? 1.f : 0.f)) / (float)B);
    dl[2 * b + 1] = to_bf16((e1 / sum - (lab == 1 ? 1.f : 0.f)) / (float)B);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = NT / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = s[0] / (float)B;
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, uint16_t *__restrict__ pb, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    pb[i] = to_bf16(pi);
  }
}

__global__ void to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = to_bf16(x[i]);
}

// x[B][C][H][W] -> y[B][H][W][C]
__global__ void nchw_to_nhwc_kernel(const uint16_t *__restrict__ x, int64_t B, int C, int H, int W,
                                    uint16_t *__restrict__ y) {
  const int64_t n = B * C * H * W;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    int64_t r = i / C;
    const int w = (int)(r % W);
    r /= W;
    const int h = (int)(r % H);
    const int64_t b = r / H;
    y[i] = x[((b * C + c) * H + h) * W + w];
  }
}

inline unsigned grid1d(int64_t n) {
  const int64_t g = (n + NT - 1) / NT;
  return (unsigned)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

inline int64_t kchunk_of(int64_t K, int splits) {
  return ((K + splits - 1) / splits + BK - 1) / BK * BK;
}

template <typename LA, typename LB>
int gemm(LA A, LB B, int64_t M, int64_t N, int64_t K, int splits, float *C, hipStream_t s) {
  const int64_t kchunk = kchunk_of(K, splits);
  const int z = (int)((K + kchunk - 1) / kchunk);
  const dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + BN - 1) / BN), (unsigned)z);
  hipLaunchKernelGGL((gemm_kernel<LA, LB>), grid, dim3(NT), 0, s, A, B, M, N, K, kchunk, C, N);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

}  // namespace

int splits_used(int64_t K, int splits) {
  const int64_t kchunk = kchunk_of(K, splits);
  return (int)((K + kchunk - 1) / kchunk);
}

int gemm_dense(const uint16_t *a, int64_t asr, int64_t asc, const uint16_t *b, int64_t bsr,
               int64_t bsc, int64_t M, int64_t N, int64_t K, int splits, float *C, hipStream_t s) {
  return gemm(Dense{a, asr, asc}, Dense{b, bsr, bsc}, M, N, K, splits, C, s);
}

int conv_forward(const uint16_t *x, int64_t B, int H, int W, int Ci, const uint16_t *w, int Co,
                 float *z, hipStream_t s) {
  // z[pix][co] = sum_k im2col(x)[pix][k] * w[co][k]
  return gemm(Im2col<false>{x, H, W, Ci}, Dense{w, 1, (int64_t)9 * Ci}, B * H * W, Co,
              (int64_t)9 * Ci, 1, z, s);
}

int conv_weight_grad(const uint16_t *dz, const uint16_t *x, int64_t B, int H, int W, int Ci, int Co,
                     int splits, float *part, hipStream_t s) {
  // dw[co][k] = sum_pix dz[pix][co] * im2col(x)[pix][k]
  return gemm(Dense{dz, 1, Co}, Im2col<false>{x, H, W, Ci}, Co, (int64_t)9 * Ci, B * H * W, splits,
              part, s);
}

int conv_input_grad(const uint16_t *dz, int64_t B, int H, int W, int Ci, int Co, const uint16_t *w,
                    float *dx, hipStream_t s) {
  // dx[pix][ci] = sum_{ky,kx,co} dz[pix shifted back][co] * w[co][ky,kx,ci]
  return gemm(Im2col<true>{dz, H, W, Co}, WFlip{w, Ci, Co}, B * H * W, Ci, (int64_t)9 * Co, 1, dx,
              s);
}

int epilogue(const float *part, int splits, int64_t M, int64_t N, const float *bias, float *out_f32,
             const float *mask, int relu, uint16_t *out_bf16, hipStream_t s) {
  hipLaunchKernelGGL(epilogue_kernel, dim3(grid1d(M * N)), dim3(NT), 0, s, part, splits, M, N, bias,
                     out_f32, mask, relu, out_bf16);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int relu_pool(const float *z, const float *bias, int64_t B, int H, int W, int C, uint16_t *a,
              uint8_t *idx, hipStream_t s) {
  hipLaunchKernelGGL(relu_pool_kernel, dim3(grid1d(B * (H / 2) * (W / 2) * C)), dim3(NT), 0, s, z,
                     bias, B, H, W, C, a, idx);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int pool_backward(const float *da, const uint8_t *idx, int64_t B, int H, int W, int C, uint16_t *dz,
                  hipStream_t s) {
  hipLaunchKernelGGL(pool_bwd_kernel, dim3(grid1d(B * (H / 2) * (W / 2) * C)), dim3(NT), 0, s, da,
                     idx, B, H, W, C, dz);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int column_sums(const uint16_t *x, int64_t rows, int C, float *part, int max_parts, float *out,
                hipStream_t s) {
  int64_t per = (rows + max_parts - 1) / max_parts;
  if (per < 1) per = 1;
  const int nparts = (int)((rows + per - 1) / per);
  hipLaunchKernelGGL(colsum_kernel, dim3(nparts > 0 ? nparts : 1), dim3(NT), 0, s, x, rows, C, per,
                     part);
  MDSM_HIP(hipGetLastError());
  hipLaunchKernelGGL(sum_parts_kernel, dim3(C), dim3(NT), 0, s, part, nparts > 0 ? nparts : 1, C,
                     out);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int softmax_ce(const float *logits, const int32_t *y, int B, uint16_t *dl, float *loss,
               hipStream_t s) {
  hipLaunchKernelGGL(softmax_ce_kernel, dim3(1), dim3(NT), 0, s, logits, y, B, dl, loss);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int adam(float *p, const float *g, float *m, float *v, uint16_t *pb, int64_t n, float lr, float b1,
         float b2, float eps, int step, float wd, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(grid1d(n)), dim3(NT), 0, s, p, g, m, v, pb, n, lr, b1, b2,
                     eps, c1, c2, wd);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1d(n)), dim3(NT), 0, s, x, y, n);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int nchw_to_nhwc(const uint16_t *x, int64_t B, int C, int H, int W, uint16_t *y, hipStream_t s) {
  hipLaunchKernelGGL(nchw_to_nhwc_kernel, dim3(grid1d(B * C * H * W)), dim3(NT), 0, s, x, B, C, H,
                     W, y);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

}  // namespace cnn
}  // namespace mdsm

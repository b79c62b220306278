// cnn_internal.h -- launchers of the triage-CNN kernels (cnn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace mdsm {
namespace cnn {

// number of K slices a split-K product with `splits` requested writes
int splits_used(int64_t K, int splits);
// C (+ slice * M * N) [m][n] = sum_k a[m*asr + k*asc] * b[k*bsr + n*bsc]
int gemm_dense(const uint16_t *a, int64_t asr, int64_t asc, const uint16_t *b, int64_t bsr,
               int64_t bsc, int64_t M, int64_t N, int64_t K, int splits, float *C, hipStream_t s);
int conv_forward(const uint16_t *x, int64_t B, int H, int W, int Ci, const uint16_t *w, int Co,
                 float *z, hipStream_t s);
int conv_weight_grad(const uint16_t *dz, const uint16_t *x, int64_t B, int H, int W, int Ci, int Co,
                     int splits, float *part, hipStream_t s);
int conv_input_grad(const uint16_t *dz, int64_t B, int H, int W, int Ci, int Co, const uint16_t *w,
                    float *dx, hipStream_t s);
int epilogue(const float *part, int splits, int64_t M, int64_t N, const float *bias, float *out_f32,
             const float *mask, int relu, uint16_t *out_bf16, hipStream_t s);
int relu_pool(const float *z, const float *bias, int64_t B, int H, int W, int C, uint16_t *a,
              uint8_t *idx, hipStream_t s);
int pool_backward(const float *da, const uint8_t *idx, int64_t B, int H, int W, int C, uint16_t *dz,
                  hipStream_t s);
int column_sums(const uint16_t *x, int64_t rows, int C, float *part, int max_parts, float *out,
                hipStream_t s);
int softmax_ce(const float *logits, const int32_t *y, int B, uint16_t *dl, float *loss,
               hipStream_t s);
int adam(float *p, const float *g, float *m, float *v, uint16_t *pb, int64_t n, float lr, float b1,
         float b2, float eps, int step, float wd, hipStream_t s);
int to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s);
int nchw_to_nhwc(const uint16_t *x, int64_t B, int C, int H, int W, uint16_t *y, hipStream_t s);

}  // namespace cnn
}  // namespace mdsm

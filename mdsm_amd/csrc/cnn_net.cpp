// cnn_net.cpp -- the event-triage network object: buffers, parameter layouts,
// forward / backward / Adam sequences over the kernels of cnn.hip.
//
// Canonical parameter layout (the C ABI's get/set functions): w1 b1 w2 b2
// w3 b3 fw1 fb1 fw2 fb2 with conv weights [co][ci][ky][kx] and fw1[o][i],
// i = c*(H/8)*(W/8) + y*(W/8) + x.  Device layout: conv weights
// [co][(ky*3+kx)*Ci + ci] and fw1 columns in NHWC order ((y*(W/8)+x)*64 +
// c), matching the NHWC activations the kernels use.  Master weights are
// FP32; BF16 copies feed the products; biases are added in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "cnn_internal.h"

using namespace mdsm;

namespace {

constexpr int CH[4] = {2, 16, 32, 64};
constexpr int HID = 64, NCLS = 2, BK = 32;

struct Layout {
  int64_t w[3], b[3], fw1, fb1, fw2, fb2, n;
  int fin;
  Layout(int H, int W) {
    int64_t off = 0;
    for (int k = 0; k < 3; ++k) {
      w[k] = off;
      off += (int64_t)CH[k + 1] * CH[k] * 9;
      b[k] = off;
      off += CH[k + 1];
    }
    fin = 64 * (H / 8) * (W / 8);
    fw1 = off;
    off += (int64_t)fin * HID;
    fb1 = off;
    off += HID;
    fw2 = off;
    off += HID * NCLS;
    fb2 = off;
    off += NCLS;
    n = off;
  }
};

int choose_splits(int64_t M, int64_t N, int64_t K) {
  const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
  int64_t want = std::max<int64_t>(1, 1024 / tiles);
  want = std::min<int64_t>(want, std::max<int64_t>(1, K / (4 * BK)));
  return (int)std::min<int64_t>(want, 1024);
}

}  // namespace

struct mdsm_cnn_s {
  int H = 0, W = 0, Bmax = 0;
  Layout L{8, 8};
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *pb = nullptr;
  uint16_t *x = nullptr, *a[4] = {nullptr, nullptr, nullptr, nullptr};
  uint8_t *idx[3] = {nullptr, nullptr, nullptr};
  float *z = nullptr, *z4 = nullptr, *logits = nullptr, *dA = nullptr, *part = nullptr;
  float *cpart = nullptr, *loss = nullptr;
  uint16_t *h = nullptr, *dl = nullptr, *dh = nullptr, *dZ = nullptr;
  int64_t part_floats = 0;
  int last_batch = 0;
  int step = 0;
  std::vector<void *> allocs;
};

namespace {

template <typename T>
bool dalloc(mdsm_cnn_s *n, T **p, int64_t count) {
  void *q = nullptr;
  if (hipMalloc(&q, (size_t)std::max<int64_t>(count, 1) * sizeof(T)) != hipSuccess) return false;
  n->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return true;
}

void free_net(mdsm_cnn_s *n) {
  for (void *q : n->allocs) (void)hipFree(q);
  delete n;
}

// canonical <-> device index maps (device index of canonical element i)
std::vector<int64_t> device_index(const mdsm_cnn_s *n) {
  const Layout &L = n->L;
  std::vector<int64_t> map(L.n);
  for (int64_t i = 0; i < L.n; ++i) map[i] = i;
  for (int k = 0; k < 3; ++k) {
    const int Ci = CH[k], Co = CH[k + 1];
    for (int co = 0; co < Co; ++co)
      for (int ci = 0; ci < Ci; ++ci)
        for (int t = 0; t < 9; ++t)
          map[L.w[k] + ((int64_t)co * Ci + ci) * 9 + t] = L.w[k] + (int64_t)co * 9 * Ci + t * Ci + ci;
  }
  const int H8 = n->H / 8, W8 = n->W / 8;
  for (int o = 0; o < HID; ++o)
    for (int c = 0; c < 64; ++c)
      for (int y = 0; y < H8; ++y)
        for (int x = 0; x < W8; ++x)
          map[L.fw1 + (int64_t)o * L.fin + ((int64_t)c * H8 + y) * W8 + x] =
              L.fw1 + (int64_t)o * L.fin + ((int64_t)y * W8 + x) * 64 + c;
  return map;
}

}  // namespace

extern "C" int mdsm_cnn_create(int height, int width, int max_batch, mdsm_cnn *out) {
  MDSM_REQUIRE(out && height >= 8 && width >= 8 && height % 8 == 0 && width % 8 == 0 &&
                   max_batch >= 1 && (int64_t)max_batch * height * width <= (1ll << 31),
               MDSM_ERR_ARG, "snippets must be multiples of 8 in both axes; batch >= 1");
  *out = nullptr;
  auto *n = new mdsm_cnn_s();
  n->H = height;
  n->W = width;
  n->Bmax = max_batch;
  n->L = Layout(height, width);
  const int64_t B = max_batch, HW = (int64_t)height * width;
  // split-K partials: the largest of the split products
  int64_t pf = 0;
  auto need = [&](int64_t M, int64_t N, int64_t K) {
    pf = std::max(pf, (int64_t)cnn::splits_used(K, choose_splits(M, N, K)) * M * N);
  };
  need(B, HID, n->L.fin);      // fc1 forward
  need(B, NCLS, HID);          // fc2 forward
  need(NCLS, HID, B);          // fc2 weight grad
  need(B, HID, NCLS);          // fc2 input grad
  need(HID, n->L.fin, B);      // fc1 weight grad
  for (int k = 0; k < 3; ++k)  // conv weight grads
    need(CH[k + 1], 9 * CH[k], B * (HW >> (2 * k)));
  n->part_floats = pf;
  bool ok = dalloc(n, &n->p, n->L.n) && dalloc(n, &n->g, n->L.n) && dalloc(n, &n->m, n->L.n) &&
            dalloc(n, &n->v, n->L.n) && dalloc(n, &n->pb, n->L.n) && dalloc(n, &n->x, B * HW * 2) &&
            dalloc(n, &n->z, B * HW * 16) && dalloc(n, &n->z4, B * HID) &&
            dalloc(n, &n->logits, B * NCLS) && dalloc(n, &n->dA, B * HW * 4) &&
            dalloc(n, &n->part, pf) && dalloc(n, &n->cpart, 1024 * 64) && dalloc(n, &n->loss, 1) &&
            dalloc(n, &n->h, B * HID) && dalloc(n, &n->dl, B * NCLS) && dalloc(n, &n->dh, B * HID) &&
            dalloc(n, &n->dZ, B * HW * 16);
  for (int k = 0; k < 3 && ok; ++k) {
    const int64_t cnt = B * (HW >> (2 * (k + 1))) * CH[k + 1];
    ok = dalloc(n, &n->a[k + 1], cnt) && dalloc(n, &n->idx[k], cnt);
  }
  if (ok) ok = hipMemset(n->m, 0, n->L.n * 4) == hipSuccess && hipMemset(n->v, 0, n->L.n * 4) == hipSuccess &&
               hipMemset(n->p, 0, n->L.n * 4) == hipSuccess && hipMemset(n->pb, 0, n->L.n * 2) == hipSuccess;
  if (!ok) {
    free_net(n);
    set_error("triage network allocation failed");
    return MDSM_ERR_NOMEM;
  }
  n->a[0] = n->x;
  *out = n;
  return MDSM_OK;
}

extern "C" int mdsm_cnn_destroy(mdsm_cnn n) {
  if (n) free_net(n);
  return MDSM_OK;
}

extern "C" int mdsm_cnn_param_count(mdsm_cnn n, uint64_t *count) {
  MDSM_REQUIRE(n && count, MDSM_ERR_ARG, "null argument");
  *count = (uint64_t)n->L.n;
  return MDSM_OK;
}

extern "C" int mdsm_cnn_shape(mdsm_cnn n, int *height, int *width, int *max_batch) {
  MDSM_REQUIRE(n, MDSM_ERR_ARG, "null network");
  if (height) *height = n->H;
  if (width) *width = n->W;
  if (max_batch) *max_batch = n->Bmax;
  return MDSM_OK;
}

extern "C" int mdsm_cnn_set_params(mdsm_cnn n, const float *host) {
  MDSM_REQUIRE(n && host, MDSM_ERR_ARG, "null argument");
  const auto map = device_index(n);
  std::vector<float> dev(n->L.n);
  for (int64_t i = 0; i < n->L.n; ++i) dev[map[i]] = host[i];
  MDSM_HIP(hipMemcpy(n->p, dev.data(), n->L.n * 4, hipMemcpyHostToDevice));
  MDSM_HIP(hipMemset(n->m, 0, n->L.n * 4));
  MDSM_HIP(hipMemset(n->v, 0, n->L.n * 4));
  n->step = 0;
  int rc = cnn::to_bf16(n->p, n->pb, n->L.n, nullptr);
  if (rc == MDSM_OK) MDSM_HIP(hipDeviceSynchronize());
  return rc;
}

static int get_canonical(mdsm_cnn n, const float *dsrc, float *host) {
  MDSM_REQUIRE(n && host, MDSM_ERR_ARG, "null argument");
  std::vector<float> dev(n->L.n);
  MDSM_HIP(hipDeviceSynchronize());
  MDSM_HIP(hipMemcpy(dev.data(), dsrc, n->L.n * 4, hipMemcpyDeviceToHost));
  const auto map = device_index(n);
  for (int64_t i = 0; i < n->L.n; ++i) host[i] = dev[map[i]];
  return MDSM_OK;
}

extern "C" int mdsm_cnn_get_params(mdsm_cnn n, float *host) { return get_canonical(n, n ? n->p : nullptr, host); }
extern "C" int mdsm_cnn_get_grads(mdsm_cnn n, float *host) { return get_canonical(n, n ? n->g : nullptr, host); }

extern "C" int mdsm_cnn_init(mdsm_cnn n, uint64_t seed) {
  MDSM_REQUIRE(n, MDSM_ERR_ARG, "null network");
  // He-normal weights (std sqrt(2 / fan_in)), zero biases
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<float> host(n->L.n, 0.f);
  const Layout &L = n->L;
  auto fill = [&](int64_t off, int64_t cnt, double fan) {
    const double sd = std::sqrt(2.0 / fan);
    for (int64_t i = 0; i < cnt; ++i) host[off + i] = (float)(nd(rng) * sd);
  };
  for (int k = 0; k < 3; ++k) fill(L.w[k], (int64_t)CH[k + 1] * CH[k] * 9, 9.0 * CH[k]);
  fill(L.fw1, (int64_t)L.fin * HID, L.fin);
  fill(L.fw2, HID * NCLS, HID);
  return mdsm_cnn_set_params(n, host.data());
}

extern "C" int mdsm_cnn_forward(mdsm_cnn n, const uint16_t *d_x, uint32_t batch, float *d_logits,
                                void *stream) {
  MDSM_REQUIRE(n && d_x && batch >= 1 && (int)batch <= n->Bmax, MDSM_ERR_ARG,
               "batch must be in 1..max_batch");
  hipStream_t s = as_stream(stream);
  const int64_t B = batch;
  const Layout &L = n->L;
  int rc = cnn::nchw_to_nhwc(d_x, B, 2, n->H, n->W, n->x, s);
  for (int k = 0; k < 3 && rc == MDSM_OK; ++k) {
    const int Hk = n->H >> k, Wk = n->W >> k;
    rc = cnn::conv_forward(n->a[k], B, Hk, Wk, CH[k], n->pb + L.w[k], CH[k + 1], n->z, s);
    if (rc == MDSM_OK) rc = cnn::relu_pool(n->z, n->p + L.b[k], B, Hk, Wk, CH[k + 1], n->a[k + 1], n->idx[k], s);
  }
  if (rc == MDSM_OK) {  // fc1: z4 = a3 fw1^T + fb1, h = relu(z4)
    const int sp = choose_splits(B, HID, L.fin);
    rc = cnn::gemm_dense(n->a[3], L.fin, 1, n->pb + L.fw1, 1, L.fin, B, HID, L.fin, sp, n->part, s);
    if (rc == MDSM_OK)
      rc = cnn::epilogue(n->part, cnn::splits_used(L.fin, sp), B, HID, n->p + L.fb1, n->z4, nullptr, 1, n->h, s);
  }
  if (rc == MDSM_OK) {  // fc2
    const int sp = choose_splits(B, NCLS, HID);
    rc = cnn::gemm_dense(n->h, HID, 1, n->pb + L.fw2, 1, HID, B, NCLS, HID, sp, n->part, s);
    if (rc == MDSM_OK)
      rc = cnn::epilogue(n->part, cnn::splits_used(HID, sp), B, NCLS, n->p + L.fb2, n->logits, nullptr, 0, nullptr, s);
  }
  if (rc == MDSM_OK && d_logits)
    MDSM_HIP(hipMemcpyAsync(d_logits, n->logits, B * NCLS * 4, hipMemcpyDeviceToDevice, s));
  n->last_batch = rc == MDSM_OK ? (int)batch : 0;
  return rc;
}

extern "C" int mdsm_cnn_backward(mdsm_cnn n, const int32_t *d_labels, uint32_t batch, float *d_loss,
                                 void *stream) {
  MDSM_REQUIRE(n && d_labels, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE((int)batch == n->last_batch, MDSM_ERR_STATE, "backward needs a forward of the same batch");
  hipStream_t s = as_stream(stream);
  const int64_t B = batch;
  const Layout &L = n->L;
  int rc = cnn::softmax_ce(n->logits, d_labels, (int)B, n->dl, n->loss, s);
  if (rc == MDSM_OK && d_loss) MDSM_HIP(hipMemcpyAsync(d_loss, n->loss, 4, hipMemcpyDeviceToDevice, s));
  // fc2: dW = dl^T h, db = colsum(dl), dh = (dl W2) * (z4 > 0)
  if (rc == MDSM_OK) {
    const int sp = choose_splits(NCLS, HID, B);
    rc = cnn::gemm_dense(n->dl, 1, NCLS, n->h, HID, 1, NCLS, HID, B, sp, n->part, s);
    if (rc == MDSM_OK)
      rc = cnn::epilogue(n->part, cnn::splits_used(B, sp), NCLS, HID, nullptr, n->g + L.fw2, nullptr, 0, nullptr, s);
  }
  if (rc == MDSM_OK) rc = cnn::column_sums(n->dl, B, NCLS, n->cpart, 1024, n->g + L.fb2, s);
  if (rc == MDSM_OK) {
    rc = cnn::gemm_dense(n->dl, NCLS, 1, n->pb + L.fw2, HID, 1, B, HID, NCLS, 1, n->part, s);
    if (rc == MDSM_OK) rc = cnn::epilogue(n->part, 1, B, HID, nullptr, nullptr, n->z4, 0, n->dh, s);
  }
  // fc1: dW = dh^T a3, db = colsum(dh), da3 = dh W1
  if (rc == MDSM_OK) {
    const int sp = choose_splits(HID, L.fin, B);
    rc = cnn::gemm_dense(n->dh, 1, HID, n->a[3], L.fin, 1, HID, L.fin, B, sp, n->part, s);
    if (rc == MDSM_OK)
      rc = cnn::epilogue(n->part, cnn::splits_used(B, sp), HID, L.fin, nullptr, n->g + L.fw1, nullptr, 0, nullptr, s);
  }
  if (rc == MDSM_OK) rc = cnn::column_sums(n->dh, B, HID, n->cpart, 1024, n->g + L.fb1, s);
  if (rc == MDSM_OK) rc = cnn::gemm_dense(n->dh, HID, 1, n->pb + L.fw1, L.fin, 1, B, L.fin, HID, 1, n->dA, s);
  // conv stack, top down: dA holds the gradient of the layer's pooled output
  for (int k = 2; k >= 0 && rc == MDSM_OK; --k) {
    const int Hk = n->H >> k, Wk = n->W >> k, Ci = CH[k], Co = CH[k + 1];
    const int64_t P = B * Hk * Wk;
    rc = cnn::pool_backward(n->dA, n->idx[k], B, Hk, Wk, Co, n->dZ, s);
    if (rc == MDSM_OK) rc = cnn::column_sums(n->dZ, P, Co, n->cpart, 1024, n->g + L.b[k], s);
    if (rc == MDSM_OK) {
      const int sp = choose_splits(Co, 9 * Ci, P);
      rc = cnn::conv_weight_grad(n->dZ, n->a[k], B, Hk, Wk, Ci, Co, sp, n->part, s);
      if (rc == MDSM_OK)
        rc = cnn::epilogue(n->part, cnn::splits_used(P, sp), Co, 9 * Ci, nullptr, n->g + L.w[k], nullptr, 0, nullptr, s);
    }
    if (rc == MDSM_OK && k > 0) rc = cnn::conv_input_grad(n->dZ, B, Hk, Wk, Ci, Co, n->pb + L.w[k], n->dA, s);
  }
  return rc;
}

extern "C" int mdsm_cnn_adam_step(mdsm_cnn n, float lr, float weight_decay, void *stream) {
  MDSM_REQUIRE(n && lr > 0, MDSM_ERR_ARG, "bad arguments");
  n->step++;
  return cnn::adam(n->p, n->g, n->m, n->v, n->pb, n->L.n, lr, 0.9f, 0.999f, 1e-8f, n->step,
                   weight_decay, as_stream(stream));
}

extern "C" int mdsm_cnn_train_step(mdsm_cnn n, const uint16_t *d_x, const int32_t *d_labels,
                                   uint32_t batch, float lr, float *d_loss, void *stream) {
  int rc = mdsm_cnn_forward(n, d_x, batch, nullptr, stream);
  if (rc == MDSM_OK) rc = mdsm_cnn_backward(n, d_labels, batch, d_loss, stream);
  if (rc == MDSM_OK) rc = mdsm_cnn_adam_step(n, lr, 0.f, stream);
  return rc;
}

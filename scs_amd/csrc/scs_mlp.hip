// Training and stand-alone inference kernels of the branching MLP.
//
// Training runs in FP32 on the vector ALUs (forward, backward, weight
// gradients, Adam), one launch sequence per mini-batch; inference uses the
// same BF16 MFMA tile routine as the search engine (scs_mlp.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "scs_mlp.h"

// Flat parameter layout = scs_model (include/scs.h).
#define P_W1 0
#define P_B1 (P_W1 + 8 * MLP_H)
#define P_W2 (P_B1 + MLP_H)
#define P_B2 (P_W2 + MLP_H * MLP_H)
#define P_W3 (P_B2 + MLP_H)
#define P_B3 (P_W3 + MLP_H)
#define P_MEAN (P_B3 + 1)
#define P_SCALE (P_MEAN + 8)
#define P_TRAIN P_MEAN  // trainable parameters precede mean/scale
#define P_TOTAL (P_SCALE + 8 + 7)

static_assert(sizeof(scs_model) == P_TOTAL * sizeof(float), "scs_model layout");

// Per-sample forward and backward pass.  Stores the activations and
// pre-activation gradients the weight-gradient kernel reduces over.
__global__ void mlp_fwd_bwd(const float* __restrict__ P, const float* __restrict__ X,
                            const float* __restrict__ Y, const u32* __restrict__ idx, u32 B,
                            float posw, float* __restrict__ xs, float* __restrict__ h1,
                            float* __restrict__ dz1, float* __restrict__ h2,
                            float* __restrict__ dz2, float* __restrict__ dy,
                            float* __restrict__ loss) {
  const u32 s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const u32 r = idx[s];
  float x[8];
  for (int k = 0; k < 8; k++) x[k] = 0.f;
  for (int k = 0; k < MLP_IN; k++) x[k] = (X[r * MLP_IN + k] - P[P_MEAN + k]) * P[P_SCALE + k];
  for (int k = 0; k < 8; k++) xs[s * 8 + k] = x[k];
  float z1[MLP_H], a1[MLP_H];
  for (int j = 0; j < MLP_H; j++) {
    float z = P[P_B1 + j];
    for (int k = 0; k < 8; k++) z += x[k] * P[P_W1 + k * MLP_H + j];
    z1[j] = z;
    a1[j] = z > 0.f ? z : 0.f;
    h1[s * MLP_H + j] = a1[j];
  }
  float z2[MLP_H];
  float out = P[P_B3];
  for (int j = 0; j < MLP_H; j++) {
    float z = P[P_B2 + j];
    for (int i = 0; i < MLP_H; i++) z += a1[i] * P[P_W2 + i * MLP_H + j];
    z2[j] = z;
    float a = z > 0.f ? z : 0.f;
    h2[s * MLP_H + j] = a;
    out += a * P[P_W3 + j];
  }
  const float t = Y[r];
  const float w = t > 0.5f ? posw : 1.f;
  // binary cross-entropy with logits, mean over the batch
  const float sp = out > 0.f ? out + log1pf(expf(-out)) : log1pf(expf(out));
  atomicAdd(loss, w * (sp - t * out) / (float)B);
  const float sig = 1.f / (1.f + expf(-out));
  const float g = w * (sig - t) / (float)B;
  dy[s] = g;
  float d2[MLP_H];
  for (int j = 0; j < MLP_H; j++) {
    d2[j] = z2[j] > 0.f ? g * P[P_W3 + j] : 0.f;
    dz2[s * MLP_H + j] = d2[j];
  }
  for (int i = 0; i < MLP_H; i++) {
    float d = 0.f;
    for (int j = 0; j < MLP_H; j++) d += P[P_W2 + i * MLP_H + j] * d2[j];
    dz1[s * MLP_H + i] = z1[i] > 0.f ? d : 0.f;
  }
}

// One thread per trainable parameter: sum its gradient over the batch.
__global__ void mlp_wgrad(u32 B, const float* __restrict__ xs, const float* __restrict__ h1,
                          const float* __restrict__ dz1, const float* __restrict__ h2,
                          const float* __restrict__ dz2, const float* __restrict__ dy,
                          float* __restrict__ G) {
  const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P_TRAIN) return;
  float g = 0.f;
  if (p < P_B1) {
    const u32 k = (p - P_W1) / MLP_H, j = (p - P_W1) % MLP_H;
    for (u32 s = 0; s < B; s++) g += xs[s * 8 + k] * dz1[s * MLP_H + j];
  } else if (p < P_W2) {
    const u32 j = p - P_B1;
    for (u32 s = 0; s < B; s++) g += dz1[s * MLP_H + j];
  } else if (p < P_B2) {
    const u32 i = (p - P_W2) / MLP_H, j = (p - P_W2) % MLP_H;
    for (u32 s = 0; s < B; s++) g += h1[s * MLP_H + i] * dz2[s * MLP_H + j];
  } else if (p < P_W3) {
    const u32 j = p - P_B2;
    for (u32 s = 0; s < B; s++) g += dz2[s * MLP_H + j];
  } else if (p < P_B3) {
    const u32 j = p - P_W3;
    for (u32 s = 0; s < B; s++) g += dy[s] * h2[s * MLP_H + j];
  } else {
    for (u32 s = 0; s < B; s++) g += dy[s];
  }
  G[p] = g;
}

__global__ void mlp_adam(float* __restrict__ P, const float* __restrict__ G,
                         float* __restrict__ m, float* __restrict__ v, float lr, float b1t,
                         float b2t) {
  const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P_TRAIN) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float g = G[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.f - b1t), vh = vv / (1.f - b2t);
  P[P_W1 + p] -= lr * mh / (sqrtf(vh) + eps);
}

// Stand-alone inference with the engine's MFMA tile routine: one wave per
// 16 rows.  X: n x 6 raw features.
__global__ void __launch_bounds__(64) mlp_infer_k(const MlpDev* M, const float* __restrict__ X,
                                                  u32 n, float* __restrict__ out) {
  __shared__ u16 h1s[16 * MLP_H];
  const u32 base = blockIdx.x * 16;
  auto feat = [&](u32 row) {
    MlpFeat f{0, 0, 0, 0, 0, 0.f};
    const u32 r = base + row;
    if (r < n) {
      f.dC = (u32)X[r * MLP_IN + 0];
      f.dR = (u32)X[r * MLP_IN + 1];
      f.dist = (u32)X[r * MLP_IN + 2];
      f.core = (u32)X[r * MLP_IN + 3];
      f.slack = (int)X[r * MLP_IN + 4];
      f.cscore = X[r * MLP_IN + 5];
    }
    return f;
  };
  const float s = mlp_tile(M, feat, h1s);
  const u32 row = threadIdx.x & 15u;
  if (threadIdx.x < 16 && base + row < n) out[base + row] = s;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
static u16 host_f2bf(float f) {
  u32 x;
  memcpy(&x, &f, 4);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (u16)(x >> 16);
}

void scs_pack_model(const scs_model* m, MlpDev* d) {
  memset(d, 0, sizeof(*d));
  for (int k = 0; k < 8; k++)
    for (int j = 0; j < MLP_H; j++) d->w1[k][j] = k < MLP_IN ? host_f2bf(m->w1[k][j]) : 0;
  for (int i = 0; i < MLP_H; i++)
    for (int j = 0; j < MLP_H; j++) d->w2[i][j] = host_f2bf(m->w2[i][j]);
  for (int j = 0; j < MLP_H; j++) {
    d->b1[j] = m->b1[j];
    d->b2[j] = m->b2[j];
    d->w3[j] = m->w3[j];
  }
  d->b3 = m->b3;
  for (int k = 0; k < 8; k++) {
    d->mean[k] = m->mean[k];
    d->scale[k] = m->scale[k];
  }
}

#define HCK(x)                          \
  do {                                  \
    if ((x) != hipSuccess) return SCS_EHIP; \
  } while (0)

extern "C" int scs_model_init(scs_model* m, uint64_t seed) {
  if (!m) return SCS_EINVAL;
  memset(m, 0, sizeof(*m));
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int k = 0; k < MLP_IN; k++)
    for (int j = 0; j < MLP_H; j++) m->w1[k][j] = nd(rng) * sqrtf(2.f / MLP_IN);
  for (int i = 0; i < MLP_H; i++)
    for (int j = 0; j < MLP_H; j++) m->w2[i][j] = nd(rng) * sqrtf(2.f / MLP_H);
  for (int j = 0; j < MLP_H; j++) m->w3[j] = nd(rng) * sqrtf(1.f / MLP_H);
  for (int k = 0; k < 8; k++) m->scale[k] = 1.f;
  return SCS_OK;
}

namespace {
struct TrainBufs {
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *X = nullptr, *Y = nullptr;
  float *xs = nullptr, *h1 = nullptr, *dz1 = nullptr, *h2 = nullptr, *dz2 = nullptr, *dy = nullptr,
        *loss = nullptr;
  u32* idx = nullptr;
  ~TrainBufs() {
    void* p[] = {P, G, M, V, X, Y, xs, h1, dz1, h2, dz2, dy, loss, idx};
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};
int alloc_train(TrainBufs& T, u32 n, u32 B) {
  HCK(hipMalloc(&T.P, sizeof(float) * P_TOTAL));
  HCK(hipMalloc(&T.G, sizeof(float) * P_TOTAL));
  HCK(hipMalloc(&T.M, sizeof(float) * P_TOTAL));
  HCK(hipMalloc(&T.V, sizeof(float) * P_TOTAL));
  HCK(hipMalloc(&T.X, sizeof(float) * MLP_IN * (u64)n));
  HCK(hipMalloc(&T.Y, sizeof(float) * (u64)n));
  HCK(hipMalloc(&T.xs, sizeof(float) * 8 * (u64)B));
  HCK(hipMalloc(&T.h1, sizeof(float) * MLP_H * (u64)B));
  HCK(hipMalloc(&T.dz1, sizeof(float) * MLP_H * (u64)B));
  HCK(hipMalloc(&T.h2, sizeof(float) * MLP_H * (u64)B));
  HCK(hipMalloc(&T.dz2, sizeof(float) * MLP_H * (u64)B));
  HCK(hipMalloc(&T.dy, sizeof(float) * (u64)B));
  HCK(hipMalloc(&T.loss, sizeof(float)));
  HCK(hipMalloc(&T.idx, sizeof(u32) * (u64)B));
  return SCS_OK;
}
int grad_batch(TrainBufs& T, u32 B, float posw) {
  HCK(hipMemset(T.loss, 0, sizeof(float)));
  mlp_fwd_bwd<<<(B + 127) / 128, 128>>>(T.P, T.X, T.Y, T.idx, B, posw, T.xs, T.h1, T.dz1, T.h2,
                                        T.dz2, T.dy, T.loss);
  HCK(hipGetLastError());
  mlp_wgrad<<<(P_TRAIN + 255) / 256, 256>>>(B, T.xs, T.h1, T.dz1, T.h2, T.dz2, T.dy, T.G);
  HCK(hipGetLastError());
  return SCS_OK;
}
}  // namespace

// Gradient of the mean weighted BCE over rows 0..n-1 (for checks).
extern "C" int scs_mlp_grad(const scs_model* m, const float* X, const float* y, uint32_t n,
                            float pos_weight, scs_model* grad, float* loss) {
  if (!m || !X || !y || !grad || n == 0) return SCS_EINVAL;
  TrainBufs T;
  int rc = alloc_train(T, n, n);
  if (rc) return rc;
  HCK(hipMemcpy(T.P, m, sizeof(float) * P_TOTAL, hipMemcpyHostToDevice));
  HCK(hipMemcpy(T.X, X, sizeof(float) * MLP_IN * (u64)n, hipMemcpyHostToDevice));
  HCK(hipMemcpy(T.Y, y, sizeof(float) * (u64)n, hipMemcpyHostToDevice));
  std::vector<u32> idx(n);
  for (u32 i = 0; i < n; i++) idx[i] = i;
  HCK(hipMemcpy(T.idx, idx.data(), sizeof(u32) * n, hipMemcpyHostToDevice));
  rc = grad_batch(T, n, pos_weight);
  if (rc) return rc;
  memset(grad, 0, sizeof(*grad));
  HCK(hipMemcpy(grad, T.G, sizeof(float) * P_TRAIN, hipMemcpyDeviceToHost));
  if (loss) HCK(hipMemcpy(loss, T.loss, sizeof(float), hipMemcpyDeviceToHost));
  return SCS_OK;
}

// Mini-batch Adam on the GPU.  Input normalisation (mean/scale) is taken
// from the data before training starts.
extern "C" int scs_mlp_train(scs_model* m, const float* X, const float* y, uint32_t n,
                             uint32_t epochs, uint32_t batch, float lr, float pos_weight,
                             uint64_t seed, float* final_loss) {
  if (!m || !X || !y || n == 0 || batch == 0) return SCS_EINVAL;
  for (int k = 0; k < MLP_IN; k++) {
    double s = 0, s2 = 0;
    for (u32 i = 0; i < n; i++) {
      s += X[i * MLP_IN + k];
      s2 += (double)X[i * MLP_IN + k] * X[i * MLP_IN + k];
    }
    const double mu = s / n, var = s2 / n - mu * mu;
    m->mean[k] = (float)mu;
    m->scale[k] = var > 1e-12 ? (float)(1.0 / sqrt(var)) : 1.f;
  }
  const u32 B = std::min<u32>(batch, n);
  TrainBufs T;
  int rc = alloc_train(T, n, B);
  if (rc) return rc;
  HCK(hipMemcpy(T.P, m, sizeof(float) * P_TOTAL, hipMemcpyHostToDevice));
  HCK(hipMemset(T.M, 0, sizeof(float) * P_TOTAL));
  HCK(hipMemset(T.V, 0, sizeof(float) * P_TOTAL));
  HCK(hipMemcpy(T.X, X, sizeof(float) * MLP_IN * (u64)n, hipMemcpyHostToDevice));
  HCK(hipMemcpy(T.Y, y, sizeof(float) * (u64)n, hipMemcpyHostToDevice));
  std::mt19937_64 rng(seed);
  std::vector<u32> perm(n);
  for (u32 i = 0; i < n; i++) perm[i] = i;
  u64 t = 0;
  float last = 0.f;
  for (u32 ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    double ep_loss = 0;
    u32 nb = 0;
    for (u32 b0 = 0; b0 + B <= n; b0 += B) {
      HCK(hipMemcpy(T.idx, perm.data() + b0, sizeof(u32) * B, hipMemcpyHostToDevice));
      rc = grad_batch(T, B, pos_weight);
      if (rc) return rc;
      t++;
      mlp_adam<<<(P_TRAIN + 255) / 256, 256>>>(T.P, T.G, T.M, T.V, lr,
                                               (float)pow(0.9, (double)t),
                                               (float)pow(0.999, (double)t));
      HCK(hipGetLastError());
      float l = 0.f;
      HCK(hipMemcpy(&l, T.loss, sizeof(float), hipMemcpyDeviceToHost));
      ep_loss += l;
      nb++;
    }
    last = nb ? (float)(ep_loss / nb) : 0.f;
  }
  HCK(hipMemcpy(m, T.P, sizeof(float) * P_TRAIN, hipMemcpyDeviceToHost));
  if (final_loss) *final_loss = last;
  return SCS_OK;
}

extern "C" int scs_mlp_infer(const scs_model* m, const float* X, uint32_t n, float* scores) {
  if (!m || !X || !scores || n == 0) return SCS_EINVAL;
  MlpDev h;
  scs_pack_model(m, &h);
  MlpDev* d = nullptr;
  float *dX = nullptr, *dO = nullptr;
  HCK(hipMalloc(&d, sizeof(MlpDev)));
  HCK(hipMalloc(&dX, sizeof(float) * MLP_IN * (u64)n));
  HCK(hipMalloc(&dO, sizeof(float) * (u64)n));
  HCK(hipMemcpy(d, &h, sizeof(MlpDev), hipMemcpyHostToDevice));
  HCK(hipMemcpy(dX, X, sizeof(float) * MLP_IN * (u64)n, hipMemcpyHostToDevice));
  mlp_infer_k<<<(n + 15) / 16, 64>>>(d, dX, n, dO);
  HCK(hipGetLastError());
  HCK(hipMemcpy(scores, dO, sizeof(float) * n, hipMemcpyDeviceToHost));
  (void)hipFree(d);
  (void)hipFree(dX);
  (void)hipFree(dO);
  return SCS_OK;
}

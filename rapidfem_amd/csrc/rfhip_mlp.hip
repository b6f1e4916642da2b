// The (dt, m) proposer of the adaptive Krylov exponential: a 12 -> 64 -> 64
// -> 2 multilayer perceptron with tanh hidden layers, FP32, and its training
// kernels (forward, backward with deterministic per-block gradient partials,
// ordered reduction, Adam). No counterpart in the reference, whose
// exponential integrator takes a fixed step and Krylov dimension
// (crates/rapidfem-td/src/session.rs:700-726).
//
// Parameter block (floats), row-major weights:
//   W1[64][12] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]          (5122)
// followed by the normalisation block
//   in_mean[12] in_inv_scale[12] out_mean[2] out_scale[2]           (28)
// The network maps normalised features xn = (x - in_mean) * in_inv_scale
// to normalised outputs yn; the caller sees y = yn * out_scale + out_mean.
// The training loss is the weighted mean of |yn - tn|^2 with
// tn = (t - out_mean) / out_scale.
//
// One 64-lane wave per block: lane j owns hidden unit j of both layers, so
// every activation is one register and every weight row one LDS row. The
// weights sit in LDS with rows padded to an odd stride, so the row reads of
// the forward pass and the column reads of the backward pass are both free
// of bank conflicts.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

#include "rfhip_impl.hpp"

namespace rfhip {
namespace {

constexpr int kIn = 12, kH = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kOut * kH, kNParam = oB3 + kOut;
constexpr int oInMean = 0, oInInv = kIn, oOutMean = 2 * kIn, oOutScale = 2 * kIn + kOut, kNNorm = 2 * kIn + 2 * kOut;
static_assert(kNParam == 5122, "parameter block layout");
constexpr int kW1S = kIn + 1;  // padded LDS strides
constexpr int kW2S = kH + 1;
constexpr int kGradBlocks = 128;  // fixed, so the reduction order never depends on the device

struct MlpLds {
  float w1[kH * kW1S];
  float w2[kH * kW2S];
  float w3[kOut * kH];
  float b1[kH], b2[kH], b3[kOut];
  float norm[kNNorm];
  float x[kIn];
  float h1[kH];
  float d2[kH];
  float y[kOut];
};

__device__ void load_params(MlpLds &L, const float *__restrict__ P, const float *__restrict__ norm) {
  const int j = threadIdx.x;
  for (int i = 0; i < kIn; ++i) L.w1[j * kW1S + i] = P[oW1 + j * kIn + i];
  for (int k = 0; k < kH; ++k) L.w2[j * kW2S + k] = P[oW2 + j * kH + k];
  for (int o = 0; o < kOut; ++o) L.w3[o * kH + j] = P[oW3 + o * kH + j];
  L.b1[j] = P[oB1 + j];
  L.b2[j] = P[oB2 + j];
  if (j < kOut) L.b3[j] = P[oB3 + j];
  if (j < kNNorm) L.norm[j] = norm[j];
  __syncthreads();
}

// Forward pass of one sample whose normalised features are in L.x: returns
// this lane's h1, h2 and leaves the normalised outputs in L.y.
__device__ void forward(MlpLds &L, float &h1, float &h2) {
  const int j = threadIdx.x;
  float a = L.b1[j];
  for (int i = 0; i < kIn; ++i) a = fmaf(L.w1[j * kW1S + i], L.x[i], a);
  h1 = tanhf(a);
  L.h1[j] = h1;
  __syncthreads();
  a = L.b2[j];
  for (int k = 0; k < kH; ++k) a = fmaf(L.w2[j * kW2S + k], L.h1[k], a);
  h2 = tanhf(a);
  // output layer: a wave-wide sum of w3[o][j] * h2[j], one DPP-free
  // butterfly over the 64 lanes per output
  for (int o = 0; o < kOut; ++o) {
    float v = L.w3[o * kH + j] * h2;
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (j == 0) L.y[o] = v + L.b3[o];
  }
  __syncthreads();
}

__global__ __launch_bounds__(64) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ norm,
                                                    const float *__restrict__ X, int64_t n, float *__restrict__ Y) {
  __shared__ MlpLds L;
  load_params(L, P, norm);
  const int j = threadIdx.x;
  for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
    if (j < kIn) L.x[j] = (X[s * kIn + j] - L.norm[oInMean + j]) * L.norm[oInInv + j];
    __syncthreads();
    float h1, h2;
    forward(L, h1, h2);
    if (j < kOut) Y[s * kOut + j] = L.y[j] * L.norm[oOutScale + j] + L.norm[oOutMean + j];
    __syncthreads();
  }
}

// Gradient of the weighted loss sum_s w_s |yn_s - tn_s|^2 * inv_wsum over the
// samples idx[0..n) (or 0..n when idx is null): block b takes samples
// b, b + gridDim.x, ... and writes its partial gradient (and its partial
// loss, entry kNParam) to partials[b * (kNParam + 1) ...].
__global__ __launch_bounds__(64) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ norm,
                                                 const float *__restrict__ X, const float *__restrict__ T,
                                                 const float *__restrict__ W, const int32_t *__restrict__ idx,
                                                 int64_t n, float inv_wsum, float *__restrict__ partials) {
  __shared__ MlpLds L;
  load_params(L, P, norm);
  const int j = threadIdx.x;
  float gw1[kIn], gw2[kH], gw3[kOut];
  #pragma unroll
  for (int i = 0; i < kIn; ++i) gw1[i] = 0.0f;
  #pragma unroll
  for (int k = 0; k < kH; ++k) gw2[k] = 0.0f;
  for (int o = 0; o < kOut; ++o) gw3[o] = 0.0f;
  float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, loss = 0.0f;
  for (int64_t q = blockIdx.x; q < n; q += gridDim.x) {
    const int64_t s = idx ? (int64_t)idx[q] : q;
    if (j < kIn) L.x[j] = (X[s * kIn + j] - L.norm[oInMean + j]) * L.norm[oInInv + j];
    __syncthreads();
    float h1, h2;
    forward(L, h1, h2);
    const float ws = (W ? W[s] : 1.0f) * inv_wsum;
    float d3[kOut];
    for (int o = 0; o < kOut; ++o) {
      const float tn = (T[s * kOut + o] - L.norm[oOutMean + o]) / L.norm[oOutScale + o];
      const float e = L.y[o] - tn;
      if (j == 0) loss = fmaf(ws * e, e, loss);
      d3[o] = 2.0f * ws * e;
      gw3[o] = fmaf(d3[o], h2, gw3[o]);
    }
    if (j == 0) gb3 += d3[0];
    if (j == 1) gb3 += d3[1];
    float d2 = 0.0f;
    for (int o = 0; o < kOut; ++o) d2 = fmaf(L.w3[o * kH + j], d3[o], d2);
    d2 *= 1.0f - h2 * h2;
    L.d2[j] = d2;
    __syncthreads();
    #pragma unroll
    for (int k = 0; k < kH; ++k) gw2[k] = fmaf(d2, L.h1[k], gw2[k]);
    gb2 += d2;
    float d1 = 0.0f;  // column j of W2 against d2
    for (int k = 0; k < kH; ++k) d1 = fmaf(L.w2[k * kW2S + j], L.d2[k], d1);
    d1 *= 1.0f - h1 * h1;
    #pragma unroll
    for (int i = 0; i < kIn; ++i) gw1[i] = fmaf(d1, L.x[i], gw1[i]);
    gb1 += d1;
    __syncthreads();  // L.x, L.h1, L.d2 and L.y are rewritten by the next sample
  }
  float *out = partials + (int64_t)blockIdx.x * (kNParam + 1);
  #pragma unroll
  for (int i = 0; i < kIn; ++i) out[oW1 + j * kIn + i] = gw1[i];
  #pragma unroll
  for (int k = 0; k < kH; ++k) out[oW2 + j * kH + k] = gw2[k];
  for (int o = 0; o < kOut; ++o) out[oW3 + o * kH + j] = gw3[o];
  out[oB1 + j] = gb1;
  out[oB2 + j] = gb2;
  if (j < kOut) out[oB3 + j] = gb3;
  if (j == 0) out[kNParam] = loss;
}

// g[p] = sum over blocks of partials[b][p], in block order (FP64 sum), for
// the kNParam gradient entries and the loss; the loss is also added to
// *loss_acc when given (the epoch's running sum).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ partials, int nblk, float *__restrict__ g,
                                                    double *__restrict__ loss_out, double *__restrict__ loss_acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kNParam) return;
  double acc = 0.0;
  for (int b = 0; b < nblk; ++b) acc += (double)partials[(int64_t)b * (kNParam + 1) + p];
  if (p < kNParam) {
    g[p] = (float)acc;
  } else {
    if (loss_out) *loss_out = acc;
    if (loss_acc) *loss_acc += acc;
  }
}

// Adam (Kingma & Ba 2015) on every parameter; c1 = 1 - beta1^t and
// c2 = 1 - beta2^t are the bias corrections of step t.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ P, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, float lr, float beta1,
                                                  float beta2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const float gp = g[p];
  const float mp = beta1 * m[p] + (1.0f - beta1) * gp;
  const float vp = beta2 * v[p] + (1.0f - beta2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

__global__ void k_zero_f64(double *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return ((splitmix(s) >> 11) + 0.5) / 9007199254740992.0; }

}  // namespace
}  // namespace rfhip

using namespace rfhip;

struct rfhip_mlp {
  rfhip_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  DevBuf params, norm, grad, adam_m, adam_v, partials, loss, io_x, io_y, data_x, data_t, data_w, perm;
  int64_t adam_t = 0;
  ~rfhip_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <class F>
int mlp_guarded(F &&f) {
  try {
    f();
    return RFHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return RFHIP_ENOMEM;
  } catch (const std::exception &e) {
    set_error(e.what());
    return RFHIP_EINVAL;
  }
}

void mlp_sync(rfhip_mlp *m) { RF_HIP(hipStreamSynchronize(m->stream)); }

// Gradient (and loss) of the samples idx[0..n) of the uploaded data into
// m->grad and m->loss[0] (and m->loss[1] += loss when accumulate).
void grad_pass(rfhip_mlp *m, const float *x, const float *t, const float *w, const int32_t *idx, int64_t n,
               float inv_wsum, bool accumulate) {
  const int nblk = (int)std::min<int64_t>(kGradBlocks, n);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(64), 0, m->stream, m->params.as<float>(), m->norm.as<float>(), x, t,
                     w, idx, n, inv_wsum, m->partials.as<float>());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, m->stream, m->partials.as<float>(),
                     nblk, m->grad.as<float>(), m->loss.as<double>(), accumulate ? m->loss.as<double>() + 1 : nullptr);
  RF_HIP(hipGetLastError());
}

void check_data(const float *x, const float *t, int64_t n) {
  RF_REQUIRE(x && t, "null data");
  RF_REQUIRE(n >= 1 && n <= (int64_t)1 << 30, "sample count must be in [1, 2^30]");
}

}  // namespace

extern "C" {

int rfhip_mlp_create(rfhip_ctx *ctx, uint64_t seed, rfhip_mlp **out) {
  return mlp_guarded([&] {
    RF_REQUIRE(ctx && out, "null argument");
    *out = nullptr;
    auto m = std::make_unique<rfhip_mlp>();
    m->ctx = ctx;
    RF_HIP(hipSetDevice(ctx->device));
    RF_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    // Glorot-uniform weights, zero biases, identity normalisation.
    std::vector<float> p(kNParam, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int fan_out, int fan_in) {
      const double a = std::sqrt(6.0 / (fan_in + fan_out));
      for (int q = 0; q < fan_out * fan_in; ++q) p[off + q] = (float)((2.0 * uniform01(s) - 1.0) * a);
    };
    fill(oW1, kH, kIn);
    fill(oW2, kH, kH);
    fill(oW3, kOut, kH);
    std::vector<float> nm(kNNorm, 0.0f);
    for (int i = 0; i < kIn; ++i) nm[oInInv + i] = 1.0f;
    for (int o = 0; o < kOut; ++o) nm[oOutScale + o] = 1.0f;
    m->params.alloc(4 * kNParam);
    m->norm.alloc(4 * kNNorm);
    m->grad.alloc(4 * kNParam);
    m->adam_m.alloc(4 * kNParam);
    m->adam_v.alloc(4 * kNParam);
    m->partials.alloc(4 * (size_t)kGradBlocks * (kNParam + 1));
    m->loss.alloc(8 * 2);
    RF_HIP(hipMemcpyAsync(m->params.p, p.data(), 4 * kNParam, hipMemcpyHostToDevice, m->stream));
    RF_HIP(hipMemcpyAsync(m->norm.p, nm.data(), 4 * kNNorm, hipMemcpyHostToDevice, m->stream));
    launch_fill(m->adam_m.p, 4 * kNParam, m->stream);
    launch_fill(m->adam_v.p, 4 * kNParam, m->stream);
    mlp_sync(m.get());
    *out = m.release();
  });
}

void rfhip_mlp_destroy(rfhip_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->stream);
  delete m;
}

int rfhip_mlp_n_params(void) { return kNParam; }

int rfhip_mlp_get_params(rfhip_mlp *m, float *params, float *norm) {
  return mlp_guarded([&] {
    RF_REQUIRE(m, "null model");
    if (params) RF_HIP(hipMemcpyAsync(params, m->params.p, 4 * kNParam, hipMemcpyDeviceToHost, m->stream));
    if (norm) RF_HIP(hipMemcpyAsync(norm, m->norm.p, 4 * kNNorm, hipMemcpyDeviceToHost, m->stream));
    mlp_sync(m);
  });
}

int rfhip_mlp_set_params(rfhip_mlp *m, const float *params, const float *norm) {
  return mlp_guarded([&] {
    RF_REQUIRE(m, "null model");
    if (norm) {
      for (int i = 0; i < kIn; ++i)
        RF_REQUIRE(std::isfinite(norm[oInMean + i]) && std::isfinite(norm[oInInv + i]), "non-finite normalisation");
      for (int o = 0; o < kOut; ++o)
        RF_REQUIRE(std::isfinite(norm[oOutMean + o]) && std::isfinite(norm[oOutScale + o]) && norm[oOutScale + o] != 0.0f,
                   "output scales must be finite and nonzero");
    }
    if (params) RF_HIP(hipMemcpyAsync(m->params.p, params, 4 * kNParam, hipMemcpyHostToDevice, m->stream));
    if (norm) RF_HIP(hipMemcpyAsync(m->norm.p, norm, 4 * kNNorm, hipMemcpyHostToDevice, m->stream));
    launch_fill(m->adam_m.p, 4 * kNParam, m->stream);
    launch_fill(m->adam_v.p, 4 * kNParam, m->stream);
    m->adam_t = 0;
    mlp_sync(m);
  });
}

int rfhip_mlp_forward(rfhip_mlp *m, const float *x, int64_t n, float *y) {
  return mlp_guarded([&] {
    RF_REQUIRE(m && x && y, "null argument");
    RF_REQUIRE(n >= 1 && n <= (int64_t)1 << 30, "sample count must be in [1, 2^30]");
    m->io_x.alloc(4 * (size_t)n * kIn);
    m->io_y.alloc(4 * (size_t)n * kOut);
    RF_HIP(hipMemcpyAsync(m->io_x.p, x, 4 * (size_t)n * kIn, hipMemcpyHostToDevice, m->stream));
    const int nblk = (int)std::min<int64_t>(n, 2048);
    hipLaunchKernelGGL(k_mlp_forward, dim3(nblk), dim3(64), 0, m->stream, m->params.as<float>(), m->norm.as<float>(),
                       m->io_x.as<float>(), n, m->io_y.as<float>());
    RF_HIP(hipGetLastError());
    RF_HIP(hipMemcpyAsync(y, m->io_y.p, 4 * (size_t)n * kOut, hipMemcpyDeviceToHost, m->stream));
    mlp_sync(m);
  });
}

int rfhip_mlp_loss_grad(rfhip_mlp *m, const float *x, const float *t, const float *w, int64_t n, double *loss,
                        float *grad) {
  return mlp_guarded([&] {
    RF_REQUIRE(m, "null model");
    check_data(x, t, n);
    double wsum = (double)n;
    if (w) {
      wsum = 0.0;
      for (int64_t s = 0; s < n; ++s) {
        RF_REQUIRE(std::isfinite(w[s]) && w[s] >= 0.0f, "sample weights must be finite and >= 0");
        wsum += w[s];
      }
      RF_REQUIRE(wsum > 0.0, "sample weights sum to zero");
    }
    upload(m->io_x, x, (size_t)n * kIn, m->stream);
    upload(m->io_y, t, (size_t)n * kOut, m->stream);
    DevBuf wd;
    if (w) upload(wd, w, (size_t)n, m->stream);
    grad_pass(m, m->io_x.as<float>(), m->io_y.as<float>(), w ? wd.as<float>() : nullptr, nullptr, n,
              (float)(1.0 / wsum), false);
    double l = 0.0;
    RF_HIP(hipMemcpyAsync(&l, m->loss.p, 8, hipMemcpyDeviceToHost, m->stream));
    if (grad) RF_HIP(hipMemcpyAsync(grad, m->grad.p, 4 * kNParam, hipMemcpyDeviceToHost, m->stream));
    mlp_sync(m);
    if (loss) *loss = l;
  });
}

int rfhip_mlp_train(rfhip_mlp *m, const float *x, const float *t, const float *w, int64_t n, int32_t epochs,
                    int32_t batch, float lr, uint64_t seed, double *epoch_loss) {
  return mlp_guarded([&] {
    RF_REQUIRE(m, "null model");
    check_data(x, t, n);
    RF_REQUIRE(epochs >= 1 && batch >= 1, "epochs and batch must be >= 1");
    RF_REQUIRE(std::isfinite(lr) && lr > 0.0f, "learning rate must be positive");
    if (w)
      for (int64_t s = 0; s < n; ++s) RF_REQUIRE(std::isfinite(w[s]) && w[s] >= 0.0f, "sample weights must be finite and >= 0");
    hipStream_t st = m->stream;
    upload(m->data_x, x, (size_t)n * kIn, st);
    upload(m->data_t, t, (size_t)n * kOut, st);
    if (w) upload(m->data_w, w, (size_t)n, st);
    m->perm.alloc(4 * (size_t)n);
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::vector<double> bw;  // per-batch weight sums, for the loss normaliser
    uint64_t rs = seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
      for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int64_t)(uniform01(rs) * (double)(i + 1)) % (i + 1)]);
      // the upload must finish before perm is reshuffled
      RF_HIP(hipMemcpyAsync(m->perm.p, perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, m->loss.as<double>(), 2);
      int64_t nb = 0;
      for (int64_t b0 = 0; b0 < n; b0 += batch, ++nb) {
        const int64_t cnt = std::min<int64_t>(batch, n - b0);
        double wsum = (double)cnt;
        if (w) {
          wsum = 0.0;
          for (int64_t q = b0; q < b0 + cnt; ++q) wsum += w[perm[q]];
        }
        if (wsum <= 0.0) continue;
        grad_pass(m, m->data_x.as<float>(), m->data_t.as<float>(), w ? m->data_w.as<float>() : nullptr,
                  m->perm.as<int32_t>() + b0, cnt, (float)(1.0 / wsum), true);
        ++m->adam_t;
        const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, st, m->params.as<float>(),
                           m->grad.as<float>(), m->adam_m.as<float>(), m->adam_v.as<float>(), lr, b1, b2, eps, c1, c2);
        RF_HIP(hipGetLastError());
      }
      double acc[2] = {0.0, 0.0};
      RF_HIP(hipMemcpyAsync(acc, m->loss.p, 16, hipMemcpyDeviceToHost, st));
      mlp_sync(m);
      if (epoch_loss) epoch_loss[ep] = nb ? acc[1] / (double)nb : 0.0;
    }
  });
}

}  // extern "C"

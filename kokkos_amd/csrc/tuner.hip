// Learned launch-shape tuner: a Kokkos Tools tuning plug-in backed by a small
// FP32 MLP (12 features -> 64 -> 64 -> 16 config logits) with hand-written
// HIP training and inference kernels.
//
// As a tool it implements the tuning callbacks of the Kokkos Tools C
// interface (core/src/impl/Kokkos_Profiling_C_Interface.h:136-276;
// kokkosp_declare_input_type, kokkosp_declare_output_type,
// kokkosp_request_values, kokkosp_begin_context, kokkosp_end_context), so it
// loads like any tuning tool (KOKKOS_TOOLS_LIBS).  When the runtime asks for
// one of the output variables "kokkos.mi355x.range_launch" (16 launch shapes
// of a RangePolicy dispatch), "kokkos.mi355x.team_shape" (16 team-size x
// vector-length pairs) or "kokkos.mi355x.mdrange_tile" (16 tiles), the tool
// turns the declared input variables of the
// request into the 12 features, runs the MLP on the GPU and answers the
// arg-max config.  Decisions are cached per feature key, so the GPU sees one
// inference launch per distinct kernel signature and extent bucket.  Without
// a model the tool leaves the runtime's heuristic value untouched.
//
// Training (kkmx_mlp_train) runs entirely in HIP kernels: forward pass,
// soft-target cross-entropy, backward pass and an Adam step per minibatch.
// Host code reaches HIP only through the C-ABI runtime (kkmx_runtime.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "kkmx_runtime.h"
#include "kkmx_tools.h"

namespace {

constexpr int F  = 12;  // features
constexpr int H  = 64;  // hidden width
constexpr int C  = 16;  // configs
// Parameter layout: W1[H][F] b1[H] W2[H][H] b2[H] W3[C][H] b3[C]
constexpr int oW1 = 0, ob1 = oW1 + H * F, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H, ob3 = oW3 + C * H;
constexpr int NP  = ob3 + C;

// ------------------------------------------------------------------ kernels
// Forward pass, one thread per sample; parameters staged in LDS.  Writes the
// hidden activations (for the backward pass) and the logits.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X, int n,
                                                     float* __restrict__ H1, float* __restrict__ H2,
                                                     float* __restrict__ Z) {
  __shared__ float w[NP];
  for (int i = threadIdx.x; i < NP; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[F], h1[H], h2[H];
#pragma unroll
  for (int k = 0; k < F; ++k) x[k] = X[s * F + k];
  for (int j = 0; j < H; ++j) {
    float a = w[ob1 + j];
#pragma unroll
    for (int k = 0; k < F; ++k) a += w[oW1 + j * F + k] * x[k];
    h1[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < H; ++j) {
    float a = w[ob2 + j];
    for (int k = 0; k < H; ++k) a += w[oW2 + j * H + k] * h1[k];
    h2[j] = a > 0.f ? a : 0.f;
  }
  for (int c = 0; c < C; ++c) {
    float a = w[ob3 + c];
    for (int k = 0; k < H; ++k) a += w[oW3 + c * H + k] * h2[k];
    Z[s * C + c] = a;
  }
  if (H1) {
    for (int j = 0; j < H; ++j) {
      H1[s * H + j] = h1[j];
      H2[s * H + j] = h2[j];
    }
  }
}

// Softmax cross-entropy against a target distribution T, then the gradients
// of the hidden activations; one thread per sample of the minibatch.
// dZ = (softmax(Z) - T) / B; dH2 = W3^T dZ * [H2 > 0]; dH1 = W2^T dH2 * [H1 > 0].
__global__ __launch_bounds__(256) void k_mlp_backward_act(const float* __restrict__ P, const float* __restrict__ Tgt,
                                                          int b, const float* __restrict__ H1,
                                                          const float* __restrict__ H2, const float* __restrict__ Z,
                                                          float* __restrict__ dZ, float* __restrict__ dH1,
                                                          float* __restrict__ dH2, float* __restrict__ loss) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b) return;
  float z[C], m = -INFINITY;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    z[c] = Z[s * C + c];
    m    = fmaxf(m, z[c]);
  }
  float sum = 0.f;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    z[c] = expf(z[c] - m);
    sum += z[c];
  }
  float l = 0.f, dz[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const float p = z[c] / sum, t = Tgt[s * C + c];
    if (t > 0.f) l -= t * logf(fmaxf(p, 1e-30f));
    dz[c]         = (p - t) / float(b);
    dZ[s * C + c] = dz[c];
  }
  atomicAdd(loss, l / float(b));
  float d2[H];
  for (int j = 0; j < H; ++j) {
    float a = 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c) a += P[oW3 + c * H + j] * dz[c];
    d2[j]             = H2[s * H + j] > 0.f ? a : 0.f;
    dH2[s * H + j]    = d2[j];
  }
  for (int k = 0; k < H; ++k) {
    float a = 0.f;
    for (int j = 0; j < H; ++j) a += P[oW2 + j * H + k] * d2[j];
    dH1[s * H + k] = H1[s * H + k] > 0.f ? a : 0.f;
  }
}

// Parameter gradients summed over the minibatch (one thread per parameter)
// followed by the Adam update (Kingma & Ba 2015) with decoupled weight decay.
__global__ __launch_bounds__(256) void k_mlp_grad_adam(float* __restrict__ P, float* __restrict__ M1,
                                                       float* __restrict__ M2, const float* __restrict__ X, int b,
                                                       const float* __restrict__ H1, const float* __restrict__ H2,
                                                       const float* __restrict__ dZ, const float* __restrict__ dH1,
                                                       const float* __restrict__ dH2, float lr, float beta1,
                                                       float beta2, float bc1, float bc2, float wd) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  float g = 0.f;
  bool weight = true;
  if (p < ob1) {  // W1[j][k]: sum_s dH1[s][j] * X[s][k]
    const int j = (p - oW1) / F, k = (p - oW1) % F;
    for (int s = 0; s < b; ++s) g += dH1[s * H + j] * X[s * F + k];
  } else if (p < oW2) {
    const int j = p - ob1;
    for (int s = 0; s < b; ++s) g += dH1[s * H + j];
    weight = false;
  } else if (p < ob2) {
    const int j = (p - oW2) / H, k = (p - oW2) % H;
    for (int s = 0; s < b; ++s) g += dH2[s * H + j] * H1[s * H + k];
  } else if (p < oW3) {
    const int j = p - ob2;
    for (int s = 0; s < b; ++s) g += dH2[s * H + j];
    weight = false;
  } else if (p < ob3) {
    const int c = (p - oW3) / H, k = (p - oW3) % H;
    for (int s = 0; s < b; ++s) g += dZ[s * C + c] * H2[s * H + k];
  } else {
    const int c = p - ob3;
    for (int s = 0; s < b; ++s) g += dZ[s * C + c];
    weight = false;
  }
  const float m = beta1 * M1[p] + (1.f - beta1) * g;
  const float v = beta2 * M2[p] + (1.f - beta2) * g * g;
  M1[p]         = m;
  M2[p]         = v;
  float w       = P[p];
  if (weight) w -= lr * wd * w;
  P[p] = w - lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

// ------------------------------------------------------------------ host helpers
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) {
    if (mx_malloc(&p, bytes ? bytes : 4, nullptr) != 0) p = nullptr;
  }
  ~DevBuf() {
    if (p) (void)mx_free(p, nullptr);
  }
  float* f() const { return static_cast<float*>(p); }
};

thread_local std::string g_err;

int launch(const void* k, int grid, int block, void** args, mx_stream_t s) {
  return mx_launch(k, grid, 1, 1, block, 1, 1, 0, args, s);
}

int forward(const float* P, const float* X, int n, float* H1, float* H2, float* Z, mx_stream_t s) {
  void* a[] = {&P, &X, &n, &H1, &H2, &Z};
  return launch(reinterpret_cast<const void*>(&k_mlp_forward), (n + 255) / 256, 256, a, s);
}

// Xavier-uniform initial weights from a SplitMix64 stream, zero biases.
void init_params(std::vector<float>& p, uint64_t seed) {
  p.assign(NP, 0.f);
  uint64_t st = seed;
  auto next   = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * 0x1.0p-53;
  };
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double a = std::sqrt(6.0 / double(fan_in + fan_out));
    for (int i = 0; i < fan_out * fan_in; ++i) p[off + i] = float((2.0 * next() - 1.0) * a);
  };
  fill(oW1, H, F);
  fill(oW2, H, H);
  fill(oW3, C, H);
}

// ------------------------------------------------------------------ tool state
struct Model {
  bool loaded = false;
  std::vector<float> params;
  float mean[F] = {}, inv_std[F] = {};
};

struct Key {
  int64_t v[8];
  bool operator==(const Key& o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int64_t x : k.v) h = (h ^ uint64_t(x)) * 0x100000001B3ull;
    return size_t(h ^ (h >> 29));
  }
};
enum Field { kPattern, kExtent, kValueBytes, kFunctorBytes, kFloating, kBytes, kFlops, kCus, kKind };

struct ToolState {
  std::mutex mtx;
  Model model;
  // Declared input ids -> feature field (looked up per request, so no string
  // compares on the dispatch path).
  std::unordered_map<size_t, int> input_field;
  // range launch, team shape, MDRange tile, batch tile shape
  size_t var_id[4]  = {0, 0, 0, 0};
  bool have_var[4]   = {false, false, false, false};
  int ncand[4]       = {16, 16, 16, 16};  // candidates each variable declared
  std::unordered_map<Key, int, KeyHash> cache;
  DevBuf* dev_params = nullptr;
  mx_stream_t stream = nullptr;
  uint64_t requests = 0, answered = 0, inferences = 0;
  // The last request answered: variable (0-3), the runtime's default value
  // (the heuristic's candidate) and the value the model gave.
  int last_var = -1, last_default = -1, last_value = -1;
};
ToolState& tool() {
  static ToolState s;
  return s;
}

// Features of one request, from the runtime's declared inputs:
//  0 log2(extent)/32          1-3 pattern one-hot (for, reduce, scan)
//  4 log2(value bytes + 1)/4  5 log2(functor bytes + 1)/10
//  6 floating-point value     7 log2(bytes/index + 1)/6
//  8 log2(flops/index + 1)/8  9 log2(extent / (CUs*2048) + 2^-20)/20
// 10 log2(extent * bytes/index + 1)/40   11 policy kind / 2
struct Inputs {
  double extent = 1, value_bytes = 0, functor_bytes = 0, floating = 0, bytes = 0, flops = 0, cus = 256;
  int pattern   = 0;
  int kind      = 0;  // 0 RangePolicy launch shape, 1 Team shape, 2 MDRange tile, 3 batch tile shape
};
void features(const Inputs& in, float* f) {
  f[0] = float(std::log2(std::max(in.extent, 1.0)) / 32.0);
  f[1] = in.pattern == 0;
  f[2] = in.pattern == 1;
  f[3] = in.pattern == 2;
  f[4] = float(std::log2(in.value_bytes + 1.0) / 4.0);
  f[5] = float(std::log2(in.functor_bytes + 1.0) / 10.0);
  f[6] = float(in.floating);
  f[7] = float(std::log2(in.bytes + 1.0) / 6.0);
  f[8] = float(std::log2(in.flops + 1.0) / 8.0);
  f[9] = float(std::log2(in.extent / (std::max(in.cus, 1.0) * 2048.0) + 0x1.0p-20) / 20.0);
  f[10] = float(std::log2(in.extent * in.bytes + 1.0) / 40.0);
  f[11] = float(in.kind) * 0.5f;
}

// Arg-max over the first `ncand` logits (the candidates the variable has).
int predict_one(const float* feat, int ncand) {
  auto& s = tool();
  float x[F];
  for (int k = 0; k < F; ++k) x[k] = (feat[k] - s.model.mean[k]) * s.model.inv_std[k];
  if (!s.stream && mx_stream_create(&s.stream) != 0) return -1;
  if (!s.dev_params) {
    s.dev_params = new DevBuf(sizeof(float) * NP);
    if (!s.dev_params->p ||
        mx_memcpy(s.dev_params->p, s.model.params.data(), sizeof(float) * NP) != 0)
      return -1;
  }
  DevBuf dx(sizeof(float) * F), dz(sizeof(float) * C);
  if (!dx.p || !dz.p) return -1;
  if (mx_memcpy(dx.p, x, sizeof(x)) != 0) return -1;
  if (forward(s.dev_params->f(), dx.f(), 1, nullptr, nullptr, dz.f(), s.stream) != 0) return -1;
  float z[C];
  if (mx_stream_sync(s.stream) != 0 || mx_memcpy(z, dz.p, sizeof(z)) != 0) return -1;
  ++s.inferences;
  int best = 0;
  for (int c = 1; c < std::min(ncand, C); ++c)
    if (z[c] > z[best]) best = c;
  return best;
}

bool load_model_file(const std::string& path, Model& m) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  char magic[8];
  f.read(magic, 8);
  if (!f || std::memcmp(magic, "KKMXMLP1", 8) != 0) return false;
  int32_t dims[4];
  f.read(reinterpret_cast<char*>(dims), sizeof(dims));
  if (!f || dims[0] != F || dims[1] != H || dims[2] != H || dims[3] != C) return false;
  f.read(reinterpret_cast<char*>(m.mean), sizeof(m.mean));
  f.read(reinterpret_cast<char*>(m.inv_std), sizeof(m.inv_std));
  m.params.resize(NP);
  f.read(reinterpret_cast<char*>(m.params.data()), sizeof(float) * NP);
  if (!f) return false;
  m.loaded = true;
  return true;
}

}  // namespace

// ---------------------------------------------------------------- C-ABI (training / inference)
extern "C" {

const char* kkmx_tuner_last_error(void) { return g_err.c_str(); }
int kkmx_mlp_param_count(void) { return NP; }

// Trains the MLP on n samples (X: n x 12 features, already normalised; T:
// n x 16 target distributions) with Adam, minibatches of `batch` in an order
// drawn from `seed`.  params (NP floats) are the initial weights when
// init == 0, else they are initialised here; the trained weights are written
// back.  loss_per_epoch (epochs floats, may be NULL) receives the mean loss.
int kkmx_mlp_train(const float* X, const float* T, int n, int epochs, int batch, float lr, float weight_decay,
                   uint64_t seed, int init, float* params, float* loss_per_epoch) {
  if (n < 1 || epochs < 1 || batch < 1 || !X || !T || !params) {
    g_err = "kkmx_mlp_train: bad arguments";
    return -1;
  }
  if (batch > n) batch = n;
  std::vector<float> p(params, params + NP);
  if (init) init_params(p, seed);
  mx_stream_t s;
  if (mx_stream_create(&s) != 0) return 1;
  DevBuf dP(sizeof(float) * NP), dM1(sizeof(float) * NP), dM2(sizeof(float) * NP);
  DevBuf dXall(sizeof(float) * size_t(n) * F), dTall(sizeof(float) * size_t(n) * C);
  DevBuf dX(sizeof(float) * size_t(batch) * F), dT(sizeof(float) * size_t(batch) * C);
  DevBuf dH1(sizeof(float) * size_t(batch) * H), dH2(sizeof(float) * size_t(batch) * H);
  DevBuf dZ(sizeof(float) * size_t(batch) * C), ddZ(sizeof(float) * size_t(batch) * C);
  DevBuf ddH1(sizeof(float) * size_t(batch) * H), ddH2(sizeof(float) * size_t(batch) * H);
  DevBuf dLoss(sizeof(float));
  int rc = 0;
  auto chk = [&](int r) {
    if (r != 0 && rc == 0) rc = r;
    return r == 0;
  };
  chk(mx_memcpy(dP.p, p.data(), sizeof(float) * NP));
  chk(mx_memset_async(dM1.p, 0, sizeof(float) * NP, s));
  chk(mx_memset_async(dM2.p, 0, sizeof(float) * NP, s));
  // Shuffled copies of the data per epoch are gathered on the host (the data
  // set is a few hundred KB) and uploaded once per epoch.
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::vector<float> Xs(size_t(n) * F), Ts(size_t(n) * C);
  uint64_t st = seed ^ 0x5DEECE66Dull;
  auto rnd    = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const float beta1 = 0.9f, beta2 = 0.999f;
  int64_t step      = 0;
  for (int e = 0; e < epochs && rc == 0; ++e) {
    for (int i = n - 1; i > 0; --i) std::swap(order[i], order[int(rnd() % uint64_t(i + 1))]);
    for (int i = 0; i < n; ++i) {
      std::memcpy(&Xs[size_t(i) * F], X + size_t(order[i]) * F, sizeof(float) * F);
      std::memcpy(&Ts[size_t(i) * C], T + size_t(order[i]) * C, sizeof(float) * C);
    }
    chk(mx_memcpy(dXall.p, Xs.data(), Xs.size() * sizeof(float)));
    chk(mx_memcpy(dTall.p, Ts.data(), Ts.size() * sizeof(float)));
    chk(mx_memset_async(dLoss.p, 0, sizeof(float), s));
    int nb = 0;
    for (int b0 = 0; b0 + batch <= n && rc == 0; b0 += batch, ++nb) {
      ++step;
      const float* xb = dXall.f() + size_t(b0) * F;
      const float* tb = dTall.f() + size_t(b0) * C;
      float *pp = dP.f(), *h1 = dH1.f(), *h2 = dH2.f(), *z = dZ.f(), *dz = ddZ.f(), *dh1 = ddH1.f(),
            *dh2 = ddH2.f(), *ls = dLoss.f(), *m1 = dM1.f(), *m2 = dM2.f();
      int b = batch;
      chk(forward(pp, xb, b, h1, h2, z, s));
      void* a2[] = {&pp, &tb, &b, &h1, &h2, &z, &dz, &dh1, &dh2, &ls};
      chk(launch(reinterpret_cast<const void*>(&k_mlp_backward_act), (b + 255) / 256, 256, a2, s));
      float lr_ = lr, b1 = beta1, b2 = beta2, wd = weight_decay;
      float bc1 = 1.f - std::pow(beta1, float(step)), bc2 = 1.f - std::pow(beta2, float(step));
      void* a3[] = {&pp, &m1, &m2, &xb, &b, &h1, &h2, &dz, &dh1, &dh2, &lr_, &b1, &b2, &bc1, &bc2, &wd};
      chk(launch(reinterpret_cast<const void*>(&k_mlp_grad_adam), (NP + 255) / 256, 256, a3, s));
    }
    if (loss_per_epoch) {
      float l = 0;
      chk(mx_stream_sync(s));
      chk(mx_memcpy(&l, dLoss.p, sizeof(float)));
      loss_per_epoch[e] = nb ? l / float(nb) : 0.f;
    }
  }
  chk(mx_stream_sync(s));
  chk(mx_memcpy(params, dP.p, sizeof(float) * NP));
  (void)mx_stream_destroy(s);
  if (rc) g_err = "kkmx_mlp_train: HIP error";
  return rc;
}

// Logits of n normalised feature vectors (n x 12) -> Z (n x 16), on the GPU.
int kkmx_mlp_predict(const float* params, const float* X, int n, float* Z) {
  if (n < 1 || !params || !X || !Z) {
    g_err = "kkmx_mlp_predict: bad arguments";
    return -1;
  }
  mx_stream_t s;
  if (mx_stream_create(&s) != 0) return 1;
  DevBuf dP(sizeof(float) * NP), dX(sizeof(float) * size_t(n) * F), dZ(sizeof(float) * size_t(n) * C);
  int rc = mx_memcpy(dP.p, params, sizeof(float) * NP);
  if (!rc) rc = mx_memcpy(dX.p, X, sizeof(float) * size_t(n) * F);
  if (!rc) rc = forward(dP.f(), dX.f(), n, nullptr, nullptr, dZ.f(), s);
  if (!rc) rc = mx_stream_sync(s);
  if (!rc) rc = mx_memcpy(Z, dZ.p, sizeof(float) * size_t(n) * C);
  (void)mx_stream_destroy(s);
  return rc;
}

// The 12 features of a request, before normalisation (for the sweep driver).
void kkmx_tuner_features(int pattern, double extent, double value_bytes, double functor_bytes, int floating,
                         double bytes_per_index, double flops_per_index, int cus, int kind, float* out) {
  Inputs in;
  in.kind          = kind;
  in.pattern       = pattern;
  in.extent        = extent;
  in.value_bytes   = value_bytes;
  in.functor_bytes = functor_bytes;
  in.floating      = floating;
  in.bytes         = bytes_per_index;
  in.flops         = flops_per_index;
  in.cus           = cus;
  features(in, out);
}

// Loads a model file (magic, dims, feature mean, 1/std, parameters).
int kkmx_tuner_load_model(const char* path) {
  auto& s = tool();
  std::lock_guard<std::mutex> lk(s.mtx);
  Model m;
  if (!path || !load_model_file(path, m)) {
    g_err = std::string("kkmx_tuner_load_model: cannot read ") + (path ? path : "(null)");
    return -1;
  }
  s.model = m;
  s.cache.clear();
  delete s.dev_params;
  s.dev_params = nullptr;
  return 0;
}

// Requests seen, requests answered by the model, GPU inference launches.
void kkmx_tuner_stats(uint64_t* requests, uint64_t* answered, uint64_t* inferences) {
  auto& s = tool();
  if (requests) *requests = s.requests;
  if (answered) *answered = s.answered;
  if (inferences) *inferences = s.inferences;
}

// The last answered request: tuning variable (0 range launch, 1 team shape,
// 2 MDRange tile, 3 batch tile shape), the heuristic's candidate the runtime
// offered as default, and the candidate the model chose.  Returns 1 if no
// request has been answered yet.
int kkmx_tuner_last_answer(int* variable, int* heuristic, int* chosen) {
  auto& s = tool();
  std::lock_guard<std::mutex> lk(s.mtx);
  if (variable) *variable = s.last_var;
  if (heuristic) *heuristic = s.last_default;
  if (chosen) *chosen = s.last_value;
  return s.last_var < 0 ? 1 : 0;
}

// ---------------------------------------------------------------- Kokkos Tools callbacks
void kokkosp_init_library(const int, const uint64_t, const uint32_t, Kokkos_Profiling_KokkosPDeviceInfo*) {
  auto& s = tool();
  std::lock_guard<std::mutex> lk(s.mtx);
  if (s.model.loaded) return;
  const char* env = std::getenv("KKMX_TUNER_MODEL");
  if (env) load_model_file(env, s.model);
}
void kokkosp_finalize_library() {}
// Tuning answers need no device synchronisation around kernels.
void kokkosp_request_tool_settings(const uint32_t, Kokkos_Tools_ToolSettings* settings) {
  settings->requires_global_fencing = false;
}

void kokkosp_declare_input_type(const char* name, const size_t id, Kokkos_Tools_VariableInfo*) {
  static const std::pair<const char*, int> names[] = {
      {"kokkos.mi355x.pattern", kPattern},         {"kokkos.mi355x.extent", kExtent},
      {"kokkos.mi355x.value_bytes", kValueBytes},   {"kokkos.mi355x.functor_bytes", kFunctorBytes},
      {"kokkos.mi355x.floating", kFloating},        {"kokkos.mi355x.bytes_per_index", kBytes},
      {"kokkos.mi355x.flops_per_index", kFlops},    {"kokkos.mi355x.cu_count", kCus},
      {"kokkos.mi355x.policy_kind", kKind}};
  auto& s = tool();
  std::lock_guard<std::mutex> lk(s.mtx);
  for (const auto& nf : names)
    if (std::strcmp(name, nf.first) == 0) s.input_field[id] = nf.second;
}
void kokkosp_declare_output_type(const char* name, const size_t id, Kokkos_Tools_VariableInfo* info) {
  auto& s = tool();
  std::lock_guard<std::mutex> lk(s.mtx);
  const char* names[4] = {"kokkos.mi355x.range_launch", "kokkos.mi355x.team_shape", "kokkos.mi355x.mdrange_tile",
                          "kokkos.mi355x.batch_shape"};
  for (int v = 0; v < 4; ++v)
    if (std::strcmp(name, names[v]) == 0) {
      s.var_id[v]   = id;
      s.have_var[v] = true;
      if (info && info->valueQuantity == kokkos_value_set && info->candidates.set.size > 0)
        s.ncand[v] = int(std::min<size_t>(info->candidates.set.size, C));
    }
}
void kokkosp_begin_context(const size_t) {}
void kokkosp_end_context(const size_t, Kokkos_Tools_VariableValue) {}

void kokkosp_request_values(const size_t, const size_t n_in, const Kokkos_Tools_VariableValue* in,
                            const size_t n_out, Kokkos_Tools_VariableValue* out) {
  auto& s = tool();
  std::lock_guard<std::mutex> lk(s.mtx);
  for (size_t o = 0; o < n_out; ++o) {
    int var = -1;
    for (int v = 0; v < 4; ++v)
      if (s.have_var[v] && out[o].type_id == s.var_id[v]) var = v;
    if (var < 0) continue;
    ++s.requests;
    if (!s.model.loaded) continue;  // the runtime's heuristic value stands
    // A model trained on RangePolicy samples only has no spread in the
    // policy-kind feature; it answers RangePolicy requests alone.
    if (var != 0 && s.model.inv_std[11] == 0.f) continue;
    Inputs x;
    x.kind = var;
    for (size_t i = 0; i < n_in; ++i) {
      auto it = s.input_field.find(in[i].type_id);
      if (it == s.input_field.end()) continue;
      const auto& v = in[i].value;
      switch (it->second) {
        case kPattern: x.pattern = int(v.int_value); break;
        case kExtent: x.extent = double(v.int_value); break;
        case kValueBytes: x.value_bytes = double(v.int_value); break;
        case kFunctorBytes: x.functor_bytes = double(v.int_value); break;
        case kFloating: x.floating = double(v.int_value); break;
        case kBytes: x.bytes = v.double_value; break;
        case kFlops: x.flops = v.double_value; break;
        case kKind: break;  // the requested variable already tells the kind
        default: x.cus = double(v.int_value); break;
      }
    }
    // Cache key: everything but the extent, which is bucketed in quarter
    // octaves.
    const int64_t bucket = int64_t(std::floor(std::log2(std::max(x.extent, 1.0)) * 4.0));
    const Key key{{x.pattern + 4 * x.kind, int64_t(x.value_bytes), int64_t(x.functor_bytes), int64_t(x.floating),
                   int64_t(x.bytes * 16), int64_t(x.flops * 16), int64_t(x.cus), bucket}};
    int c;
    auto hit = s.cache.find(key);
    if (hit != s.cache.end()) {
      c = hit->second;
    } else {
      Inputs xb = x;
      xb.extent = std::exp2(double(bucket) / 4.0);
      float f[F];
      features(xb, f);
      c            = predict_one(f, s.ncand[var]);
      s.cache[key] = c;
    }
    if (c >= 0 && c < s.ncand[var]) {
      s.last_var             = var;
      s.last_default         = int(out[o].value.int_value);
      s.last_value           = c;
      out[o].value.int_value = c;
      ++s.answered;
    }
  }
}

}  // extern "C"

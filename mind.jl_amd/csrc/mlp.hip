// Training and inference kernels of the surrogate MLP (BF16 weights and
// activations as MFMA operands, FP32 accumulation, FP32 master weights and
// Adam moments). The network has no counterpart in the reference: it stands in
// for a micro population's input -> output map inside micro-macro coupling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/mind.h"
#include "mlp.h"

namespace mind {
int fail(int code, const std::string& msg);
}

namespace mind {
namespace {

#define MTRY(expr)                                                                          \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return ::mind::fail(MIND_EHIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

// C(i, j) [+]= sum_k A(i, k) B(k, j) with A(i, k) = A[i*sa0 + k*sa1] and
// B(k, j) = B[k*sb0 + j*sb1] in BF16, C row-major FP32 with row stride ldc.
// One wave per 16x16 output tile.
__global__ void __launch_bounds__(64) k_gemm(const uint16_t* A, int64_t sa0, int64_t sa1, const uint16_t* B,
                                             int64_t sb0, int64_t sb1, float* C, int64_t ldc, int M, int N,
                                             int K, int accumulate) {
  const int lane = threadIdx.x;
  const int i = blockIdx.x * 16 + (lane & 15);
  const int j = blockIdx.y * 16 + (lane & 15);
  const int kb = (lane >> 4) * 4;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + kb + q;
      a[q] = (short)((i < M && k < K) ? A[(int64_t)i * sa0 + (int64_t)k * sa1] : 0);
      b[q] = (short)((j < N && k < K) ? B[(int64_t)k * sb0 + (int64_t)j * sb1] : 0);
    }
    acc = mfma16(a, b, acc);
  }
  for (int q = 0; q < 4; ++q) {
    const int row = blockIdx.x * 16 + kb + q;
    const int col = blockIdx.y * 16 + (lane & 15);
    if (row < M && col < N) {
      float* c = C + (int64_t)row * ldc + col;
      *c = accumulate ? *c + acc[q] : acc[q];
    }
  }
}

__global__ void k_gather_batch(const float* X, const float* Y, const int32_t* idx, int B, int nin, int kp,
                               int nout, uint16_t* act0, float* yt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < B * kp) {
    const int i = t / kp, k = t % kp;
    act0[t] = k < nin ? f32_to_bf16(X[(int64_t)idx[i] * nin + k]) : 0;
  }
  if (t < B * nout) {
    const int i = t / nout, k = t % nout;
    yt[t] = Y[(int64_t)idx[i] * nout + k];
  }
}

__global__ void k_input_bf16(const float* x, int B, int nin, int kp, uint16_t* act0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kp) return;
  const int i = t / kp, k = t % kp;
  act0[t] = k < nin ? f32_to_bf16(x[(int64_t)i * nin + k]) : 0;
}

// z += b; act = gelu(z) (hidden) or z (output), padded columns zero.
__global__ void k_bias_act(float* z, const float* b, int B, int N, int np, uint16_t* act, int hidden) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * np) return;
  const int j = t % np;
  if (j >= N) {
    z[t] = 0.0f;
    if (act) act[t] = 0;
    return;
  }
  const float x = z[t] + b[j];
  z[t] = x;
  if (act) act[t] = f32_to_bf16(hidden ? gelu(x) : x);
}

__global__ void k_mse_grad(const float* out, const float* y, int B, int nout, int np, float* dz, float* loss) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * np) return;
  const int i = t / np, j = t % np;
  if (j >= nout) {
    dz[t] = 0.0f;
    return;
  }
  const float e = out[t] - y[i * nout + j];
  dz[t] = 2.0f * e / (float)(B * nout);
  atomicAdd(loss, e * e / (float)(B * nout));
}

__global__ void k_to_bf16(const float* x, int64_t n, uint16_t* y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = f32_to_bf16(x[t]);
}

// Bias gradient: column sums. A workgroup of kColSlices x 64 threads takes 64
// columns; slice q sums rows q, q + kColSlices, ... (coalesced across the
// columns), then the slices' partials are added in slice order: a fixed order,
// so the result is deterministic.
constexpr int kColSlices = 16;
__global__ void __launch_bounds__(64 * kColSlices) k_colsum(const float* dz, int B, int N, int np, float* gb) {
  __shared__ float part[kColSlices][64];
  const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + c;
  float s = 0.0f;
  if (j < N)
    for (int i = q; i < B; i += kColSlices) s += dz[(int64_t)i * np + j];
  part[q][c] = s;
  __syncthreads();
  if (q == 0 && j < N) {
    float t = part[0][c];
    for (int k = 1; k < kColSlices; ++k) t += part[k][c];
    gb[j] = t;
  }
}

// dz = dh * gelu'(z) on the real columns of a hidden layer.
__global__ void k_gelu_back(const float* dh, const float* z, int B, int N, int np, float* dz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * np) return;
  const int j = t % np;
  dz[t] = j < N ? dh[t] * gelu_grad(z[t]) : 0.0f;
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// Padded BF16 copy of one layer's weights from the flat FP32 parameters.
__global__ void k_pack_weights(const float* w, int K, int N, int kp, int np, uint16_t* wb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kp * np) return;
  const int k = t / np, j = t % np;
  wb[t] = (k < K && j < N) ? f32_to_bf16(w[(int64_t)k * N + j]) : 0;
}

// y (rows x n_out, row-major) = the first n_out columns of z (row stride np) + b.
__global__ void k_bias_out(const float* z, const float* b, int B, int N, int np, float* y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * N) return;
  const int i = t / N, j = t % N;
  y[t] = z[(int64_t)i * np + j] + b[j];
}

inline unsigned nb(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

// Scratch memory of mind_mlp_forward_noalloc: two BF16 activation buffers and one
// FP32 pre-activation buffer of kScratchRows rows each at the widest padded
// layer. One array per GPU for the life of the process; every call on a GPU
// shares it, so calls that overlap on one GPU race.
constexpr int64_t kScratchBytes = int64_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_dev[64] = {};  // device address of `scratch` on each GPU

// Device address of this GPU's `scratch` (looked up once per device, with that
// device current).
int scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= 64) return fail(MIND_EARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MIND_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_dev[device]);
  return 0;
}

}  // namespace

struct Mlp {
  MlpShape sh{};
  int device = 0;  // current device at creation: where every buffer lives
  int64_t max_batch = 0, nparam = 0;
  std::vector<int64_t> w_off, b_off;
  float *master = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  std::vector<uint16_t*> wbf;     // per layer [kp x np]
  std::vector<uint16_t*> act;     // act[0] input, act[l+1] output of hidden layer l, [B x np]
  std::vector<float*> z;          // pre-activations [B x np]
  float *dz = nullptr, *dh = nullptr, *yt = nullptr, *loss = nullptr;
  uint16_t* dzb = nullptr;
  float *X = nullptr, *Y = nullptr, *loss_steps = nullptr;  // training set, owned outside allocs
  int32_t* perm = nullptr;
  int64_t n_data = 0, adam_t = 0;
  std::vector<void*> allocs;
  hipStream_t stream = nullptr;

  template <typename T>
  int alloc(T** p, int64_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<int64_t>(1, count) * sizeof(T)) != hipSuccess)
      return fail(MIND_EOOM, "device allocation failed");
    allocs.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
  }
  ~Mlp() {
    for (void* p : allocs) (void)hipFree(p);
    for (void* p : {(void*)X, (void*)Y, (void*)loss_steps, (void*)perm})
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int kp(int l) const { return MlpShape::pad(sh.k(l)); }
  int np(int l) const { return MlpShape::pad(sh.n(l)); }
};

namespace {

int pack_all(Mlp& M) {
  for (int l = 0; l < M.sh.nlin(); ++l)
    hipLaunchKernelGGL(k_pack_weights, dim3(nb((int64_t)M.kp(l) * M.np(l))), dim3(256), 0, M.stream,
                       M.master + M.w_off[l], M.sh.k(l), M.sh.n(l), M.kp(l), M.np(l), M.wbf[l]);
  MTRY(hipGetLastError());
  return 0;
}

void gemm(Mlp& M, const uint16_t* A, int64_t sa0, int64_t sa1, const uint16_t* B, int64_t sb0, int64_t sb1,
          float* C, int64_t ldc, int rows, int cols, int K, int accumulate) {
  hipLaunchKernelGGL(k_gemm, dim3((rows + 15) / 16, (cols + 15) / 16), dim3(64), 0, M.stream, A, sa0, sa1, B,
                     sb0, sb1, C, ldc, rows, cols, K, accumulate);
}

// Forward pass over B rows of act[0]; leaves z[l] and act[l] for backward.
void forward(Mlp& M, int B) {
  const int L = M.sh.nlin();
  for (int l = 0; l < L; ++l) {
    const int kp = M.kp(l), np = M.np(l);
    gemm(M, M.act[l], kp, 1, M.wbf[l], np, 1, M.z[l], np, B, np, kp, 0);
    const bool hidden = l < L - 1;
    hipLaunchKernelGGL(k_bias_act, dim3(nb((int64_t)B * np)), dim3(256), 0, M.stream, M.z[l],
                       M.master + M.b_off[l], B, M.sh.n(l), np, hidden ? M.act[l + 1] : nullptr, hidden ? 1 : 0);
  }
}

// Backward pass from the MSE loss against yt; writes the flat gradient.
void backward(Mlp& M, int B) {
  const int L = M.sh.nlin();
  hipMemsetAsync(M.loss, 0, sizeof(float), M.stream);
  hipLaunchKernelGGL(k_mse_grad, dim3(nb((int64_t)B * M.np(L - 1))), dim3(256), 0, M.stream, M.z[L - 1], M.yt, B,
                     M.sh.n_out, M.np(L - 1), M.dz, M.loss);
  for (int l = L - 1; l >= 0; --l) {
    const int kp = M.kp(l), np = M.np(l), K = M.sh.k(l), N = M.sh.n(l);
    hipLaunchKernelGGL(k_to_bf16, dim3(nb((int64_t)B * np)), dim3(256), 0, M.stream, M.dz, (int64_t)B * np, M.dzb);
    // dW = act^T dz  (K x N, written unpadded into the flat gradient)
    gemm(M, M.act[l], 1, kp, M.dzb, np, 1, M.grad + M.w_off[l], N, K, N, B, 0);
    hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64), dim3(64 * kColSlices), 0, M.stream, M.dz, B, N, np,
                       M.grad + M.b_off[l]);
    if (l == 0) break;
    // dh = dz W^T  (B x kp), then through the previous layer's GELU
    gemm(M, M.dzb, np, 1, M.wbf[l], 1, np, M.dh, kp, B, kp, np, 0);
    hipLaunchKernelGGL(k_gelu_back, dim3(nb((int64_t)B * kp)), dim3(256), 0, M.stream, M.dh, M.z[l - 1], B,
                       M.sh.n(l - 1), kp, M.dz);
  }
}

void adam(Mlp& M, float lr) {
  M.adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)M.adam_t), c2 = 1.0f - powf(b2, (float)M.adam_t);
  hipLaunchKernelGGL(k_adam, dim3(nb(M.nparam)), dim3(256), 0, M.stream, M.master, M.grad, M.m, M.v, M.nparam, lr,
                     b1, b2, 1e-8f, c1, c2);
}

}  // namespace
}  // namespace mind

using mind::Mlp;

extern "C" {

int mind_mlp_create(int32_t n_in, int32_t hidden, int32_t layers, int32_t n_out, int64_t max_batch,
                    mind_mlp** out) {
  if (n_in <= 0 || hidden <= 0 || layers <= 0 || n_out <= 0 || max_batch <= 0 || !out)
    return mind::fail(MIND_EARG, "invalid MLP shape");
  Mlp* M = new Mlp();
  M->sh = mind::MlpShape{n_in, hidden, layers, n_out};
  if (hipGetDevice(&M->device) != hipSuccess) {
    delete M;
    return mind::fail(MIND_EHIP, "no current device");
  }
  M->max_batch = (max_batch + 15) / 16 * 16;
  int rc = 0;
  if (hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess) {
    delete M;
    return mind::fail(MIND_EHIP, "stream creation failed");
  }
  int64_t off = 0, maxnp = 0;
  for (int l = 0; l < M->sh.nlin(); ++l) {
    M->w_off.push_back(off);
    off += (int64_t)M->sh.k(l) * M->sh.n(l);
    M->b_off.push_back(off);
    off += M->sh.n(l);
    maxnp = std::max<int64_t>(maxnp, std::max(M->kp(l), M->np(l)));
  }
  M->nparam = off;
  const int64_t B = M->max_batch;
  if ((rc = M->alloc(&M->master, off)) || (rc = M->alloc(&M->grad, off)) || (rc = M->alloc(&M->m, off)) ||
      (rc = M->alloc(&M->v, off))) {
    delete M;
    return rc;
  }
  M->wbf.resize(M->sh.nlin());
  M->z.resize(M->sh.nlin());
  M->act.resize(M->sh.nlin());
  for (int l = 0; l < M->sh.nlin(); ++l) {
    if ((rc = M->alloc(&M->wbf[l], (int64_t)M->kp(l) * M->np(l))) || (rc = M->alloc(&M->z[l], B * M->np(l))) ||
        (rc = M->alloc(&M->act[l], B * M->kp(l)))) {
      delete M;
      return rc;
    }
  }
  if ((rc = M->alloc(&M->dz, B * maxnp)) || (rc = M->alloc(&M->dh, B * maxnp)) ||
      (rc = M->alloc(&M->dzb, B * maxnp)) || (rc = M->alloc(&M->yt, B * n_out)) || (rc = M->alloc(&M->loss, 1))) {
    delete M;
    return rc;
  }
  (void)hipMemset(M->master, 0, sizeof(float) * off);
  (void)hipMemset(M->m, 0, sizeof(float) * off);
  (void)hipMemset(M->v, 0, sizeof(float) * off);
  *out = reinterpret_cast<mind_mlp*>(M);
  return 0;
}

void mind_mlp_destroy(mind_mlp* mlp) { delete reinterpret_cast<Mlp*>(mlp); }

int mind_mlp_param_count(const mind_mlp* mlp, int64_t* n) {
  *n = reinterpret_cast<const Mlp*>(mlp)->nparam;
  return 0;
}

int mind_mlp_set_params(mind_mlp* mlp, const float* params, int64_t n) {
  Mlp& M = *reinterpret_cast<Mlp*>(mlp);
  if (n != M.nparam) return mind::fail(MIND_EARG, "parameter count mismatch");
  MTRY(hipMemcpy(M.master, params, sizeof(float) * n, hipMemcpyHostToDevice));
  MTRY(hipMemset(M.m, 0, sizeof(float) * n));
  MTRY(hipMemset(M.v, 0, sizeof(float) * n));
  M.adam_t = 0;
  int rc = mind::pack_all(M);
  if (rc) return rc;
  MTRY(hipStreamSynchronize(M.stream));
  return 0;
}

int mind_mlp_get_params(mind_mlp* mlp, float* params, int64_t n) {
  Mlp& M = *reinterpret_cast<Mlp*>(mlp);
  if (n != M.nparam) return mind::fail(MIND_EARG, "parameter count mismatch");
  MTRY(hipStreamSynchronize(M.stream));
  MTRY(hipMemcpy(params, M.master, sizeof(float) * n, hipMemcpyDeviceToHost));
  return 0;
}

// Forward pass of `batch` rows of x (row-major, n_in floats each) into y (n_out each).
int mind_mlp_forward(mind_mlp* mlp, const float* x, int64_t batch, float* y) {
  Mlp& M = *reinterpret_cast<Mlp*>(mlp);
  if (batch <= 0 || batch > M.max_batch) return mind::fail(MIND_EARG, "batch exceeds the MLP's capacity");
  const int B = (int)batch, L = M.sh.nlin();
  float* dx = M.dh;  // staging: B x n_in floats fit in B x maxnp
  MTRY(hipMemcpyAsync(dx, x, sizeof(float) * B * M.sh.n_in, hipMemcpyHostToDevice, M.stream));
  hipLaunchKernelGGL(mind::k_input_bf16, dim3(mind::nb((int64_t)B * M.kp(0))), dim3(256), 0, M.stream, dx, B,
                     M.sh.n_in, M.kp(0), M.act[0]);
  mind::forward(M, B);
  MTRY(hipGetLastError());
  std::vector<float> out((size_t)B * M.np(L - 1));
  MTRY(hipMemcpyAsync(out.data(), M.z[L - 1], sizeof(float) * out.size(), hipMemcpyDeviceToHost, M.stream));
  MTRY(hipStreamSynchronize(M.stream));
  for (int i = 0; i < B; ++i)
    for (int j = 0; j < M.sh.n_out; ++j) y[(int64_t)i * M.sh.n_out + j] = out[(size_t)i * M.np(L - 1) + j];
  return 0;
}

// mind_mlp_forward without allocation: x (batch x n_in) and y (batch x n_out)
// are device pointers on the MLP's device, the activations live in the
// library's static scratch array. With stream NULL the call runs on the MLP's
// stream and returns when y is written; otherwise it is only enqueued on
// `stream` (capturable). No hipMalloc, hipFree or host allocation.
int mind_mlp_forward_noalloc(const mind_mlp* mlp, const float* x, int64_t batch, float* y, void* stream) {
  if (!mlp || !x || !y) return mind::fail(MIND_EARG, "null argument");
  const Mlp& M = *reinterpret_cast<const Mlp*>(mlp);
  const int L = M.sh.nlin();
  int64_t width = 0;  // widest padded layer edge
  for (int l = 0; l < L; ++l) width = std::max<int64_t>(width, std::max(M.kp(l), M.np(l)));
  const int64_t max_rows = mind::kScratchBytes / (width * (2 * (int64_t)sizeof(uint16_t) + (int64_t)sizeof(float)));
  if (batch <= 0) return mind::fail(MIND_EARG, "batch must be positive");
  if (batch > max_rows)
    return mind::fail(MIND_EARG, "batch of " + std::to_string(batch) + " rows exceeds the scratch limit of " +
                                     std::to_string(max_rows) + " rows for this MLP");
  int cur = -1;
  MTRY(hipGetDevice(&cur));
  if (cur != M.device) MTRY(hipSetDevice(M.device));
  unsigned char* base = nullptr;
  int rc = mind::scratch_address(M.device, &base);
  if (rc) return rc;
  const int B = (int)batch;
  uint16_t* act[2] = {reinterpret_cast<uint16_t*>(base),
                      reinterpret_cast<uint16_t*>(base + (size_t)B * width * sizeof(uint16_t))};
  float* z = reinterpret_cast<float*>(base + 2 * (size_t)B * width * sizeof(uint16_t));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : M.stream;
  hipLaunchKernelGGL(mind::k_input_bf16, dim3(mind::nb((int64_t)B * M.kp(0))), dim3(256), 0, st, x, B, M.sh.n_in,
                     M.kp(0), act[0]);
  for (int l = 0; l < L; ++l) {
    const int kp = M.kp(l), np = M.np(l);
    hipLaunchKernelGGL(mind::k_gemm, dim3((B + 15) / 16, (np + 15) / 16), dim3(64), 0, st, act[l & 1], (int64_t)kp,
                       (int64_t)1, M.wbf[l], (int64_t)np, (int64_t)1, z, (int64_t)np, B, np, kp, 0);
    if (l < L - 1)
      hipLaunchKernelGGL(mind::k_bias_act, dim3(mind::nb((int64_t)B * np)), dim3(256), 0, st, z,
                         M.master + M.b_off[l], B, M.sh.n(l), np, act[(l + 1) & 1], 1);
    else
      hipLaunchKernelGGL(mind::k_bias_out, dim3(mind::nb((int64_t)B * M.sh.n_out)), dim3(256), 0, st, z,
                         M.master + M.b_off[l], B, M.sh.n_out, np, y);
  }
  MTRY(hipGetLastError());
  if (!stream) MTRY(hipStreamSynchronize(st));
  if (cur != M.device) MTRY(hipSetDevice(cur));
  return 0;
}

// Loss and flat gradient of the MSE on one batch (gradient check entry point).
int mind_mlp_grad(mind_mlp* mlp, const float* x, const float* y, int64_t batch, float* loss, float* grads) {
  Mlp& M = *reinterpret_cast<Mlp*>(mlp);
  if (batch <= 0 || batch > M.max_batch) return mind::fail(MIND_EARG, "batch exceeds the MLP's capacity");
  const int B = (int)batch;
  MTRY(hipMemcpyAsync(M.dh, x, sizeof(float) * B * M.sh.n_in, hipMemcpyHostToDevice, M.stream));
  MTRY(hipMemcpyAsync(M.yt, y, sizeof(float) * B * M.sh.n_out, hipMemcpyHostToDevice, M.stream));
  hipLaunchKernelGGL(mind::k_input_bf16, dim3(mind::nb((int64_t)B * M.kp(0))), dim3(256), 0, M.stream, M.dh, B,
                     M.sh.n_in, M.kp(0), M.act[0]);
  mind::forward(M, B);
  mind::backward(M, B);
  MTRY(hipGetLastError());
  MTRY(hipMemcpyAsync(loss, M.loss, sizeof(float), hipMemcpyDeviceToHost, M.stream));
  MTRY(hipMemcpyAsync(grads, M.grad, sizeof(float) * M.nparam, hipMemcpyDeviceToHost, M.stream));
  MTRY(hipStreamSynchronize(M.stream));
  return 0;
}

// Minibatch Adam on a dataset held on the device. The row order of each epoch is
// a seeded shuffle; epoch_loss receives the mean batch loss of every epoch.
int mind_mlp_train(mind_mlp* mlp, const float* X, const float* Y, int64_t n, int64_t batch, int32_t epochs,
                   float lr, uint64_t seed, float* epoch_loss) {
  Mlp& M = *reinterpret_cast<Mlp*>(mlp);
  if (batch <= 0 || batch > M.max_batch || n < batch) return mind::fail(MIND_EARG, "invalid batch or data size");
  const int64_t steps = n / batch;
  if (n != M.n_data) {
    if (M.X) (void)hipFree(M.X);
    if (M.Y) (void)hipFree(M.Y);
    if (M.perm) (void)hipFree(M.perm);
    if (M.loss_steps) (void)hipFree(M.loss_steps);
    M.X = M.Y = M.loss_steps = nullptr;
    M.perm = nullptr;
    M.n_data = 0;
    if (hipMalloc(&M.X, sizeof(float) * n * M.sh.n_in) != hipSuccess ||
        hipMalloc(&M.Y, sizeof(float) * n * M.sh.n_out) != hipSuccess ||
        hipMalloc(&M.perm, sizeof(int32_t) * n) != hipSuccess ||
        hipMalloc(&M.loss_steps, sizeof(float) * steps) != hipSuccess)
      return mind::fail(MIND_EOOM, "device allocation failed");
    M.n_data = n;
  }
  MTRY(hipMemcpy(M.X, X, sizeof(float) * n * M.sh.n_in, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(M.Y, Y, sizeof(float) * n * M.sh.n_out, hipMemcpyHostToDevice));
  std::vector<int32_t> perm(n);
  std::mt19937_64 rng(seed);
  const int B = (int)batch;
  std::vector<float> losses(steps);
  for (int e = 0; e < epochs; ++e) {
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    std::shuffle(perm.begin(), perm.end(), rng);
    MTRY(hipMemcpyAsync(M.perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, M.stream));
    for (int64_t s = 0; s < steps; ++s) {
      hipLaunchKernelGGL(mind::k_gather_batch, dim3(mind::nb((int64_t)B * std::max(M.kp(0), M.sh.n_out))), dim3(256),
                         0, M.stream, M.X, M.Y, M.perm + s * B, B, M.sh.n_in, M.kp(0), M.sh.n_out, M.act[0], M.yt);
      mind::forward(M, B);
      mind::backward(M, B);
      mind::adam(M, lr);
      int rc = mind::pack_all(M);
      if (rc) return rc;
      MTRY(hipMemcpyAsync(M.loss_steps + s, M.loss, sizeof(float), hipMemcpyDeviceToDevice, M.stream));
    }
    MTRY(hipMemcpyAsync(losses.data(), M.loss_steps, sizeof(float) * steps, hipMemcpyDeviceToHost, M.stream));
    MTRY(hipStreamSynchronize(M.stream));
    double acc = 0.0;
    for (float x : losses) acc += x;
    if (epoch_loss) epoch_loss[e] = (float)(acc / (double)steps);
  }
  return 0;
}

}  // extern "C"

// Micro-macro coupling with a learned surrogate in place of the micro
// populations. The whole run is one persistent single-workgroup kernel: per
// macro step it forms the delayed coupling (as k_macro_pre), accumulates each
// ROI's drive, runs the MLP for all ROIs at bin ends on the matrix cores, and
// advances Wilson-Cowan by RK4 (as k_macro_post). The surrogate has no
// counterpart in the reference; the coupling and REGION arithmetic are the
// same operations as the plain path's (src/runtime/step.jl:156-204).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mind.h"
#include "macro.h"
#include "mlp.h"
#include "state.h"

namespace mind {
namespace {

constexpr int kThreads = 1024;  // 16 waves
constexpr int kWaves = kThreads / 64;
constexpr int kMaxWidth = 512;  // padded n_in and hidden
constexpr int kMaxLayers = 8;
constexpr int kMaxIn = 256;  // network inputs
constexpr int64_t kStepsPerLaunch = 1 << 16;

struct SurDev {
  int32_t n, history, layers, n_out;  // layers = number of linear layers
  double dt_macro, v_rest;
  int64_t cap, bin_steps;
  const double* W;
  const int64_t* dsteps;
  double* hist;      // [slot * n + roi]
  const double* wc;  // [roi * WC_NPARAMS + k]
  double *E, *I, *coupling, *rate, *meanv, *acc;
  double* drive;     // [roi * history + k], ring
  int64_t* clock;    // [0] macro step, [1] bins so far, [2] OOD bins, [3] step of the first OOD bin (-1)
  // network: per linear layer padded K and N (multiples of 16), packed BF16 MFMA
  // B fragments [tile][kstep][lane][4] and padded FP32 bias
  int32_t kp[kMaxLayers], np[kMaxLayers];
  const bf16x4* frag[kMaxLayers];
  const float* bias[kMaxLayers];
  const float *x_mean, *x_std, *y_mean, *y_std, *x_lo, *x_hi;
  const float *skip_w, *skip_b;  // optional linear skip [n_in][n_out], [n_out]
  int32_t n_in;
  double* trace;  // [5][n][samples]
  int64_t samples, trace_start;
};

__device__ inline void record(const SurDev& s, int r, int64_t sample, double e, double i, double rate,
                              double c, double v) {
  if (!s.trace) return;
  const int64_t S = s.samples;
  s.trace[(0 * (int64_t)s.n + r) * S + sample] = e;
  s.trace[(1 * (int64_t)s.n + r) * S + sample] = i;
  s.trace[(2 * (int64_t)s.n + r) * S + sample] = rate;
  s.trace[(3 * (int64_t)s.n + r) * S + sample] = c;
  s.trace[(4 * (int64_t)s.n + r) * S + sample] = v;
}

// One 16-row forward pass: act[0] holds the normalised BF16 inputs; the result of
// the last layer lands in out[16][16] (FP32, before de-normalisation).
__device__ void mlp_rows(const SurDev& s, uint16_t (*act)[16][kMaxWidth + 4], float (*out)[16]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kq = (lane >> 4) * 4;
  for (int l = 0; l < s.layers; ++l) {
    const uint16_t(*in)[kMaxWidth + 4] = act[l & 1];
    uint16_t(*nxt)[kMaxWidth + 4] = act[(l + 1) & 1];
    const int ks_n = s.kp[l] / 16, tiles = s.np[l] / 16;
    const bool last = l == s.layers - 1;
    for (int t = wave; t < tiles; t += kWaves) {
      f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      const bf16x4* fr = s.frag[l] + (int64_t)t * ks_n * 64 + lane;
      for (int ks = 0; ks < ks_n; ++ks) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(&in[row][ks * 16 + kq]);
        c = mfma16(a, fr[(int64_t)ks * 64], c);
      }
      const int col = t * 16 + row;
      const float b = s.bias[l][col];
      for (int q = 0; q < 4; ++q) {
        const float x = c[q] + b;
        if (last) {
          out[kq + q][col] = x;
        } else {
          nxt[kq + q][col] = f32_to_bf16(gelu(x));
        }
      }
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kThreads) k_surrogate_run(SurDev s, int64_t count) {
  __shared__ uint16_t act[2][16][kMaxWidth + 4];
  __shared__ float out[16][16];
  __shared__ float xin[16][kMaxIn];
  __shared__ int ood_row[16];
  const int tid = threadIdx.x;
  const int64_t m0 = s.clock[0];
  int64_t nbins = s.clock[1];
  const int n = s.n, H = s.history;
  for (int64_t j = 0; j < count; ++j) {
    const int64_t ms = m0 + j;
    const int64_t slot = ms % s.cap;
    // delayed coupling, sources in increasing order; drive accumulation
    for (int i = tid; i < n; i += kThreads) {
      double c = 0.0;
      for (int q = 0; q < n; ++q) {
        const double w = s.W[(int64_t)i * n + q];
        if (w == 0.0) continue;
        const int64_t sl = ((ms - s.dsteps[(int64_t)i * n + q]) % s.cap + s.cap) % s.cap;
        c = c + edge_value(w, s.hist[sl * n + q]);
      }
      s.coupling[i] = c;
      s.acc[i] = s.acc[i] + drive_rate(s.wc[(int64_t)i * WC_NPARAMS + WC_RATE_SCALE], s.hist[slot * n + i]);
    }
    if ((ms + 1) % s.bin_steps == 0) {
      for (int i = tid; i < n; i += kThreads) {
        const double d = s.acc[i] / (double)s.bin_steps;
        s.acc[i] = 0.0;
        if (nbins == 0)
          for (int k = 0; k < H; ++k) s.drive[(int64_t)i * H + k] = d;
        else
          s.drive[(int64_t)i * H + nbins % H] = d;
      }
      __syncthreads();
      for (int r0 = 0; r0 < n; r0 += 16) {
        if (tid < 16) ood_row[tid] = 0;
        __syncthreads();
        // inputs: 16 rows x kp[0] columns, oldest drive bin first
        for (int e = tid; e < 16 * s.kp[0]; e += kThreads) {
          const int row = e / s.kp[0], k = e % s.kp[0], r = r0 + row;
          float x = 0.0f;
          bool valid = r < n && k < s.n_in;
          if (valid) {
            double raw;
            if (k < H)
              raw = s.drive[(int64_t)r * H + (nbins + 1 + k) % H];
            else if (k == H)
              raw = s.rate[r];
            else
              raw = s.meanv[r];
            const float f = (float)raw;
            if (!(f >= s.x_lo[k] && f <= s.x_hi[k])) ood_row[row] = 1;
            x = (f - s.x_mean[k]) / s.x_std[k];
            xin[row][k] = x;
          }
          act[0][row][k] = valid ? f32_to_bf16(x) : (uint16_t)0;
        }
        __syncthreads();
        mlp_rows(s, act, out);
        if (tid < 16 && r0 + tid < n) {
          const int r = r0 + tid;
          float o0 = out[tid][0], o1 = out[tid][1];
          if (s.skip_w) {
            float k0 = 0.0f, k1 = 0.0f;
            for (int k = 0; k < s.n_in; ++k) {
              k0 = k0 + xin[tid][k] * s.skip_w[k * 2 + 0];
              k1 = k1 + xin[tid][k] * s.skip_w[k * 2 + 1];
            }
            o0 = o0 + (k0 + s.skip_b[0]);
            o1 = o1 + (k1 + s.skip_b[1]);
          }
          const float y0 = o0 * s.y_std[0] + s.y_mean[0];
          const float y1 = o1 * s.y_std[1] + s.y_mean[1];
          s.rate[r] = y0 > 0.0f ? (double)y0 : 0.0;
          s.meanv[r] = (double)y1;
          if (ood_row[tid]) {
            atomicAdd((unsigned long long*)&s.clock[2], 1ull);
            atomicCAS((unsigned long long*)&s.clock[3], (unsigned long long)-1ll, (unsigned long long)ms);
          }
        }
        __syncthreads();
      }
      ++nbins;
    }
    for (int i = tid; i < n; i += kThreads) {
      const double* p = s.wc + (int64_t)i * WC_NPARAMS;
      double e = s.E[i], ii = s.I[i];
      region_advance(p, e, ii, s.coupling[i], s.rate[i], s.dt_macro);
      s.E[i] = e;
      s.I[i] = ii;
      s.hist[((ms + 1) % s.cap) * n + i] = e;
      record(s, i, ms + 1 - s.trace_start, e, ii, s.rate[i], s.coupling[i], s.meanv[i]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    s.clock[0] = m0 + count;
    s.clock[1] = nbins;
  }
}

__global__ void k_surrogate_init(SurDev s, const double* e0, const double* i0) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) {
    s.clock[0] = 0;
    s.clock[1] = 0;
    s.clock[2] = 0;
    s.clock[3] = -1;
  }
  if (r >= s.n) return;
  s.E[r] = e0[r];
  s.I[r] = i0[r];
  s.coupling[r] = 0.0;
  s.rate[r] = 0.0;
  s.acc[r] = 0.0;
  s.meanv[r] = s.v_rest;
  for (int64_t k = 0; k < s.cap; ++k) s.hist[k * s.n + r] = e0[r];
  for (int k = 0; k < s.history; ++k) s.drive[(int64_t)r * s.history + k] = 0.0;
}

__global__ void k_surrogate_record(SurDev s, int64_t sample) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < s.n) record(s, r, sample, s.E[r], s.I[r], s.rate[r], s.coupling[r], s.meanv[r]);
}

}  // namespace

struct Surrogate {
  SurDev d{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  double *e0 = nullptr, *i0 = nullptr;
  bool initialized = false;

  template <typename T>
  int upload(T** out, const T* host, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)) != hipSuccess)
      return fail(MIND_EOOM, "device allocation failed");
    allocs.push_back(p);
    if (host && count && hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      return fail(MIND_EHIP, "surrogate upload failed");
    *out = static_cast<T*>(p);
    return 0;
  }
  ~Surrogate() {
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace mind

using mind::Surrogate;

extern "C" {

int mind_build_surrogate(const mind_macro_desc* md, const mind_surrogate_desc* nd, int32_t device,
                         mind_surrogate** out) {
  using namespace mind;
  if (!md || !nd || !out || !nd->params) return fail(MIND_EARG, "null argument");
  const int n = md->n_roi;
  if (n <= 0) return fail(MIND_EARG, "the macro model needs at least one ROI");
  if (nd->history < 1 || nd->n_in != nd->history + 2 || nd->n_out != 2 || nd->layers < 1 ||
      nd->layers + 1 > kMaxLayers || nd->hidden < 1 || MlpShape::pad(nd->hidden) > kMaxWidth ||
      MlpShape::pad(nd->n_in) > kMaxWidth || nd->n_in > kMaxIn || !nd->skip_w != !nd->skip_b)
    return fail(MIND_EARG, "unsupported surrogate network shape");
  const MlpShape sh{nd->n_in, nd->hidden, nd->layers, nd->n_out};
  int64_t nparam = 0;
  for (int l = 0; l < sh.nlin(); ++l) nparam += (int64_t)sh.k(l) * sh.n(l) + sh.n(l);
  if (nd->n_params != nparam) return fail(MIND_EARG, "surrogate parameter count mismatch");
  const std::vector<double> W(md->weights, md->weights + (size_t)n * n);
  const std::vector<double> wc(md->wc_params, md->wc_params + (size_t)n * WC_NPARAMS);
  std::vector<int64_t> ds((size_t)n * n);
  int64_t maxd = 0;
  for (size_t q = 0; q < ds.size(); ++q) {
    if (!std::isfinite(md->delays[q]) || md->delays[q] < 0) return fail(MIND_EARG, "invalid macro delay");
    ds[q] = (int64_t)std::nearbyint(md->delays[q] / md->dt_macro);
    if (W[q] != 0.0) maxd = std::max(maxd, ds[q]);
  }
  int64_t bin = -1;
  for (int r = 0; r < n; ++r) {
    const int64_t b = std::max<int64_t>(1, (int64_t)std::nearbyint(md->bin_ms[r] / md->dt_macro));
    if (bin >= 0 && b != bin) return fail(MIND_EUNSUPPORTED, "the surrogate needs one bin width for all ROIs");
    bin = b;
  }
  if (!(md->dt_macro > 0.0)) return fail(MIND_EARG, "dt_macro must be positive");
  for (int k = 0; k < nd->n_in; ++k)
    if (!(nd->x_std[k] > 0.0f)) return fail(MIND_EARG, "input scales must be positive");
  if (hipSetDevice(device) != hipSuccess) return fail(MIND_EHIP, "invalid device");
  Surrogate* S = new Surrogate();
  S->device = device;
  SurDev& d = S->d;
  d.n = n;
  d.history = nd->history;
  d.layers = sh.nlin();
  d.n_out = nd->n_out;
  d.n_in = nd->n_in;
  d.dt_macro = md->dt_macro;
  d.v_rest = nd->v_rest;
  d.cap = maxd + 1;
  d.bin_steps = bin;
  std::vector<double> e0(n), i0(n);
  for (int r = 0; r < n; ++r) {
    e0[r] = wc[(size_t)r * WC_NPARAMS + WC_E0];
    i0[r] = wc[(size_t)r * WC_NPARAMS + WC_I0];
  }
  int rc = 0;
  double* pd = nullptr;
  int64_t* p64 = nullptr;
  auto fin = [&](int code) {
    delete S;
    return code;
  };
  if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess)
    return fin(fail(MIND_EHIP, "stream creation failed"));
  if ((rc = S->upload(&pd, W.data(), W.size()))) return fin(rc);
  d.W = pd;
  if ((rc = S->upload(&p64, ds.data(), ds.size()))) return fin(rc);
  d.dsteps = p64;
  if ((rc = S->upload(&d.hist, (const double*)nullptr, (size_t)d.cap * n))) return fin(rc);
  if ((rc = S->upload(&pd, wc.data(), wc.size()))) return fin(rc);
  d.wc = pd;
  for (double** p : {&d.E, &d.I, &d.coupling, &d.rate, &d.meanv, &d.acc})
    if ((rc = S->upload(p, (const double*)nullptr, (size_t)n))) return fin(rc);
  if ((rc = S->upload(&d.drive, (const double*)nullptr, (size_t)n * nd->history))) return fin(rc);
  if ((rc = S->upload(&d.clock, (const int64_t*)nullptr, 4))) return fin(rc);
  if ((rc = S->upload(&S->e0, e0.data(), e0.size()))) return fin(rc);
  if ((rc = S->upload(&S->i0, i0.data(), i0.size()))) return fin(rc);
  // network: BF16 B fragments in MFMA lane order, padded FP32 biases
  int64_t off = 0;
  for (int l = 0; l < sh.nlin(); ++l) {
    const int K = sh.k(l), N = sh.n(l), kp = MlpShape::pad(K), np = MlpShape::pad(N);
    d.kp[l] = kp;
    d.np[l] = np;
    std::vector<uint16_t> frag((size_t)kp * np);
    for (int t = 0; t < np / 16; ++t)
      for (int ks = 0; ks < kp / 16; ++ks)
        for (int lane = 0; lane < 64; ++lane)
          for (int q = 0; q < 4; ++q) {
            const int k = ks * 16 + (lane >> 4) * 4 + q, j = t * 16 + (lane & 15);
            const float w = (k < K && j < N) ? nd->params[off + (int64_t)k * N + j] : 0.0f;
            frag[(((size_t)t * (kp / 16) + ks) * 64 + lane) * 4 + q] = f32_to_bf16(w);
          }
    off += (int64_t)K * N;
    std::vector<float> b(np, 0.0f);
    for (int j = 0; j < N; ++j) b[j] = nd->params[off + j];
    off += N;
    uint16_t* pf = nullptr;
    float* pb = nullptr;
    if ((rc = S->upload(&pf, frag.data(), frag.size()))) return fin(rc);
    if ((rc = S->upload(&pb, b.data(), b.size()))) return fin(rc);
    d.frag[l] = reinterpret_cast<const bf16x4*>(pf);
    d.bias[l] = pb;
  }
  float* pf = nullptr;
  if ((rc = S->upload(&pf, nd->x_mean, (size_t)nd->n_in))) return fin(rc);
  d.x_mean = pf;
  if ((rc = S->upload(&pf, nd->x_std, (size_t)nd->n_in))) return fin(rc);
  d.x_std = pf;
  if ((rc = S->upload(&pf, nd->x_lo, (size_t)nd->n_in))) return fin(rc);
  d.x_lo = pf;
  if ((rc = S->upload(&pf, nd->x_hi, (size_t)nd->n_in))) return fin(rc);
  d.x_hi = pf;
  if ((rc = S->upload(&pf, nd->y_mean, (size_t)nd->n_out))) return fin(rc);
  d.y_mean = pf;
  if ((rc = S->upload(&pf, nd->y_std, (size_t)nd->n_out))) return fin(rc);
  d.y_std = pf;
  d.skip_w = d.skip_b = nullptr;
  if (nd->skip_w) {
    if ((rc = S->upload(&pf, nd->skip_w, (size_t)nd->n_in * nd->n_out))) return fin(rc);
    d.skip_w = pf;
    if ((rc = S->upload(&pf, nd->skip_b, (size_t)nd->n_out))) return fin(rc);
    d.skip_b = pf;
  }
  *out = reinterpret_cast<mind_surrogate*>(S);
  return 0;
}

int mind_initialize_surrogate(mind_surrogate* sur) {
  if (!sur) return mind::fail(MIND_EARG, "null argument");
  Surrogate& S = *reinterpret_cast<Surrogate*>(sur);
  HIP_TRY(hipSetDevice(S.device));
  hipLaunchKernelGGL(mind::k_surrogate_init, dim3((S.d.n + 255) / 256), dim3(256), 0, S.stream, S.d, S.e0, S.i0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(S.stream));
  S.initialized = true;
  return 0;
}

int mind_run_surrogate(mind_surrogate* sur, double stop, mind_macro_result* out, int64_t* ood_windows,
                       double* first_ood_time) {
  using namespace mind;
  if (!sur || !out) return fail(MIND_EARG, "null argument");
  Surrogate& S = *reinterpret_cast<Surrogate*>(sur);
  if (!S.initialized) return fail(MIND_EARG, "surrogate is not initialized");
  HIP_TRY(hipSetDevice(S.device));
  int64_t clock[4];
  HIP_TRY(hipMemcpy(clock, S.d.clock, sizeof(clock), hipMemcpyDeviceToHost));
  const double dtm = S.d.dt_macro;
  const double now = (double)clock[0] * dtm;
  const double duration = stop - now;
  if (!std::isfinite(duration)) return fail(MIND_EARG, "stop time must be finite");
  const int64_t count = (int64_t)std::nearbyint(duration / dtm);
  const double big = std::max(std::max(std::fabs(stop), std::fabs(now)), dtm);
  const double tol = 64.0 * (std::nextafter(big, INFINITY) - big);
  if (!(count >= 0 && std::fabs(duration - (double)count * dtm) <= tol))
    return fail(MIND_EARG, "stop time must be an integer number of steps after the current time");
  const int64_t samples = count + 1;
  const size_t tbytes = sizeof(double) * 5 * (size_t)S.d.n * samples;
  void* buf = nullptr;
  if (hipMalloc(&buf, tbytes) != hipSuccess) return fail(MIND_EOOM, "device allocation failed");
  SurDev d = S.d;
  d.trace = static_cast<double*>(buf);
  d.samples = samples;
  d.trace_start = clock[0];
  hipLaunchKernelGGL(k_surrogate_record, dim3((d.n + 255) / 256), dim3(256), 0, S.stream, d, (int64_t)0);
  hipError_t e = hipStreamSynchronize(S.stream);
  const auto t0 = std::chrono::steady_clock::now();
  for (int64_t done = 0; e == hipSuccess && done < count; done += kStepsPerLaunch) {
    hipLaunchKernelGGL(k_surrogate_run, dim3(1), dim3(kThreads), 0, S.stream, d,
                       std::min<int64_t>(kStepsPerLaunch, count - done));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(S.stream);
  const auto t1 = std::chrono::steady_clock::now();
  if (e != hipSuccess) {
    (void)hipFree(buf);
    return fail(MIND_EHIP, std::string("surrogate run: ") + hipGetErrorString(e));
  }
  std::memset(out, 0, sizeof(*out));
  out->n_samples = samples;
  out->n_roi = d.n;
  out->time = (double*)malloc(sizeof(double) * samples);
  out->traces = (double*)malloc(tbytes);
  for (int64_t j = 0; j < samples; ++j) out->time[j] = (double)(clock[0] + j) * dtm;
  e = hipMemcpy(out->traces, buf, tbytes, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(MIND_EHIP, "trace download failed");
  HIP_TRY(hipMemcpy(clock, S.d.clock, sizeof(clock), hipMemcpyDeviceToHost));
  if (ood_windows) *ood_windows = clock[2];
  if (first_ood_time) *first_ood_time = clock[3] < 0 ? -1.0 : (double)(clock[3] + 1) * dtm;
  out->run_seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

void mind_surrogate_destroy(mind_surrogate* sur) { delete reinterpret_cast<Surrogate*>(sur); }

}  // extern "C"

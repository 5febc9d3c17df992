// Micro-macro coupling on one GPU (reference advance_macro!, src/runtime/step.jl:156-204,
// and the kernels of src/runtime/cuda/macro.jl).
//
// One macro step is five launches, captured once as a hipGraph and replayed:
//   k_macro_pre   delayed coupling from the history ring, then every MACRO2MICRO
//                 binding draws its Poisson events for [start, stop) (rekeyed per
//                 macro step like input_trigger_kernel!) and routes them into the
//                 delivery ring with tickets in binding order;
//   window        gather / cell kernel / exchange over dt_macro / dt_micro micro steps;
//   k_macro_post  MICRO2MACRO spike counts of the window's emissions, population
//                 rate at bin boundaries, mean soma voltage, Wilson-Cowan RK4 per
//                 ROI, publication of E into the history ring, macro recording.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.h"
#include "macro.h"
#include "state.h"

namespace mind {

namespace {

// Random123 stream of MACRO2MICRO binding b at macro step s (cuda/macro.jl:288-296):
// keys (seed ^ stream id, logical id, seed>>32 ^ stream id>>32 ^ ROI), zero Philox key.
__device__ inline Stream input_stream(const Dev& d, const MacroDev& m, int64_t b, int64_t s) {
  Stream st;
  const uint64_t sid = m.stream_id[b];
  st.counter = 0;
  st.key[0] = (uint32_t)(d.seed ^ sid);
  st.key[1] = (uint32_t)(uint64_t)b;
  st.key[2] = (uint32_t)((d.seed >> 32) ^ (sid >> 32) ^ (uint64_t)m.in_roi[b]);
  stream_setseq(st, 0, (double)s);
  return st;
}

__device__ inline double input_rate(const MacroDev& m, int64_t b, int64_t slot) {
  const int r = m.in_roi[b];
  return drive_rate(m.wc[(int64_t)r * WC_NPARAMS + WC_RATE_SCALE], m.hist[slot * m.n + r]);
}

constexpr int kB = 1024;

__device__ inline int64_t scan_block(int64_t x, int64_t* lds, int64_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t incl = x;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int w = 0; w < kB / 64; ++w) {
      const int64_t t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[kB / 64] = acc;
  }
  __syncthreads();
  const int64_t r = lds[wave] + incl - x;
  *total = lds[kB / 64];
  __syncthreads();
  return r;
}

}  // namespace

__global__ void __launch_bounds__(kB) k_macro_pre(Dev d, MacroDev m) {
  __shared__ int64_t scan[kB / 64 + 1];
  const int tid = threadIdx.x;
  const int64_t s = d.clock->mstep;
  const int64_t k0 = d.clock->step;
  const double start = (double)s * m.dt_macro, stop = (double)(s + 1) * m.dt_macro;
  const int64_t slot = s % m.cap;
  // 1. delayed coupling (prepare_coupling_kernel! + macro_edge_kernel!): sources in order
  for (int i = tid; i < m.n; i += kB) {
    double c = 0.0;
    for (int j = 0; j < m.n; ++j) {
      const double w = m.W[(int64_t)i * m.n + j];
      if (w == 0.0) continue;
      const int64_t sl = ((s - m.dsteps[(int64_t)i * m.n + j]) % m.cap + m.cap) % m.cap;
      c = c + edge_value(w, m.hist[sl * m.n + j]);
    }
    m.coupling[i] = c;
  }
  // 2. Poisson events of each binding: count, prefix, then route with tickets
  for (int64_t b = tid; b < m.nb; b += kB) {
    const double rate = input_rate(m, b, slot);
    int64_t cnt = 0;
    if (rate > 0.0) {
      Stream st = input_stream(d, m, b, s);
      const double scale = 1000.0 / rate;
      double tn = start + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
      while (tn < stop) {
        ++cnt;
        tn = tn + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
      }
    }
    m.in_count[b] = cnt;
  }
  __syncthreads();
  int64_t running = 0;
  for (int64_t base = 0; base < m.nb; base += kB) {
    const int64_t b = base + tid;
    int64_t total;
    const int64_t pre = scan_block(b < m.nb ? m.in_count[b] : 0, scan, &total);
    if (b < m.nb) m.in_base[b] = running + pre;
    running += total;
  }
  const uint64_t ticket0 = d.clock->ticket;
  for (int64_t b = tid; b < m.nb; b += kB) {
    const double rate = input_rate(m, b, slot);
    if (!(rate > 0.0)) continue;
    Stream st = input_stream(d, m, b, s);
    const double scale = 1000.0 / rate;
    const int64_t src = m.first_virtual + b;
    const int64_t e = d.fanout[src];
    int64_t q = m.in_base[b];
    double tn = start + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
    while (tn < stop) {
      const double T = tn + d.edge_delay[e];
      const int64_t j = call_of(T, 2 * k0, d.dt);
      const int64_t bucket = (j / 2) % d.ring_slots;
      const unsigned long long pos = atomicAdd(&d.ring_count[bucket], 1ull);
      if ((int64_t)pos < d.ring_cap)
        d.ring[bucket * d.ring_cap + (int64_t)pos] = QEvent{T, ticket0 + (uint64_t)q + 1, j, (int32_t)e, 0};
      else
        raise_error(d.error, MIND_EQUEUE);
      if (d.record_spikes) {
        const unsigned long long sp = atomicAdd(d.spike_count, 1ull);
        if ((int64_t)sp < d.spike_cap) {
          d.spike_t[sp] = tn;
          d.spike_sid[sp] = d.source_sid[src];
        } else {
          raise_error(d.error, MIND_EQUEUE);
        }
      }
      ++q;
      tn = tn + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
    }
  }
  __syncthreads();
  if (tid == 0) d.clock->ticket = ticket0 + (uint64_t)running;
}

__device__ inline void record_macro(const MacroDev& m, int r, int64_t sample) {
  if (!m.trace) return;
  const int64_t S = m.samples;
  m.trace[(0 * (int64_t)m.n + r) * S + sample] = m.E[r];
  m.trace[(1 * (int64_t)m.n + r) * S + sample] = m.I[r];
  m.trace[(2 * (int64_t)m.n + r) * S + sample] = m.rate[r];
  m.trace[(3 * (int64_t)m.n + r) * S + sample] = m.coupling[r];
  m.trace[(4 * (int64_t)m.n + r) * S + sample] = m.meanv[r];
}

__device__ inline double mean_soma(const Dev& d, const MacroDev& m, int r) {
  double sum = 0.0;
  for (int64_t q = m.out_off[r]; q < m.out_off[r + 1]; ++q) sum += d.v[m.out_node[q]];
  const int64_t n = m.out_off[r + 1] - m.out_off[r];
  return n > 0 ? sum / (double)n : 0.0;
}

__global__ void __launch_bounds__(kB) k_macro_post(Dev d, MacroDev m) {
  extern __shared__ int lds_counts[];
  const int tid = threadIdx.x;
  const int64_t s = d.clock->mstep;
  const double start = (double)s * m.dt_macro, stop = (double)(s + 1) * m.dt_macro;
  for (int r = tid; r < m.n; r += kB) lds_counts[r] = 0;
  __syncthreads();
  const int64_t ne = d.clock->emitted;
  for (int64_t q = tid; q < ne; q += kB) {
    const Emit em = d.emit_all[q];
    const int r = m.src_roi[em.source];
    if (r >= 0 && em.t >= start && em.t < stop) atomicAdd(&lds_counts[r], 1);
  }
  __syncthreads();
  for (int r = tid; r < m.n; r += kB) {
    m.acc[r] = m.acc[r] + (double)lds_counts[r];
    if ((s + 1) % m.bin_steps[r] == 0) {
      const double nsrc = (double)(m.out_off[r + 1] - m.out_off[r]);
      const double bin = (double)m.bin_steps[r] * m.dt_macro;
      m.rate[r] = bin_rate(m.acc[r], nsrc, bin);
      m.acc[r] = 0.0;
    }
    m.meanv[r] = mean_soma(d, m, r);
    const double* p = m.wc + (int64_t)r * WC_NPARAMS;
    double e = m.E[r], i = m.I[r];
    region_advance(p, e, i, m.coupling[r], m.rate[r], m.dt_macro);
    m.E[r] = e;
    m.I[r] = i;
    m.hist[((s + 1) % m.cap) * m.n + r] = e;
    record_macro(m, r, s + 1 - m.trace_start);
  }
  __syncthreads();
  if (tid == 0) d.clock->mstep = s + 1;
}

__global__ void k_macro_init(Dev d, MacroDev m, const double* e0, const double* i0) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m.n) return;
  m.E[r] = e0[r];
  m.I[r] = i0[r];
  m.coupling[r] = 0.0;
  m.rate[r] = 0.0;
  m.acc[r] = 0.0;
  for (int64_t k = 0; k < m.cap; ++k) m.hist[k * m.n + r] = e0[r];
  m.meanv[r] = mean_soma(d, m, r);
}

__global__ void k_macro_record(MacroDev m, int64_t sample) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < m.n) record_macro(m, r, sample);
}

// Window of K macro steps s .. s+K-1 inside one population-rate bin (the last
// step at most ends the bin). The population rate stays fixed until the bin
// ends and every coupling delay is at least K macro steps, so the ROI states of
// the window's steps follow from what is known at its start: for each step in
// order, the delayed coupling, the Poisson events of every MACRO2MICRO binding
// (tickets in step, then binding order) and, but for the last step, the RK4
// advance and its publication into the history ring. The same operations as
// one macro step at a time, in the same order per ROI.
__global__ void __launch_bounds__(kB) k_macro_pre_bin(Dev d, MacroDev m, int32_t K) {
  __shared__ int64_t scan[kB / 64 + 1];
  const int tid = threadIdx.x;
  const int64_t s0 = d.clock->mstep;
  const int64_t k0 = d.clock->step;
  const uint64_t ticket0 = d.clock->ticket;
  int64_t routed = 0;
  double* tE = m.traj;
  double* tI = m.traj + (int64_t)m.kwin * m.n;
  double* tC = m.traj + 2 * (int64_t)m.kwin * m.n;
  for (int mi = 0; mi < K; ++mi) {
    const int64_t s = s0 + mi;
    const double start = (double)s * m.dt_macro, stop = (double)(s + 1) * m.dt_macro;
    const int64_t slot = s % m.cap;
    for (int i = tid; i < m.n; i += kB) {
      double c = 0.0;
      for (int j = 0; j < m.n; ++j) {
        const double w = m.W[(int64_t)i * m.n + j];
        if (w == 0.0) continue;
        const int64_t sl = ((s - m.dsteps[(int64_t)i * m.n + j]) % m.cap + m.cap) % m.cap;
        c = c + edge_value(w, m.hist[sl * m.n + j]);
      }
      tC[(int64_t)mi * m.n + i] = c;
    }
    for (int64_t b = tid; b < m.nb; b += kB) {
      const double rate = input_rate(m, b, slot);
      int64_t cnt = 0;
      if (rate > 0.0) {
        Stream st = input_stream(d, m, b, s);
        const double scale = 1000.0 / rate;
        double tn = start + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
        while (tn < stop) {
          ++cnt;
          tn = tn + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
        }
      }
      m.in_count[b] = cnt;
    }
    __syncthreads();
    int64_t running = 0;
    for (int64_t base = 0; base < m.nb; base += kB) {
      const int64_t b = base + tid;
      int64_t total;
      const int64_t pre = scan_block(b < m.nb ? m.in_count[b] : 0, scan, &total);
      if (b < m.nb) m.in_base[b] = running + pre;
      running += total;
    }
    for (int64_t b = tid; b < m.nb; b += kB) {
      const double rate = input_rate(m, b, slot);
      if (!(rate > 0.0)) continue;
      Stream st = input_stream(d, m, b, s);
      const double scale = 1000.0 / rate;
      const int64_t src = m.first_virtual + b;
      const int64_t e = d.fanout[src];
      int64_t q = routed + m.in_base[b];
      double tn = start + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
      while (tn < stop) {
        const double T = tn + d.edge_delay[e];
        const int64_t j = call_of(T, 2 * k0, d.dt);
        const int64_t bucket = (j / 2) % d.ring_slots;
        const unsigned long long pos = atomicAdd(&d.ring_count[bucket], 1ull);
        if ((int64_t)pos < d.ring_cap)
          d.ring[bucket * d.ring_cap + (int64_t)pos] = QEvent{T, ticket0 + (uint64_t)q + 1, j, (int32_t)e, 0};
        else
          raise_error(d.error, MIND_EQUEUE);
        if (d.record_spikes) {
          const unsigned long long sp = atomicAdd(d.spike_count, 1ull);
          if ((int64_t)sp < d.spike_cap) {
            d.spike_t[sp] = tn;
            d.spike_sid[sp] = d.source_sid[src];
          } else {
            raise_error(d.error, MIND_EQUEUE);
          }
        }
        ++q;
        tn = tn + -mind_log(stream_uniform(st, 0)) * 1.0 * scale;
      }
    }
    routed += running;
    __syncthreads();
    if (mi + 1 < K) {  // the advance of a step that does not end the bin
      for (int r = tid; r < m.n; r += kB) {
        const double* p = m.wc + (int64_t)r * WC_NPARAMS;
        double e = m.E[r], i = m.I[r];
        region_advance(p, e, i, tC[(int64_t)mi * m.n + r], m.rate[r], m.dt_macro);
        m.E[r] = e;
        m.I[r] = i;
        m.hist[((s + 1) % m.cap) * m.n + r] = e;
        tE[(int64_t)mi * m.n + r] = e;
        tI[(int64_t)mi * m.n + r] = i;
      }
      __syncthreads();
    }
  }
  if (tid == 0) d.clock->ticket = ticket0 + (uint64_t)routed;
}

// After the micro window of K macro steps: per step, the outputs' spike counts
// (emissions of the step's micro steps within its time span), the bin's rate at
// its end, the mean soma voltage from the step's snapshot, the last step's RK4
// advance and publication, and the macro records of every step.
__global__ void __launch_bounds__(kB) k_macro_post_bin(Dev d, MacroDev m, int32_t K) {
  extern __shared__ int lds_counts[];  // [K][n]
  const int tid = threadIdx.x;
  const int64_t s0 = d.clock->mstep;
  const int64_t k0 = d.clock->step - (int64_t)K * m.ratio;
  for (int q = tid; q < K * m.n; q += kB) lds_counts[q] = 0;
  __syncthreads();
  const int64_t ne = d.clock->emitted;
  for (int64_t q = tid; q < ne; q += kB) {
    const Emit em = d.emit_all[q];
    const int r = m.src_roi[em.source];
    if (r < 0) continue;
    const int64_t mi = (em.key / 3 - k0) / m.ratio;
    if (mi < 0 || mi >= K) continue;
    const double start = (double)(s0 + mi) * m.dt_macro, stop = (double)(s0 + mi + 1) * m.dt_macro;
    if (em.t >= start && em.t < stop) atomicAdd(&lds_counts[mi * m.n + r], 1);
  }
  __syncthreads();
  const double* tE = m.traj;
  const double* tI = m.traj + (int64_t)m.kwin * m.n;
  const double* tC = m.traj + 2 * (int64_t)m.kwin * m.n;
  for (int r = tid; r < m.n; r += kB) {
    const double* p = m.wc + (int64_t)r * WC_NPARAMS;
    for (int mi = 0; mi < K; ++mi) {
      const int64_t s = s0 + mi;
      m.acc[r] = m.acc[r] + (double)lds_counts[mi * m.n + r];
      if ((s + 1) % m.bin_steps[r] == 0) {
        const double nsrc = (double)(m.out_off[r + 1] - m.out_off[r]);
        const double bin = (double)m.bin_steps[r] * m.dt_macro;
        m.rate[r] = bin_rate(m.acc[r], nsrc, bin);
        m.acc[r] = 0.0;
      }
      double sum = 0.0;
      for (int64_t q = m.out_off[r]; q < m.out_off[r + 1]; ++q) sum += m.snap[(int64_t)mi * m.n_out + q];
      const int64_t cnt = m.out_off[r + 1] - m.out_off[r];
      m.meanv[r] = cnt > 0 ? sum / (double)cnt : 0.0;
      m.coupling[r] = tC[(int64_t)mi * m.n + r];
      if (mi + 1 == K) {
        double e = m.E[r], i = m.I[r];
        region_advance(p, e, i, m.coupling[r], m.rate[r], m.dt_macro);
        m.E[r] = e;
        m.I[r] = i;
        m.hist[((s + 1) % m.cap) * m.n + r] = e;
        record_macro(m, r, s + 1 - m.trace_start);
      } else if (m.trace) {
        const int64_t S = m.samples, at = s + 1 - m.trace_start;
        m.trace[(0 * (int64_t)m.n + r) * S + at] = tE[(int64_t)mi * m.n + r];
        m.trace[(1 * (int64_t)m.n + r) * S + at] = tI[(int64_t)mi * m.n + r];
        m.trace[(2 * (int64_t)m.n + r) * S + at] = m.rate[r];
        m.trace[(3 * (int64_t)m.n + r) * S + at] = m.coupling[r];
        m.trace[(4 * (int64_t)m.n + r) * S + at] = m.meanv[r];
      }
    }
  }
  __syncthreads();
  if (tid == 0) d.clock->mstep = s0 + K;
}

namespace {

template <typename T>
int mupload(MacroHost& H, T** out, const std::vector<T>& v) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)) != hipSuccess)
    return fail(MIND_EOOM, "device allocation failed");
  H.allocations.push_back(p);
  if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MIND_EHIP, "macro upload failed");
  *out = static_cast<T*>(p);
  return 0;
}

}  // namespace

void macro_release(State& S) {
  MacroHost* H = S.macro;
  if (!H) return;
  if (H->graph) (void)hipGraphExecDestroy(H->graph);
  for (void* p : H->allocations) (void)hipFree(p);
  delete H;
  S.macro = nullptr;
}

int macro_setup(State& S, const mind_macro_desc& md) {
  const Layout& L = S.L;
  if (S.serial)
    return fail(MIND_EUNSUPPORTED, "micro-macro coupling needs micro NetCon delays of at least 1.5 dt and no "
                                   "connections into NetStims");
  MacroHost* H = new MacroHost();
  S.macro = H;
  MacroDev& m = H->m;
  const int n = md.n_roi;
  if (n <= 0 || n > kB) return fail(MIND_EARG, "the macro model needs 1..1024 ROIs");
  m.n = n;
  m.dt_macro = md.dt_macro;
  const double ratio = md.dt_macro / L.dt;
  if (!(std::fabs(ratio - std::nearbyint(ratio)) < 1e-9) || std::nearbyint(ratio) < 1)
    return fail(MIND_EARG, "dt_macro must be an integer multiple of dt_micro");
  m.ratio = (int32_t)std::nearbyint(ratio);
  if (S.wmax < m.ratio)
    return fail(MIND_EUNSUPPORTED, "micro NetCon delays must cover a whole macro step on the device path");
  const int32_t micro_window = S.wmax;  // longest micro window the micro delays allow
  S.wmax = m.ratio;
  std::vector<double> W(md.weights, md.weights + (size_t)n * n), wc(md.wc_params, md.wc_params + (size_t)n * WC_NPARAMS);
  std::vector<int64_t> ds((size_t)n * n);
  int64_t maxd = 0;
  for (size_t q = 0; q < ds.size(); ++q) {
    if (!std::isfinite(md.delays[q]) || md.delays[q] < 0) return fail(MIND_EARG, "invalid macro delay");
    ds[q] = (int64_t)std::nearbyint(md.delays[q] / md.dt_macro);  // pack_macro (macro/model.jl:225)
    if (W[q] != 0.0) maxd = std::max(maxd, ds[q]);
  }
  m.cap = maxd + 1;
  std::vector<int64_t> bins(n);
  for (int r = 0; r < n; ++r) bins[r] = std::max<int64_t>(1, (int64_t)std::nearbyint(md.bin_ms[r] / md.dt_macro));
  // Windows of one bin: every ROI with the same bin of K >= 2 macro steps, every
  // coupling delay at least K macro steps (so no step of a window reads an E the
  // window produces), and micro delays that allow a window of K macro steps.
  int64_t min_coupling = INT64_MAX;
  for (size_t q = 0; q < ds.size(); ++q)
    if (W[q] != 0.0) min_coupling = std::min(min_coupling, ds[q]);
  const int64_t K = bins[0];
  bool same = true;
  for (int r = 0; r < n; ++r) same = same && bins[r] == K;
  H->bins = same && K >= 2 && min_coupling >= K && (int64_t)micro_window >= K * m.ratio &&
            K * m.ratio <= (int64_t)1 << 20 && K * (int64_t)n * (int64_t)sizeof(int) <= 64 * 1024 &&
            (S.window_steps_opt == 0 || S.window_steps_opt >= K * m.ratio);
  m.kwin = H->bins ? (int32_t)K : 1;
  if (H->bins) S.wmax = (int32_t)(K * m.ratio);
  const int64_t nb = md.input_offsets[n];
  std::vector<int32_t> in_roi(nb);
  std::vector<uint64_t> sid(nb);
  std::vector<int64_t> roi_stream(n, -1);
  int64_t next = 0;
  for (int r = 0; r < n; ++r)
    for (int64_t b = md.input_offsets[r]; b < md.input_offsets[r + 1]; ++b) in_roi[b] = r;
  for (int64_t b = 0; b < nb; ++b) {  // one random identity per (mechanism, ROI, parameters)
    if (roi_stream[in_roi[b]] < 0) roi_stream[in_roi[b]] = next++;
    sid[b] = (uint64_t)roi_stream[in_roi[b]];
  }
  std::vector<int64_t> out_off(md.output_offsets, md.output_offsets + n + 1);
  std::vector<int32_t> out_node, src_roi(L.nsources, -1);
  for (int r = 0; r < n; ++r)
    for (int64_t q = out_off[r]; q < out_off[r + 1]; ++q) {
      const int32_t src = md.output_source[q];
      src_roi[src] = r;
      int32_t node = -1;
      for (size_t k = 0; k < L.det_source.size(); ++k)
        if (L.det_source[k] == src) node = L.det_node[k];
      out_node.push_back(node);
    }
  H->E0.resize(n);
  H->I0.resize(n);
  for (int r = 0; r < n; ++r) {
    H->E0[r] = wc[(size_t)r * WC_NPARAMS + WC_E0];
    H->I0[r] = wc[(size_t)r * WC_NPARAMS + WC_I0];
  }
  int rc;
  double* pd;
  int64_t* p64;
  int32_t* p32;
  uint64_t* pu;
  if ((rc = mupload(*H, &pd, W))) return rc; m.W = pd;
  if ((rc = mupload(*H, &p64, ds))) return rc; m.dsteps = p64;
  if ((rc = mupload(*H, &pd, std::vector<double>((size_t)m.cap * n, 0.0)))) return rc; m.hist = pd;
  if ((rc = mupload(*H, &pd, wc))) return rc; m.wc = pd;
  const std::vector<double> zeros(n, 0.0);
  if ((rc = mupload(*H, &m.E, zeros))) return rc;
  if ((rc = mupload(*H, &m.I, zeros))) return rc;
  if ((rc = mupload(*H, &m.coupling, zeros))) return rc;
  if ((rc = mupload(*H, &m.rate, zeros))) return rc;
  if ((rc = mupload(*H, &m.acc, zeros))) return rc;
  if ((rc = mupload(*H, &m.meanv, zeros))) return rc;
  if ((rc = mupload(*H, &p64, bins))) return rc; m.bin_steps = p64;
  if ((rc = mupload(*H, &p64, out_off))) return rc; m.out_off = p64;
  if ((rc = mupload(*H, &p32, out_node))) return rc; m.out_node = p32;
  if ((rc = mupload(*H, &p32, src_roi))) return rc; m.src_roi = p32;
  m.nb = nb;
  m.first_virtual = L.first_virtual;
  if ((rc = mupload(*H, &p32, in_roi))) return rc; m.in_roi = p32;
  if ((rc = mupload(*H, &pu, sid))) return rc; m.stream_id = pu;
  if ((rc = mupload(*H, &m.in_count, std::vector<int64_t>(nb, 0)))) return rc;
  if ((rc = mupload(*H, &m.in_base, std::vector<int64_t>(nb, 0)))) return rc;
  if (H->bins) {
    m.n_out = out_off[n];
    if ((rc = mupload(*H, &m.traj, std::vector<double>((size_t)3 * m.kwin * n, 0.0)))) return rc;
    double* snap;
    if ((rc = mupload(*H, &snap, std::vector<double>((size_t)std::max<int64_t>(1, m.kwin * m.n_out), 0.0)))) return rc;
    m.snap = snap;
    std::vector<int32_t> det_snap(L.det_source.size(), -1);
    for (size_t k = 0; k < L.det_source.size(); ++k)
      for (int64_t q = 0; q < m.n_out; ++q)
        if (md.output_source[q] == L.det_source[k]) det_snap[k] = (int32_t)q;
    if ((rc = mupload(*H, &p32, det_snap))) return rc;
    S.d.snap = snap;
    S.d.det_snap = p32;
    S.d.snap_every = m.ratio;
    S.d.n_snap = m.n_out;
  }
  // E0 and I0 stay the last two uploads (macro_initialize)
  if ((rc = mupload(*H, &pd, H->E0))) return rc;
  if ((rc = mupload(*H, &pd, H->I0))) return rc;
  return 0;
}

int macro_initialize(State& S) {
  MacroHost* H = S.macro;
  if (!H) return 0;
  // E0 and I0 are the last two uploads of macro_setup
  const double* e0 = static_cast<const double*>(H->allocations[H->allocations.size() - 2]);
  const double* i0 = static_cast<const double*>(H->allocations[H->allocations.size() - 1]);
  hipLaunchKernelGGL(k_macro_init, dim3((H->m.n + 255) / 256), dim3(256), 0, S.stream, S.d, H->m, e0, i0);
  HIP_TRY(hipGetLastError());
  return 0;
}

int launch_macro_step(State& S) {
  MacroHost& H = *S.macro;
  hipLaunchKernelGGL(k_macro_pre, dim3(1), dim3(kB), 0, S.stream, S.d, H.m);
  int rc = launch_window(S, H.m.ratio);
  if (rc) return rc;
  hipLaunchKernelGGL(k_macro_post, dim3(1), dim3(kB), sizeof(int) * H.m.n, S.stream, S.d, H.m);
  HIP_TRY(hipGetLastError());
  return 0;
}

// K macro steps inside one rate bin as one micro window.
int launch_macro_window(State& S, int32_t K) {
  MacroHost& H = *S.macro;
  hipLaunchKernelGGL(k_macro_pre_bin, dim3(1), dim3(kB), 0, S.stream, S.d, H.m, K);
  int rc = launch_window(S, K * H.m.ratio);
  if (rc) return rc;
  hipLaunchKernelGGL(k_macro_post_bin, dim3(1), dim3(kB), sizeof(int) * (size_t)K * H.m.n, S.stream, S.d, H.m, K);
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace mind

using mind::State;

extern "C" {

int mind_build_macro(const mind_model_desc* micro, const mind_macro_desc* macro,
                     const mind_build_options* options, mind_state** out) {
  if (!micro || !macro || !options || !out) return mind::fail(MIND_EARG, "null argument");
  if (options->secondorder != 0)
    return mind::fail(MIND_EARG, "secondorder is not supported with micro-macro coupling");
  State* S = new State();
  std::string err = mind::pack(*micro, options->dt, options->seed, S->L);
  if (err.empty()) err = mind::pack_inputs(*micro, *macro, S->L);
  if (!err.empty()) {
    delete S;
    return mind::fail(MIND_EARG, err);
  }
  int rc = mind::setup_state(*S, *options);
  if (!rc) rc = mind::macro_setup(*S, *macro);
  if (!rc && options->method == MIND_METHOD_BATCHED) rc = mind::batched_setup(*S, options->cells_per_group);
  if (rc) {
    delete S;
    return rc;
  }
  *out = reinterpret_cast<mind_state*>(S);
  return 0;
}

int mind_run_macro(mind_state* state, double stop, int32_t record_spikes, mind_macro_result* out) {
  if (!state || !out) return mind::fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  if (!S.macro) return mind::fail(MIND_EARG, "state has no macro model");
  if (!S.initialized) return mind::fail(MIND_EARG, "simulation is not initialized");
  HIP_TRY(hipSetDevice(S.device));
  mind::MacroHost& H = *S.macro;
  mind::Dev& d = S.d;
  int64_t mstep = 0;
  HIP_TRY(hipMemcpy(&mstep, &d.clock->mstep, sizeof(int64_t), hipMemcpyDeviceToHost));
  const double dtm = H.m.dt_macro;
  const double now = (double)mstep * dtm;
  const double duration = stop - now;
  if (!std::isfinite(duration)) return mind::fail(MIND_EARG, "stop time must be finite");
  const int64_t count = (int64_t)std::nearbyint(duration / dtm);
  const double big = std::max(std::max(std::fabs(stop), std::fabs(now)), dtm);
  const double tol = 64.0 * (std::nextafter(big, INFINITY) - big);
  if (!(count >= 0 && std::fabs(duration - (double)count * dtm) <= tol))
    return mind::fail(MIND_EARG, "stop time must be an integer number of steps after the current time");
  const int64_t samples = count + 1;
  void* buf = nullptr;
  if (hipMalloc(&buf, sizeof(double) * 5 * H.m.n * samples) != hipSuccess)
    return mind::fail(MIND_EOOM, "device allocation failed");
  H.m.trace = static_cast<double*>(buf);
  H.m.samples = samples;
  H.m.trace_start = mstep;
  d.nrec = 0;
  d.record_spikes = record_spikes ? 1 : 0;
  if (S.step != 0 || !record_spikes)
    HIP_TRY(hipMemsetAsync(d.spike_count, 0, sizeof(unsigned long long), S.stream));
  if (H.graph) {
    (void)hipGraphExecDestroy(H.graph);
    H.graph = nullptr;
  }
  if (S.graph) {
    (void)hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
  }
  hipLaunchKernelGGL(mind::k_macro_record, dim3((H.m.n + 255) / 256), dim3(256), 0, S.stream, H.m, (int64_t)0);
  HIP_TRY(hipStreamSynchronize(S.stream));
  const auto t0 = std::chrono::steady_clock::now();
  int rc = 0;
  // windows: one macro step, or (bins) the macro steps up to the next bin end
  for (int64_t j = 0, ms = mstep; j < count;) {
    const int32_t K = H.bins ? (int32_t)std::min<int64_t>(H.m.kwin - ms % H.m.kwin, count - j) : 1;
    if ((rc = mind::stage_injections(S, K * H.m.ratio))) break;
    const bool full = K == H.m.kwin;
    if (!S.profiling && full) {
      if (!H.graph) {
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
        rc = H.bins ? mind::launch_macro_window(S, K) : mind::launch_macro_step(S);
        hipError_t e = hipStreamEndCapture(S.stream, &g);
        if (rc) break;
        if (e != hipSuccess) {
          rc = mind::fail(MIND_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
          break;
        }
        HIP_TRY(hipGraphInstantiate(&H.graph, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
      }
      HIP_TRY(hipGraphLaunch(H.graph, S.stream));
    } else if ((rc = H.bins ? mind::launch_macro_window(S, K) : mind::launch_macro_step(S))) {
      break;
    }
    S.step += (int64_t)K * H.m.ratio;
    j += K;
    ms += K;
  }
  HIP_TRY(hipStreamSynchronize(S.stream));
  const auto t1 = std::chrono::steady_clock::now();
  if (!rc) rc = mind::check_device_error(S);
  d.record_spikes = 0;
  if (H.graph) {
    (void)hipGraphExecDestroy(H.graph);
    H.graph = nullptr;
  }
  if (rc) {
    (void)hipFree(buf);
    H.m.trace = nullptr;
    return rc;
  }
  out->n_samples = samples;
  out->n_roi = H.m.n;
  out->time = (double*)malloc(sizeof(double) * samples);
  for (int64_t j = 0; j < samples; ++j) out->time[j] = (double)(mstep + j) * dtm;
  out->traces = (double*)malloc(sizeof(double) * 5 * H.m.n * samples);
  HIP_TRY(hipMemcpy(out->traces, buf, sizeof(double) * 5 * H.m.n * samples, hipMemcpyDeviceToHost));
  (void)hipFree(buf);
  H.m.trace = nullptr;
  rc = mind::collect_spikes(S, &out->n_spikes, &out->spike_times, &out->spike_sids);
  out->run_seconds = std::chrono::duration<double>(t1 - t0).count();
  return rc;
}

void mind_macro_result_free(mind_macro_result* r) {
  if (!r) return;
  free(r->time);
  free(r->traces);
  free(r->spike_times);
  free(r->spike_sids);
  memset(r, 0, sizeof(*r));
}

}  // extern "C"

// MI355X runtime of the MIND simulation step: device state, kernels and the C ABI.
//
// Execution model. The reference runs ~12 launches per micro step with a
// blocking device->host read of the outbox every step and a CPU binary heap
// for events (src/runtime/step.jl:91-151, src/runtime/events.jl:4-165). Here
// the whole loop stays on the device and advances in *windows* of W steps,
// W*dt <= min NetCon delay - dt/2: no event produced inside a window can be
// due inside it, so cells are independent for W steps. Per window:
//   gather   : events due in the window -> per-cell lists sorted by
//              (delivery call, target, time, ticket)
//   cells    : W steps of detect / deliver / nrn_cur / cable / nrn_state /
//              record / deliver for every cell (plain or batched kernel)
//   exchange : NetStim events of the window in heap order, detector spikes
//              compacted by wave ballot + prefix sum, heap tickets re-derived
//              in emission order, NetCon fanout routed into a ring of
//              per-step delivery buckets.
// The three launches of a full window are captured once in a hipGraph and
// replayed; the clock lives in device memory.
//
// Event order contract (the reference's CPU heap, src/runtime/utils/events.jl:7-62):
// a target applies its events in (time, ticket) order, tickets being
// insertion order: per step, first the detector spikes of that step in source
// order (each routed over its CSR edges in edge order, events.jl:34-53), then
// NetStim events of the midpoint delivery, then of the end-of-step delivery,
// in the heap's (time, ticket) order (events.jl:92-101, 160-162).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "device.h"
#include "state.h"
#include "cell_kernels.h"

namespace mind {

// ------------------------------------------------------------ initialization
__global__ void k_init_nodes(Dev d, const double* v_init) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.nodes) d.v[i] = v_init[i];
  if (i < d.ndet) d.det_above[i] = 0;
}

__global__ void k_init_netstim(Dev d, const uint64_t* logical, uint64_t digest, uint64_t seq0) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d.ns_rows) return;
  const GroupDev& G = d.g[d.ns_group];
  Stream s;
  s.counter = 0;
  s.key[0] = (uint32_t)logical[r];
  s.key[1] = logical[r] >> 32;
  s.key[2] = (uint64_t)(uint32_t)digest;
  stream_setseq(s, d.seed, 0.0);
  const NetStimOut o = phase_netstim_initial(cols(G), r, s, d.seed);
  d.ns_stream[r] = s;
  d.ns_valid[r] = 0;
  if (o.send) {
    if (!(o.send_delay >= 0.0) || !mind_isfinite(o.send_delay)) {
      raise_error(d.error, MIND_EEVENT);
      return;
    }
    d.ns_t[r] = 0.0 + o.send_delay;
    d.ns_flag[r] = o.send_flag;
    d.ns_seq[r] = seq0 + (uint64_t)r + 1;  // INITIAL flush orders NetStims by rank, after the pattern spikes
    d.ns_edge[r] = 0;
    d.ns_valid[r] = 1;
  }
}

__global__ void k_record_sample(Dev d, int64_t sample) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.nrec) return;
  const double x = d.rec_kind[q] == 0 ? d.v[d.rec_index[q]] : stored_value(d, (int32_t)q);
  d.rec_buf[q * d.rec_stride + sample] = x;
}

// ------------------------------------------------------------ block helpers
constexpr int kBlock = 1024;

// Exclusive scan of one value per thread over the block; returns the prefix
// and writes the block total to *total.
__device__ inline int64_t block_scan(int64_t x, int64_t* lds, int64_t* total) {
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
    for (int w = 0; w < kBlock / 64; ++w) {
      const int64_t t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[kBlock / 64] = acc;
  }
  __syncthreads();
  const int64_t r = lds[wave] + incl - x;
  *total = lds[kBlock / 64];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ exchange
constexpr int kHeapCap = 2048;  // NetStim and injected events of a window held in LDS; more use Dev::ns_heap

__device__ inline bool heap_less(const HeapItem& a, const HeapItem& b) {
  return a.t < b.t || (a.t == b.t && a.seq < b.seq);
}

__device__ void heap_push(HeapItem* h, int& n, HeapItem x) {
  int i = n++;
  while (i > 0) {
    const int p = (i - 1) >> 1;
    if (!heap_less(x, h[p])) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = x;
}

__device__ HeapItem heap_pop(HeapItem* h, int& n) {
  const HeapItem top = h[0];
  const HeapItem last = h[--n];
  int i = 0;
  while (2 * i + 1 < n) {
    int c = 2 * i + 1;
    if (c + 1 < n && heap_less(h[c + 1], h[c])) ++c;
    if (!heap_less(h[c], last)) break;
    h[i] = h[c];
    i = c;
  }
  if (n > 0) h[i] = last;
  return top;
}

// One workgroup. n == 0 processes the time-zero delivery of initialize!.
__global__ void __launch_bounds__(kBlock) k_exchange(Dev d, int32_t n) {
  __shared__ HeapItem heap[kHeapCap];
  __shared__ int64_t scan[kBlock / 64 + 1];
  __shared__ int heap_n;
  __shared__ int64_t n_ns, n_det;
  const int tid = threadIdx.x;
  const int64_t k0 = d.clock->step;
  const int64_t call_lo = n == 0 ? -1 : 2 * k0;
  const int64_t call_hi = n == 0 ? -1 : 2 * (k0 + n) - 1;
  const double dt = d.dt;
  const double cut_hi = call_cutoff(call_hi, dt);
  // A. NetStim events and injected source spikes of the window, in heap order
  //    (time, insertion order); an injected spike k enters as row -(k + 1).
  const int32_t ninj = *d.inj_count;
  if (tid == 0) { heap_n = 0; n_ns = 0; }
  __syncthreads();
  // every due row and injection gets a slot in HBM (one pending event per row,
  // at most kInjCap injections, so ns_heap never overflows); the first kHeapCap
  // also in LDS, used when all of them fit
  for (int64_t r = tid; r < d.ns_rows; r += kBlock) {
    if (d.ns_valid[r] && d.ns_t[r] <= cut_hi) {
      const int slot = atomicAdd(&heap_n, 1);
      const HeapItem it{d.ns_t[r], d.ns_seq[r], (int32_t)r};
      if (slot < kHeapCap) heap[slot] = it;
      d.ns_heap[slot] = it;
    }
  }
  for (int k = tid; k < ninj && k < kInjCap; k += kBlock) {
    const int slot = atomicAdd(&heap_n, 1);
    const HeapItem it{d.inj[k].t, d.inj[k].seq, -(k + 1)};
    if (slot < kHeapCap) heap[slot] = it;
    d.ns_heap[slot] = it;
  }
  __syncthreads();
  if (tid == 0 && (d.ns_rows > 0 || ninj > 0)) {
    const int hn = heap_n;
    // a popped row pushes at most its own next event, so the heap never grows
    // beyond its initial size
    HeapItem* hp = hn <= kHeapCap ? heap : d.ns_heap;
    int m = 0;
    for (int i = 0; i < hn; ++i) heap_push(hp, m, hp[i]);
    const Cols col = d.ns_group >= 0 ? cols(d.g[d.ns_group]) : Cols{nullptr, 0};
    int64_t count = 0;
    uint64_t seq = d.clock->ns_seq;
    // artificial-cell records of step k take the state after the events of
    // calls <= 2k (the record precedes the end-of-step delivery)
    int64_t sampled = n > 0 ? k0 : INT64_MAX;
    while (m > 0) {
      const HeapItem it = heap_pop(hp, m);
      const int32_t r = it.row;
      const double t = it.t;
      if (d.n_art_rec) {
        const int64_t jt = call_of(t, call_lo, dt);
        for (; sampled < k0 + n && 2 * sampled < jt; ++sampled) record_artificial(d, sampled);
      }
      if (r < 0) {  // injected source spike: routed at its delivery call, like a net_event
        const int64_t j = call_of(t, call_lo, dt);
        if (count < d.emit_cap) d.emit_ns[count] = Emit{emit_key(j), t, d.inj[-r - 1].source, 0};
        else raise_error(d.error, MIND_EQUEUE);
        ++count;
        continue;
      }
      const double flag = d.ns_flag[r];
      const int32_t edge = d.ns_edge[r];
      d.ns_valid[r] = 0;
      const int64_t j = call_of(t, call_lo, dt);
      const double w = edge == 0 ? 0.0 : d.edge_weight[edge - 1];
      Stream s = d.ns_stream[r];
      const NetStimOut o = phase_netstim_receive(col, r, s, d.seed, t, w, flag);
      d.ns_stream[r] = s;
      if (o.spiked && d.ns_source[r] >= 0) {
        if (count < d.emit_cap) d.emit_ns[count] = Emit{emit_key(j), o.spike_time, d.ns_source[r], 0};
        else raise_error(d.error, MIND_EQUEUE);
        ++count;
      }
      if (o.send) {
        if (!(o.send_delay >= 0.0) || !mind_isfinite(o.send_delay)) {
          raise_error(d.error, MIND_EEVENT);
          continue;
        }
        const double tn = t + o.send_delay;
        d.ns_t[r] = tn;
        d.ns_flag[r] = o.send_flag;
        d.ns_seq[r] = ++seq;
        d.ns_edge[r] = edge;
        d.ns_valid[r] = 1;
        if (tn <= cut_hi) heap_push(hp, m, HeapItem{tn, seq, r});
      }
    }
    d.clock->ns_seq = seq;
    n_ns = count < d.emit_cap ? count : d.emit_cap;
    if (d.n_art_rec)
      for (; sampled < k0 + n; ++sampled) record_artificial(d, sampled);
  }
  __syncthreads();
  // B. Detector spikes of the window in (step, detector) order. The flags of all
  //    steps (row-major [step][detector], one byte each, 0 or 1) are cut into one
  //    contiguous chunk per thread, in thread order, a multiple of 8 bytes long:
  //    each thread counts its set flags eight at a time, a block prefix sum gives
  //    each chunk its first position, and a second walk writes every set flag
  //    at its position.
  int64_t ndet_total = 0;
  {
    const int64_t nf = (int64_t)n * d.ndet;
    const int64_t chunk = ((nf + kBlock - 1) / kBlock + 7) / 8 * 8;
    const int64_t f0 = std::min<int64_t>((int64_t)tid * chunk, nf), f1 = std::min<int64_t>(f0 + chunk, nf);
    int64_t count = 0;
    for (int64_t f = f0; f < f1; f += 8) {
      if (f + 8 <= f1) {
        count += __popcll(*reinterpret_cast<const uint64_t*>(d.det_flags + f));
      } else {
        for (int64_t g = f; g < f1; ++g) count += d.det_flags[g] != 0;
      }
    }
    int64_t total;
    int64_t pos = block_scan(count, scan, &total);
    for (int64_t f = f0; f < f1 && count > 0; f += 8) {
      uint64_t w = 0;
      if (f + 8 <= f1) {
        w = *reinterpret_cast<const uint64_t*>(d.det_flags + f);
      } else {
        for (int64_t g = f; g < f1; ++g) w |= (uint64_t)(d.det_flags[g] != 0) << (8 * (g - f));
      }
      for (; w != 0; w &= w - 1) {
        const int64_t fi = f + (__builtin_ctzll(w) >> 3);  // flag bytes are 0 or 1
        const int64_t st = fi / d.ndet, q = fi - st * d.ndet;
        if (pos < d.emit_cap)
          d.emit_det[pos] = Emit{3 * (k0 + st), (double)(k0 + st) * dt + 1.0e-10, d.det_source[q], 0};
        else
          raise_error(d.error, MIND_EQUEUE);
        ++pos;
      }
    }
    ndet_total = total;
  }
  if (tid == 0) n_det = ndet_total < d.emit_cap ? ndet_total : d.emit_cap;
  __syncthreads();
  // C. Merge both emission lists by key (they never share a key).
  const int64_t nd = n_det, ns = n_ns;
  for (int64_t i = tid; i < nd + ns; i += kBlock) {
    Emit e;
    int64_t pos;
    if (i < nd) {
      e = d.emit_det[i];
      int64_t lo = 0, hi = ns;  // count of NetStim emissions with key < e.key
      while (lo < hi) { const int64_t mid = (lo + hi) >> 1; if (d.emit_ns[mid].key < e.key) lo = mid + 1; else hi = mid; }
      pos = i + lo;
    } else {
      e = d.emit_ns[i - nd];
      int64_t lo = 0, hi = nd;
      while (lo < hi) { const int64_t mid = (lo + hi) >> 1; if (d.emit_det[mid].key < e.key) lo = mid + 1; else hi = mid; }
      pos = (i - nd) + lo;
    }
    d.emit_all[pos] = e;
  }
  __syncthreads();
  // D. Tickets: prefix sum of fanout sizes in emission order.
  const int64_t ne = nd + ns;
  int64_t running = 0;
  for (int64_t base = 0; base < ne; base += kBlock) {
    const int64_t i = base + tid;
    const int64_t f = i < ne ? d.fanout[d.emit_all[i].source + 1] - d.fanout[d.emit_all[i].source] : 0;
    int64_t total;
    const int64_t pre = block_scan(f, scan, &total);
    if (i < ne) d.emit_base[i] = running + pre;
    running += total;
  }
  if (tid == 0) d.emit_base[ne] = running;
  __syncthreads();
  // E. Route every edge of every emission into its delivery bucket.
  const uint64_t ticket0 = d.clock->ticket;
  for (int64_t q = tid; q < running; q += kBlock) {
    int64_t lo = 0, hi = ne;  // emission m with base[m] <= q < base[m+1]
    while (hi - lo > 1) { const int64_t mid = (lo + hi) >> 1; if (d.emit_base[mid] <= q) lo = mid; else hi = mid; }
    const Emit& em = d.emit_all[lo];
    const int64_t e = d.fanout[em.source] + (q - d.emit_base[lo]);
    const double T = em.t + d.edge_delay[e];
    if (!mind_isfinite(T)) { raise_error(d.error, MIND_ENONFINITE); continue; }
    const int64_t j = call_of(T, call_hi + 1, dt);
    if (T <= cut_hi) { raise_error(d.error, MIND_EUNSUPPORTED); continue; }
    const int64_t step = j / 2;
    if (step - (k0 + n) >= d.ring_slots - 1) { raise_error(d.error, MIND_EQUEUE); continue; }
    const int64_t b = step % d.ring_slots;
    const unsigned long long pos = atomicAdd(&d.ring_count[b], 1ull);
    if ((int64_t)pos >= d.ring_cap) { raise_error(d.error, MIND_EQUEUE); continue; }
    d.ring[b * d.ring_cap + (int64_t)pos] = QEvent{T, ticket0 + (uint64_t)q + 1, j, (int32_t)e, 0};
  }
  // F. Spike recording.
  if (d.record_spikes) {
    for (int64_t i = tid; i < ne; i += kBlock) {
      const unsigned long long pos = atomicAdd(d.spike_count, 1ull);
      if ((int64_t)pos < d.spike_cap) {
        d.spike_t[pos] = d.emit_all[i].t;
        d.spike_sid[pos] = d.source_sid[d.emit_all[i].source];
      } else {
        raise_error(d.error, MIND_EQUEUE);
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    d.clock->ticket = ticket0 + (uint64_t)running;
    d.clock->step = k0 + n;
    d.clock->emitted = ne;
    *d.inj_count = 0;
  }
}

// ------------------------------------------------------------ gather
__device__ inline bool win_less(const QEvent& a, const QEvent& b, const int32_t* edge_target) {
  if (a.call != b.call) return a.call < b.call;
  const int32_t ta = edge_target[a.edge], tb = edge_target[b.edge];
  if (ta != tb) return ta < tb;
  if (a.t != b.t) return a.t < b.t;
  return a.ticket < b.ticket;
}

// Restores the max-heap (by win_less) of w[0, m) below `root`.
__device__ inline void sift_down(QEvent* w, int64_t root, int64_t m, const int32_t* edge_target) {
  const QEvent x = w[root];
  for (;;) {
    int64_t c = 2 * root + 1;
    if (c >= m) break;
    if (c + 1 < m && win_less(w[c], w[c + 1], edge_target)) ++c;
    if (!win_less(x, w[c], edge_target)) break;
    w[root] = w[c];
    root = c;
  }
  w[root] = x;
}

// Sorts one bucket's events w[0, m) by win_less: insertion sort for short lists,
// heapsort (O(m log m), in place) for long ones. win_less is a strict total
// order (tickets are unique), so both give the one sorted order.
__device__ inline void sort_bucket(QEvent* w, int64_t m, const int32_t* edge_target) {
  if (m <= 32) {
    for (int64_t i = 1; i < m; ++i) {
      const QEvent x = w[i];
      int64_t j = i - 1;
      while (j >= 0 && win_less(x, w[j], edge_target)) {
        w[j + 1] = w[j];
        --j;
      }
      w[j + 1] = x;
    }
    return;
  }
  for (int64_t i = m / 2 - 1; i >= 0; --i) sift_down(w, i, m, edge_target);
  for (int64_t e = m - 1; e > 0; --e) {
    const QEvent top = w[0];
    w[0] = w[e];
    w[e] = top;
    sift_down(w, 0, e, edge_target);
  }
}

__global__ void __launch_bounds__(kBlock) k_gather(Dev d, int32_t n) {
  __shared__ int64_t scan[kBlock / 64 + 1];
  const int tid = threadIdx.x;
  const int64_t k1 = d.clock->step;
  const int64_t nb = d.nbuckets;
  // 1. count events per bucket
  for (int64_t c = tid; c < nb; c += kBlock) d.seg_cursor[c] = 0;
  __syncthreads();
  for (int s = 0; s < n; ++s) {
    const int64_t b = (k1 + s) % d.ring_slots;
    // a writer that overflowed the bucket raised MIND_EQUEUE and stored nothing
    // past ring_cap: read only the stored events
    const int64_t cnt = std::min<int64_t>((int64_t)d.ring_count[b], d.ring_cap);
    for (int64_t i = tid; i < cnt; i += kBlock) {
      const int32_t cell = d.point_cell[d.edge_target[d.ring[b * d.ring_cap + i].edge]];
      atomicAdd(&d.seg_cursor[d.cell_bucket[cell]], 1ull);
    }
  }
  __syncthreads();
  // 2. exclusive scan over buckets
  int64_t running = 0;
  for (int64_t base = 0; base < nb; base += kBlock) {
    const int64_t c = base + tid;
    const int64_t x = c < nb ? (int64_t)d.seg_cursor[c] : 0;
    int64_t total;
    const int64_t pre = block_scan(x, scan, &total);
    if (c < nb) {
      d.seg_off[c] = running + pre;
      d.seg_cursor[c] = (unsigned long long)(running + pre);
    }
    running += total;
  }
  if (tid == 0) {
    d.seg_off[nb] = running;
    if (running <= d.win_cap) d.clock->delivered += running;
  }
  __syncthreads();
  if (running > d.win_cap) {
    if (tid == 0) raise_error(d.error, MIND_EQUEUE);
    for (int64_t c = tid; c <= nb; c += kBlock) d.seg_off[c] = 0;
    return;
  }
  // 3. scatter
  for (int s = 0; s < n; ++s) {
    const int64_t b = (k1 + s) % d.ring_slots;
    const int64_t cnt = std::min<int64_t>((int64_t)d.ring_count[b], d.ring_cap);
    for (int64_t i = tid; i < cnt; i += kBlock) {
      const QEvent ev = d.ring[b * d.ring_cap + i];
      const int32_t cell = d.point_cell[d.edge_target[ev.edge]];
      const unsigned long long pos = atomicAdd(&d.seg_cursor[d.cell_bucket[cell]], 1ull);
      d.win[pos] = ev;
    }
  }
  __syncthreads();
  // 4. deterministic order inside each bucket
  for (int64_t c = tid; c < nb; c += kBlock) {
    const int64_t a = d.seg_off[c];
    sort_bucket(d.win + a, d.seg_off[c + 1] - a, d.edge_target);
  }
  __syncthreads();
  for (int s = tid; s < n; s += kBlock) d.ring_count[(k1 + s) % d.ring_slots] = 0;
}

// ============================================================ host side
thread_local std::string g_error;
constexpr int64_t kMaxWindow = 1000;          // steps per window at most
constexpr int64_t kSerialWindow = 16;         // steps per replayed graph on the event-exact path
constexpr int64_t kRingEvents = int64_t(1) << 27;  // default ring bound: 4 GiB of 32-byte events

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

State::~State() { release(); }

namespace {


}  // namespace



// Event buckets of the window lists: S.cell_bucket maps each cell to one of nb buckets.
int set_buckets(State& S, int64_t nb) {
  Dev& d = S.d;
  int rc;
  int32_t* p32;
  if ((rc = S.upload(&p32, S.cell_bucket))) return rc;
  d.cell_bucket = p32;
  d.nbuckets = nb;
  if ((rc = S.alloc(&d.seg_off, nb + 1))) return rc;
  if ((rc = S.alloc(&d.seg_cursor, nb + 1))) return rc;
  HIP_TRY(hipMemset(d.seg_off, 0, sizeof(int64_t) * (nb + 1)));
  return 0;
}

int setup_state(State& S, const mind_build_options& o) {
  Layout& L = S.L;
  Dev& d = S.d;
  S.device = o.device;
  S.method = o.method;
  S.window_steps_opt = o.window_steps;
  HIP_TRY(hipSetDevice(o.device));
  HIP_TRY(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
  int rck = rtc_kernels(S);  // the cell kernels of a model with runtime-compiled mechanisms
  if (rck) return rck;
  d.dt = L.dt;
  d.celsius = L.celsius;
  d.seed = L.seed;
  d.ncells = L.ncells;
  d.nodes = L.nodes;
  int rc = 0;
  std::vector<double> cdt(L.nodes);
  for (int64_t i = 0; i < L.nodes; ++i) cdt[i] = L.cap[i] / L.assembly_dt();
  d.vscale = L.secondorder ? 2.0 : 1.0;
  {
    int64_t* p64; int32_t* p32; double* pd;
    if ((rc = S.upload(&p64, L.cell_off))) return rc; d.cell_off = p64;
    if ((rc = S.upload(&p32, L.parent))) return rc; d.parent = p32;
    if ((rc = S.upload(&pd, L.area))) return rc; d.area = pd;
    if ((rc = S.upload(&pd, cdt))) return rc; d.cdt = pd;
    if ((rc = S.upload(&pd, L.axial))) return rc; d.axial = pd;
    if ((rc = S.upload(&pd, L.adiag))) return rc; d.adiag = pd;
    if ((rc = S.upload(&pd, L.v_init))) return rc; S.v_init = pd;
    if ((rc = S.alloc(&d.v, L.nodes))) return rc;
    if ((rc = S.alloc(&d.rhs, L.nodes))) return rc;
    if ((rc = S.alloc(&d.diag, L.nodes))) return rc;
    if ((rc = S.upload(&p64, L.child_off))) return rc; d.child_off = p64;
    if ((rc = S.upload(&p32, L.child))) return rc; d.child = p32;
  }
  // mechanism groups
  d.ngroups = (int32_t)L.groups.size();
  d.ns_group = L.group_of_mech[MIND_MECH_NETSTIM];
  S.groups_host.resize(L.groups.size());
  S.group_initial.resize(L.groups.size());
  for (size_t gi = 0; gi < L.groups.size(); ++gi) {
    const MechGroup& G = L.groups[gi];
    GroupDev& H = S.groups_host[gi];
    H.mech = G.mech;
    H.point = mech_is_point(G.mech) ? 1 : 0;
    H.rows = G.rows;
    if ((rc = S.alloc(&H.data, G.data.size()))) return rc;
    if ((rc = S.upload(&S.group_initial[gi], G.data))) return rc;
    int32_t* p32; int64_t* p64; uint8_t* p8;
    if ((rc = S.upload(&p32, G.node))) return rc; H.node = p32;
    if ((rc = S.upload(&p64, G.cell_rows))) return rc; H.cell_rows = p64;
    if ((rc = S.upload(&p8, G.run_start))) return rc; H.run_start = p8;
    H.nslot = (int16_t)G.nslot;
    H.nread = (int16_t)G.nread;
    H.nwrite = (int16_t)G.nwrite;
    H.nrecv = (int16_t)G.nrecv;
    H.ion_row = nullptr;
    H.ion_bind = nullptr;
    H.ion_state = 0;
    H.ion_fast = 0;
    if (G.nslot) {
      bool reads_sums = false, currents_only = G.nwrite <= 2;
      for (int k = 0; k < G.nread; ++k)
        reads_sums |= G.ion_bind[k].column == MIND_ION_I || G.ion_bind[k].column == MIND_ION_DIDV;
      for (int k = G.nread; k < G.nread + G.nwrite; ++k)
        currents_only &= G.ion_bind[k].column == MIND_ION_I && G.ion_bind[k].aux < 2;
      H.ion_state = (L.ion_ordered || reads_sums) ? 1 : 0;
      H.ion_fast = (!H.point && G.nslot <= 2 && G.nread <= 2 && currents_only) ? 1 : 0;
      IonBind* pb;
      if ((rc = S.upload(&p32, G.ion_row))) return rc; H.ion_row = p32;
      if ((rc = S.upload(&pb, G.ion_bind))) return rc; H.ion_bind = pb;
    }
  }
  // ion rows
  d.nion = L.nion;
  d.ion_ordered = L.ion_ordered ? 1 : 0;
  d.ion_correct = L.secondorder == 2 ? 1 : 0;
  {
    std::vector<double> valence, conc0;
    for (const Species& sp : L.species) {
      valence.push_back(sp.valence);
      conc0.push_back(sp.internal);
      conc0.push_back(sp.external);
    }
    double* pd; int32_t* p32; int64_t* p64; uint8_t* p8;
    if ((rc = S.alloc(&d.ion, std::max<size_t>(1, L.ion_data.size())))) return rc;
    if ((rc = S.upload(&S.ion_initial, L.ion_data))) return rc;
    if ((rc = S.upload(&p8, L.ion_style))) return rc; d.ion_style = p8;
    d.ion_styles = 0;
    for (uint8_t st : L.ion_style) d.ion_styles |= st;
    if ((rc = S.upload(&p32, L.ion_species))) return rc; d.ion_species = p32;
    if ((rc = S.upload(&pd, valence))) return rc; d.ion_valence = pd;
    if ((rc = S.upload(&pd, conc0))) return rc; d.ion_conc0 = pd;
    if ((rc = S.upload(&p64, L.node_ion_off))) return rc; d.node_ion_off = p64;
  }
  {
    GroupDev* pg;
    if ((rc = S.upload(&pg, S.groups_host))) return rc;
    d.g = pg;
  }
  // the core window kernels take models whose groups all have inlined paths
  S.core = !S.ks && !L.ion_ordered && !(L.secondorder == 2 && L.nion);
  for (const GroupDev& H : S.groups_host) {
    const int m = H.mech;
    S.core = S.core && (m == MIND_MECH_PAS || m == MIND_MECH_HH || m == MIND_MECH_EXPSYN ||
                        m == MIND_MECH_EXP2SYN || m == MIND_MECH_ICLAMP || m == MIND_MECH_NETSTIM);
    if (H.nslot) S.core = S.core && m == MIND_MECH_HH && H.ion_fast && !H.ion_state && !H.nrecv;
    if (m == MIND_MECH_PAS || m == MIND_MECH_HH) S.core = S.core && !H.point;
  }
  {
    int32_t* p32;
    if ((rc = S.upload(&p32, L.point_group))) return rc; d.point_group = p32;
    if ((rc = S.upload(&p32, L.point_row))) return rc; d.point_row = p32;
    if ((rc = S.upload(&p32, L.point_cell))) return rc; d.point_cell = p32;
  }
  // detectors
  d.ndet = (int64_t)L.det_node.size();
  {
    int32_t* p32; int64_t* p64; double* pd;
    if ((rc = S.upload(&p32, L.det_node))) return rc; d.det_node = p32;
    if ((rc = S.upload(&pd, L.det_threshold))) return rc; d.det_thr = pd;
    if ((rc = S.alloc(&d.det_above, d.ndet))) return rc;
    if ((rc = S.upload(&p64, L.cell_det_off))) return rc; d.cell_det_off = p64;
    if ((rc = S.upload(&p32, L.cell_det))) return rc; d.cell_det = p32;
    if ((rc = S.upload(&p32, L.det_source))) return rc; d.det_source = p32;
  }
  // window length: no event produced inside a window may be due inside it
  const double min_delay = L.fanout[L.first_virtual] == 0 ? INFINITY : L.min_micro_delay;
  // at most kMaxWindow steps per window: detector flags, emission buffers and the
  // ring horizon grow with it
  int64_t wmax = kMaxWindow;
  if (std::isfinite(min_delay)) wmax = (int64_t)std::min(floor(min_delay / L.dt - 0.5), (double)kMaxWindow);
  if (o.window_steps > 0) wmax = std::min<int64_t>(wmax, o.window_steps);
  // delays shorter than 1.5 dt (so that an event can fall due in the step that
  // produced it) or connections into artificial cells (NetStim inputs are
  // feedback receivers) run on the event-exact path, one step per window
  bool into_artificial = false;
  for (int32_t t : L.edge_target) {
    if (L.point_group[t] < 0)  // PatternStim has no NET_RECEIVE
      return fail(MIND_EARG, "a connection targets a PatternStim, which receives no events");
    if (L.point_cell[t] < 0) into_artificial = true;
  }
  S.serial = wmax < 1 || into_artificial;
  if (S.serial) wmax = kSerialWindow;
  S.wmax = (int32_t)wmax;
  if ((rc = S.alloc(&d.det_flags, (size_t)d.ndet * wmax))) return rc;
  // edges
  {
    int64_t* p64; int32_t* p32; double* pd;
    if ((rc = S.upload(&p64, L.fanout))) return rc; d.fanout = p64;
    if ((rc = S.upload(&p32, L.edge_target))) return rc; d.edge_target = p32;
    if ((rc = S.upload(&pd, L.edge_delay))) return rc; d.edge_delay = pd;
    if ((rc = S.upload(&pd, L.edge_weight))) return rc; d.edge_weight = pd;
    if ((rc = S.upload(&p64, L.source_sid))) return rc; d.source_sid = p64;
  }
  // NetStim rows
  d.ns_rows = d.ns_group >= 0 ? L.groups[d.ns_group].rows : 0;
  if ((rc = S.alloc(&d.ns_stream, d.ns_rows))) return rc;
  if ((rc = S.alloc(&d.ns_t, d.ns_rows))) return rc;
  if ((rc = S.alloc(&d.ns_heap, d.ns_rows + kInjCap))) return rc;
  if ((rc = S.alloc(&d.ns_flag, d.ns_rows))) return rc;
  if ((rc = S.alloc(&d.ns_seq, d.ns_rows))) return rc;
  if ((rc = S.alloc(&d.ns_edge, d.ns_rows))) return rc;
  if ((rc = S.alloc(&d.ns_valid, d.ns_rows))) return rc;
  {
    int32_t* p32;
    if ((rc = S.upload(&p32, L.netstim_source))) return rc; d.ns_source = p32;
    std::vector<uint64_t> logical = d.ns_group >= 0 ? L.groups[d.ns_group].logical : std::vector<uint64_t>();
    if ((rc = S.upload(&S.ns_logical, logical))) return rc;
  }
  // delivery ring: one bucket per step up to the longest delay past a window
  const int64_t E = (int64_t)L.edge_target.size();
  d.ring_slots = (int64_t)ceil(L.max_delay / L.dt) + wmax + 4;
  // default capacity per step bucket: every edge firing in one step, bounded so
  // that the whole ring stays within kRingEvents events (long delays mean many
  // buckets); an overflowing bucket raises MIND_EQUEUE and event_capacity sets
  // the bucket size explicitly
  d.ring_cap = o.event_capacity > 0 ? o.event_capacity
                                    : std::min<int64_t>(std::max<int64_t>(E + 1024, 4096), 1 << 20);
  if (o.event_capacity <= 0 && d.ring_slots * d.ring_cap > kRingEvents)
    d.ring_cap = std::max<int64_t>(4096, kRingEvents / d.ring_slots);
  if ((rc = S.alloc(&d.ring, (size_t)d.ring_slots * d.ring_cap))) return rc;
  if ((rc = S.alloc(&d.ring_count, d.ring_slots))) return rc;
  HIP_TRY(hipMemset(d.ring_count, 0, sizeof(unsigned long long) * d.ring_slots));
  d.win_cap = std::min<int64_t>(d.ring_cap * wmax, (int64_t)1 << 26);
  if ((rc = S.alloc(&d.win, d.win_cap))) return rc;
  // event buckets: one per cell (the batched plan regroups them)
  S.cell_bucket.resize(L.ncells);
  std::iota(S.cell_bucket.begin(), S.cell_bucket.end(), 0);
  if ((rc = set_buckets(S, L.ncells))) return rc;
  d.emit_cap = d.ndet * wmax + 64 * d.ns_rows + 1024;
  if ((rc = S.alloc(&d.emit_det, d.emit_cap))) return rc;
  if ((rc = S.alloc(&d.emit_ns, d.emit_cap))) return rc;
  if ((rc = S.alloc(&d.emit_all, 2 * d.emit_cap))) return rc;
  if ((rc = S.alloc(&d.emit_base, 2 * d.emit_cap + 1))) return rc;
  d.spike_cap = 1 << 22;
  if ((rc = S.alloc(&d.spike_t, d.spike_cap))) return rc;
  if ((rc = S.alloc(&d.spike_sid, d.spike_cap))) return rc;
  if ((rc = S.alloc(&d.spike_count, 1))) return rc;
  if ((rc = S.alloc(&d.clock, 1))) return rc;
  if ((rc = S.alloc(&d.error, 1))) return rc;
  HIP_TRY(hipMemset(d.error, 0, sizeof(int32_t)));
  if ((rc = S.alloc(&d.inj, kInjCap))) return rc;
  if ((rc = S.alloc(&d.inj_count, 1))) return rc;
  HIP_TRY(hipMemset(d.inj_count, 0, sizeof(int32_t)));
  if ((rc = S.alloc(&S.inj_ev, kInjCap))) return rc;
  if (S.serial) {
    double* pd;
    int32_t* p32;
    if ((rc = S.upload(&pd, L.pattern_time))) return rc;
    S.pattern_time = pd;
    if ((rc = S.upload(&p32, L.pattern_source))) return rc;
    S.pattern_source = p32;
    const int64_t cap = o.event_capacity > 0 ? o.event_capacity : std::max<int64_t>(int64_t(1) << 20, 8 * E);
    if ((rc = serial_setup(S, cap))) return rc;
  }
  return 0;
}

namespace {
inline unsigned blocks_for(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }
}  // namespace

int check_device_error(State& S) {
  HIP_TRY(hipStreamSynchronize(S.stream));
  int32_t code = 0;
  HIP_TRY(hipMemcpy(&code, S.d.error, sizeof(code), hipMemcpyDeviceToHost));
  if (code == 0) return 0;
  S.initialized = false;
  switch (code) {
    case MIND_EEVENT: return fail(code, "invalid self-event delay or target");
    case MIND_ESINGULAR: return fail(code, "singular or non-finite cable system");
    case MIND_ENONFINITE: return fail(code, "non-finite numerical state");
    case MIND_EQUEUE: return fail(code, "event ring, window list or spike buffer capacity exceeded");
    case MIND_EUNSUPPORTED:
      return fail(code, "an event fell due inside the window that produced it");
    default: return fail(code, "device error " + std::to_string(code));
  }
}

// Copies the spikes recorded on the device and sorts them by (time, SID)
// (collect_result, src/runtime/recording.jl:279-283).
int collect_spikes(State& S, int64_t* n_out, double** times, int64_t** sids) {
  Dev& d = S.d;
  unsigned long long ns = 0;
  HIP_TRY(hipMemcpy(&ns, d.spike_count, sizeof(ns), hipMemcpyDeviceToHost));
  if ((int64_t)ns > d.spike_cap) ns = d.spike_cap;
  std::vector<double> st(ns);
  std::vector<int64_t> ss(ns);
  if (ns) {
    HIP_TRY(hipMemcpy(st.data(), d.spike_t, sizeof(double) * ns, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ss.data(), d.spike_sid, sizeof(int64_t) * ns, hipMemcpyDeviceToHost));
  }
  std::vector<int64_t> order(ns);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return st[a] < st[b] || (st[a] == st[b] && ss[a] < ss[b]);
  });
  *n_out = (int64_t)ns;
  *times = (double*)malloc(sizeof(double) * std::max<unsigned long long>(1, ns));
  *sids = (int64_t*)malloc(sizeof(int64_t) * std::max<unsigned long long>(1, ns));
  for (unsigned long long i = 0; i < ns; ++i) {
    (*times)[i] = st[order[i]];
    (*sids)[i] = ss[order[i]];
  }
  return 0;
}

int timer_start(State& S) {
  if (S.events_used + 2 > S.events.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      S.events.push_back(e);
    }
  }
  HIP_TRY(hipEventRecord(S.events[S.events_used], S.stream));
  return 0;
}

int timer_stop(State& S) {
  HIP_TRY(hipEventRecord(S.events[S.events_used + 1], S.stream));
  S.events_used += 2;
  S.kernel_launches += 1;
  return 0;
}

// Injected connection events (already ticketed) into their delivery-ring buckets.
__global__ void k_inject_events(Dev d, const QEvent* ev, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const QEvent e = ev[i];
  const int64_t b = (e.call / 2) % d.ring_slots;
  const unsigned long long pos = atomicAdd(&d.ring_count[b], 1ull);
  if ((int64_t)pos >= d.ring_cap) {
    raise_error(d.error, MIND_EQUEUE);
    return;
  }
  d.ring[b * d.ring_cap + (int64_t)pos] = e;
}

// Before a window of n steps from S.step: injected connection events whose step
// is inside the delivery ring's horizon go into the ring; injected source spikes
// due in the window go to the window's exchange. Nothing to do (and no
// synchronisation) without injected input.
int stage_injections(State& S, int32_t n) {
  if (S.pending_events.empty() && S.pending_spikes.empty()) return 0;
  Dev& d = S.d;
  HIP_TRY(hipStreamSynchronize(S.stream));
  std::vector<QEvent> now, later;
  for (const QEvent& e : S.pending_events) (e.call / 2 - S.step < d.ring_slots - 1 ? now : later).push_back(e);
  for (size_t a = 0; a < now.size(); a += kInjCap) {
    const int32_t m = (int32_t)std::min<size_t>(kInjCap, now.size() - a);
    HIP_TRY(hipMemcpy(S.inj_ev, now.data() + a, sizeof(QEvent) * m, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_inject_events, dim3(blocks_for(m, 256)), dim3(256), 0, S.stream, d, S.inj_ev, m);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(S.stream));  // the staging buffer is reused
  }
  S.pending_events.swap(later);
  const double cut_hi = call_cutoff(2 * (S.step + n) - 1, S.L.dt);
  std::vector<InjSpike> due, rest;
  for (const InjSpike& q : S.pending_spikes) (q.t <= cut_hi ? due : rest).push_back(q);
  if ((int64_t)due.size() > kInjCap) return fail(MIND_EQUEUE, "more injected spikes due in one window than it takes");
  if (!due.empty()) {
    const int32_t m = (int32_t)due.size();
    HIP_TRY(hipMemcpy(d.inj, due.data(), sizeof(InjSpike) * m, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.inj_count, &m, sizeof(int32_t), hipMemcpyHostToDevice));
  }
  S.pending_spikes.swap(rest);
  return 0;
}

// The longest window of at most n steps from S.step in which at most kInjCap
// injected source spikes fall due (a window hands its exchange at most that many);
// 0 if more than kInjCap fall due in the first step.
int32_t injection_window(const State& S, int32_t n) {
  if ((int64_t)S.pending_spikes.size() <= kInjCap) return n;
  std::vector<double> t;
  t.reserve(S.pending_spikes.size());
  for (const InjSpike& q : S.pending_spikes) t.push_back(q.t);
  std::nth_element(t.begin(), t.begin() + kInjCap, t.end());
  const double first_over = t[kInjCap];  // the earliest time by which more than kInjCap are due
  while (n > 0 && first_over <= call_cutoff(2 * (S.step + n) - 1, S.L.dt)) --n;
  return n;
}

// The cell kernel of n steps (plain or batched).
int cells_launch(State& S, int32_t n) {
  if (S.method == MIND_METHOD_BATCHED) return batched_launch(S, n);
  int rc = 0;
  if (S.profiling && (rc = timer_start(S))) return rc;
  HIP_TRY(launch(S, K_PLAIN, (const void*)k_plain, dim3(blocks_for(S.d.ncells, 64)), dim3(64), 0, S.stream, S.d, n));
  if (S.profiling && (rc = timer_stop(S))) return rc;
  if (S.profiling) S.kernel_steps += n;
  return 0;
}

int launch_window(State& S, int32_t n) {
  if (S.serial) return serial_launch(S, n);
  Dev& d = S.d;
  hipLaunchKernelGGL(k_gather, dim3(1), dim3(kBlock), 0, S.stream, d, n);
  int rc = cells_launch(S, n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_exchange, dim3(1), dim3(kBlock), 0, S.stream, d, n);
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace mind

using mind::State;
using mind::fail;

extern "C" {

const char* mind_last_error(void) { return mind::g_error.c_str(); }

const char* mind_build_info(void) { return mind::kMechanismSource; }

int mind_build(const mind_model_desc* model, const mind_build_options* options, mind_state** out) {
  if (!model || !options || !out) return fail(MIND_EARG, "null argument");
  if (options->secondorder < 0 || options->secondorder > 2) return fail(MIND_EARG, "secondorder must be 0, 1 or 2");
  State* S = new State();
  std::string err = mind::pack(*model, options->dt, options->seed, S->L);
  if (!err.empty()) {
    delete S;
    return fail(MIND_EARG, err);
  }
  S->L.secondorder = options->secondorder;
  int rc = mind::setup_state(*S, *options);
  if (!rc && options->method == MIND_METHOD_BATCHED) rc = mind::batched_setup(*S, options->cells_per_group);
  if (rc) {
    delete S;
    return rc;
  }
  *out = reinterpret_cast<mind_state*>(S);
  return 0;
}

void mind_state_destroy(mind_state* state) { delete reinterpret_cast<State*>(state); }

int mind_initialize(mind_state* state) {
  State& S = *reinterpret_cast<State*>(state);
  mind::Dev& d = S.d;
  HIP_TRY(hipSetDevice(S.device));
  S.initialized = false;
  S.step = 0;
  S.pending_events.clear();  // injected input does not survive initialize!
  S.pending_spikes.clear();
  // PatternStim spikes: source events enqueued before any INITIAL block runs
  // (reset_events!), so they take the first insertion orders
  const uint64_t npattern = S.L.pattern_time.size();
  if (!S.serial)  // (the event-exact path enqueues them itself)
    for (uint64_t k = 0; k < npattern; ++k)
      S.pending_spikes.push_back(mind::InjSpike{S.L.pattern_time[k], k + 1, S.L.pattern_source[k], 0});
  HIP_TRY(hipMemsetAsync(d.inj_count, 0, sizeof(int32_t), S.stream));
  HIP_TRY(hipMemsetAsync(d.error, 0, sizeof(int32_t), S.stream));
  mind::Clock clock{0, 0, npattern, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(d.clock, &clock, sizeof(clock), hipMemcpyHostToDevice, S.stream));
  HIP_TRY(hipMemsetAsync(d.ring_count, 0, sizeof(unsigned long long) * d.ring_slots, S.stream));
  for (size_t gi = 0; gi < S.groups_host.size(); ++gi)
    HIP_TRY(hipMemcpyAsync(S.groups_host[gi].data, S.group_initial[gi],
                           sizeof(double) * S.L.groups[gi].data.size(), hipMemcpyDeviceToDevice, S.stream));
  hipLaunchKernelGGL(mind::k_init_nodes, dim3(mind::blocks_for(std::max(d.nodes, d.ndet), 256)), dim3(256), 0,
                     S.stream, d, S.v_init);
  if (d.nion) {  // ion rows restored, concentrations reset (cinit), initial reversals
    HIP_TRY(hipMemcpyAsync(d.ion, S.ion_initial, sizeof(double) * S.L.ion_data.size(), hipMemcpyDeviceToDevice,
                           S.stream));
    hipLaunchKernelGGL(mind::k_init_ions, dim3(mind::blocks_for(d.nion, 256)), dim3(256), 0, S.stream, d);
  }
  for (int gi = 0; gi < d.ngroups; ++gi) {
    const mind::MechGroup& G = S.L.groups[gi];
    if (G.mech == MIND_MECH_NETSTIM) continue;
    HIP_TRY(mind::launch(S, mind::K_INIT_GROUP, (const void*)mind::k_init_group,
                         dim3(mind::blocks_for(G.rows, 256)), dim3(256), 0, S.stream, d, gi));
    bool writes_concentration = false;
    for (int k = G.nread; k < G.nread + G.nwrite; ++k)
      writes_concentration |= G.ion_bind[k].column == MIND_ION_CI || G.ion_bind[k].column == MIND_ION_CO;
    if (writes_concentration)
      hipLaunchKernelGGL(mind::k_init_reversals, dim3(mind::blocks_for(G.rows, 256)), dim3(256), 0, S.stream, d,
                         gi);
  }
  if (d.ns_rows > 0 && !S.serial) {
    hipLaunchKernelGGL(mind::k_init_netstim, dim3(mind::blocks_for(d.ns_rows, 256)), dim3(256), 0, S.stream, d,
                       S.ns_logical, mind::kNetStimDigest, npattern);
    // the netstim sequence counter continues after the INITIAL insertions
    mind::Clock c2{0, 0, npattern + (uint64_t)d.ns_rows, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(d.clock, &c2, sizeof(c2), hipMemcpyHostToDevice, S.stream));
  }
  // Spikes of the time-zero delivery are kept for the first run (step.jl:47-49).
  d.record_spikes = 1;
  HIP_TRY(hipMemsetAsync(d.spike_count, 0, sizeof(unsigned long long), S.stream));
  // time-zero delivery (NetStim events and pattern spikes at t <= 0), then currents at t = 0
  if (S.serial) {
    if (int rcs = mind::serial_initialize(S)) return rcs;
  } else {
    if (int rcs = mind::stage_injections(S, 0)) return rcs;
    hipLaunchKernelGGL(mind::k_exchange, dim3(1), dim3(mind::kBlock), 0, S.stream, d, 0);
  }
  if (d.nion)
    hipLaunchKernelGGL(mind::k_prepare_ions, dim3(mind::blocks_for(d.nion, 256)), dim3(256), 0, S.stream, d);
  for (int gi = 0; gi < d.ngroups; ++gi) {
    if (S.L.groups[gi].mech == MIND_MECH_NETSTIM) continue;
    HIP_TRY(mind::launch(S, mind::K_CURRENTS_FIELDS, (const void*)mind::k_currents_fields,
                         dim3(mind::blocks_for(S.L.groups[gi].rows, 256)), dim3(256), 0, S.stream, d, gi, 0.0));
  }
  HIP_TRY(hipGetLastError());
  d.record_spikes = 0;
  int rc = mind::macro_initialize(S);
  if (!rc) rc = mind::check_device_error(S);
  if (rc) return rc;
  S.initialized = true;
  return 0;
}

static int resolve_records(State& S, const mind_record* recs, int64_t n, std::vector<int32_t>& kind,
                           std::vector<int32_t>& group, std::vector<int32_t>& field, std::vector<int64_t>& index,
                           std::vector<int32_t>& cell) {
  const mind::Layout& L = S.L;
  for (int64_t q = 0; q < n; ++q) {
    const mind_record& r = recs[q];
    if (r.kind == 0) {
      if (r.cell < 0 || r.cell >= L.ncells) return fail(MIND_EARG, "record on an unknown cell");
      kind.push_back(0); group.push_back(-1); field.push_back(-1);
      index.push_back(L.location_node(r.cell, r.section, r.x));
      cell.push_back(r.cell);
    } else if (r.kind == 1) {
      if (r.point < 0 || r.point >= (int32_t)L.point_group.size()) return fail(MIND_EARG, "unknown point process");
      const int g = L.point_group[r.point];
      if (g < 0) return fail(MIND_EARG, "a PatternStim has no fields to record");
      if (r.field < 0 || r.field >= mind::field_count(L.groups[g].mech)) return fail(MIND_EARG, "unknown field");
      kind.push_back(1); group.push_back(g); field.push_back(r.field);
      index.push_back(L.point_row[r.point]);
      cell.push_back(L.point_cell[r.point]);
    } else if (r.kind == 2) {
      if (r.mech < 0 || r.mech >= MIND_MECH_SLOTS || L.group_of_mech[r.mech] < 0)
        return fail(MIND_EARG, "mechanism not in model");
      const int g = L.group_of_mech[r.mech];
      const int64_t node = L.range_node(r.cell, r.section, r.x);
      const std::vector<int32_t>& nodes = L.groups[g].node;
      auto it = std::lower_bound(nodes.begin(), nodes.end(), (int32_t)node);
      if (it == nodes.end() || *it != node)
        return fail(MIND_EARG, "location must contain exactly one matching density mechanism");
      if (r.field < 0 || r.field >= mind::field_count(r.mech)) return fail(MIND_EARG, "unknown field");
      kind.push_back(1); group.push_back(g); field.push_back(r.field);
      index.push_back(it - nodes.begin());
      cell.push_back(r.cell);
    } else if (r.kind == 3) {  // ion_ref (references.jl:156-178)
      if (r.cell < 0 || r.cell >= L.ncells) return fail(MIND_EARG, "record on an unknown cell");
      if (r.mech < 0 || r.mech >= (int32_t)L.species.size()) return fail(MIND_EARG, "unknown ion species");
      if (r.field < MIND_ION_E || r.field > MIND_ION_DIDV) return fail(MIND_EARG, "unknown ion field");
      const int64_t q = L.ion_row_at(L.range_node(r.cell, r.section, r.x), r.mech);
      if (q < 0) return fail(MIND_EARG, "ion " + L.species[r.mech].name + " is not present at this location");
      kind.push_back(3); group.push_back(-1); field.push_back(r.field);
      index.push_back(q);
      cell.push_back(r.cell);
    } else {
      return fail(MIND_EARG, "unknown record kind");
    }
  }
  return 0;
}

int mind_run(mind_state* state, double stop, const mind_record* records, int64_t n_records,
             int32_t record_spikes, mind_result* out) {
  if (!state || !out) return fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  mind::Dev& d = S.d;
  HIP_TRY(hipSetDevice(S.device));
  if (!S.initialized) return fail(MIND_EARG, "simulation is not initialized");
  const double dt = S.L.dt;
  const double now = (double)S.step * dt;
  const double duration = stop - now;
  if (!std::isfinite(duration)) return fail(MIND_EARG, "stop time must be finite");
  const int64_t count = (int64_t)nearbyint(duration / dt);
  const double big = std::max(std::max(fabs(stop), fabs(now)), dt);
  const double tol = 64.0 * (nextafter(big, INFINITY) - big);
  if (!(count >= 0 && fabs(duration - (double)count * dt) <= tol))
    return fail(MIND_EARG, "stop time must be an integer number of steps after the current time");
  // records
  std::vector<int32_t> kind, group, field, cell;
  std::vector<int64_t> index;
  int rc = resolve_records(S, records, n_records, kind, group, field, index, cell);
  if (rc) return rc;
  for (void* p : S.run_allocations) hipFree(p);
  S.run_allocations.clear();
  auto run_upload = [&](auto** dst, const auto& host) -> int {
    using T = typename std::remove_reference<decltype(host)>::type::value_type;
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(1, host.size()) * sizeof(T)) != hipSuccess)
      return fail(MIND_EOOM, "device allocation failed");
    S.run_allocations.push_back(p);
    if (!host.empty() && hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      return fail(MIND_EHIP, "record upload failed");
    *dst = static_cast<T*>(p);
    return 0;
  };
  const int64_t samples = count + 1;
  // records of cells, by cell (sampled by the cell kernels); records of
  // artificial cells (cell -1), sampled by the event kernels
  std::vector<int64_t> cell_rec_off(S.L.ncells + 1, 0);
  std::vector<int32_t> cell_rec, art_rec;
  for (int64_t q = 0; q < n_records; ++q)
    if (cell[q] >= 0) cell_rec_off[cell[q] + 1] += 1;
    else art_rec.push_back((int32_t)q);
  for (int c = 0; c < S.L.ncells; ++c) cell_rec_off[c + 1] += cell_rec_off[c];
  cell_rec.resize(cell_rec_off[S.L.ncells]);
  {
    std::vector<int64_t> cur(cell_rec_off.begin(), cell_rec_off.end() - 1);
    for (int64_t q = 0; q < n_records; ++q)
      if (cell[q] >= 0) cell_rec[cur[cell[q]]++] = (int32_t)q;
  }
  {
    int32_t* p32; int64_t* p64;
    if ((rc = run_upload(&p32, kind))) return rc; d.rec_kind = p32;
    if ((rc = run_upload(&p32, group))) return rc; d.rec_group = p32;
    if ((rc = run_upload(&p32, field))) return rc; d.rec_field = p32;
    if ((rc = run_upload(&p64, index))) return rc; d.rec_index = p64;
    if ((rc = run_upload(&p64, cell_rec_off))) return rc; d.cell_rec_off = p64;
    if ((rc = run_upload(&p32, cell_rec))) return rc; d.cell_rec = p32;
    if ((rc = run_upload(&p32, art_rec))) return rc; d.art_rec = p32;
    d.n_art_rec = (int64_t)art_rec.size();
    void* p = nullptr;
    if (hipMalloc(&p, std::max<int64_t>(1, n_records * samples) * sizeof(double)) != hipSuccess)
      return fail(MIND_EOOM, "device allocation failed");
    S.run_allocations.push_back(p);
    d.rec_buf = static_cast<double*>(p);
  }
  d.nrec = n_records;
  d.rec_stride = samples;
  d.rec_start = S.step;
  d.record_spikes = record_spikes ? 1 : 0;
  if (S.step != 0 || !record_spikes)
    HIP_TRY(hipMemsetAsync(d.spike_count, 0, sizeof(unsigned long long), S.stream));
  if (S.graph) {  // kernel arguments changed: rebuild the window graph
    hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
  }
  if (n_records)
    hipLaunchKernelGGL(mind::k_record_sample, dim3(mind::blocks_for(n_records, 256)), dim3(256), 0, S.stream, d,
                       (int64_t)0);
  HIP_TRY(hipStreamSynchronize(S.stream));
  const auto t0 = std::chrono::steady_clock::now();
  int64_t left = count;
  while (left > 0) {
    // a window short enough that the injected spikes due in it fit its exchange
    const int32_t n = mind::injection_window(S, (int32_t)std::min<int64_t>(left, S.wmax));
    if (n == 0) return fail(MIND_EQUEUE, "more injected spikes due in one step than a window takes");
    if ((rc = mind::stage_injections(S, n))) return rc;
    if (n == S.wmax && !S.profiling) {
      if (!S.graph) {
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
        rc = mind::launch_window(S, n);
        hipError_t e = hipStreamEndCapture(S.stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return fail(MIND_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
        HIP_TRY(hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0));
        hipGraphDestroy(g);
      }
      HIP_TRY(hipGraphLaunch(S.graph, S.stream));
    } else {
      rc = mind::launch_window(S, n);
      if (rc) return rc;
    }
    left -= n;
    S.step += n;
  }
  HIP_TRY(hipStreamSynchronize(S.stream));
  const auto t1 = std::chrono::steady_clock::now();
  for (size_t e = 0; e + 1 < S.events_used; e += 2) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, S.events[e], S.events[e + 1]));
    S.kernel_ms += ms;
  }
  S.events_used = 0;
  rc = mind::check_device_error(S);
  d.record_spikes = 0;
  if (rc) return rc;
  if (S.graph) {  // the graph holds this run's record buffers
    hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
  }
  // results
  out->n_samples = samples;
  out->time = (double*)malloc(sizeof(double) * samples);
  for (int64_t j = 0; j < samples; ++j) out->time[j] = (double)(d.rec_start + j) * dt;
  out->n_records = n_records;
  out->traces = (double*)malloc(sizeof(double) * std::max<int64_t>(1, n_records * samples));
  if (n_records)
    HIP_TRY(hipMemcpy(out->traces, d.rec_buf, sizeof(double) * n_records * samples, hipMemcpyDeviceToHost));
  rc = mind::collect_spikes(S, &out->n_spikes, &out->spike_times, &out->spike_sids);
  if (rc) return rc;
  out->run_seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

int mind_step(mind_state* state) {
  State& S = *reinterpret_cast<State*>(state);
  HIP_TRY(hipSetDevice(S.device));
  if (!S.initialized) return fail(MIND_EARG, "simulation is not initialized");
  S.d.nrec = 0;
  S.d.n_art_rec = 0;
  S.d.record_spikes = 0;
  if (S.graph) {
    hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
  }
  int rc = mind::stage_injections(S, 1);
  if (!rc) rc = mind::launch_window(S, 1);
  if (rc) return rc;
  S.step += 1;
  return mind::check_device_error(S);
}

void mind_result_free(mind_result* r) {
  if (!r) return;
  free(r->time);
  free(r->traces);
  free(r->spike_times);
  free(r->spike_sids);
  memset(r, 0, sizeof(*r));
}

int mind_node_count(const mind_state* state, int64_t* nodes, int64_t* compartments) {
  const State& S = *reinterpret_cast<const State*>(state);
  *nodes = S.L.nodes;
  *compartments = S.L.compartments;
  return 0;
}

int mind_read_voltage(mind_state* state, double* out, int64_t n) {
  State& S = *reinterpret_cast<State*>(state);
  if (n != S.L.nodes) return fail(MIND_EARG, "size mismatch");
  HIP_TRY(hipStreamSynchronize(S.stream));
  HIP_TRY(hipMemcpy(out, S.d.v, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int mind_row_count(const mind_state* state, int32_t mech, int64_t* rows) {
  const State& S = *reinterpret_cast<const State*>(state);
  if (mech < 0 || mech >= MIND_MECH_SLOTS || S.L.group_of_mech[mech] < 0) return fail(MIND_EARG, "mechanism not in model");
  *rows = S.L.groups[S.L.group_of_mech[mech]].rows;
  return 0;
}

int mind_read_field(mind_state* state, int32_t mech, int32_t field, double* out, int64_t n) {
  State& S = *reinterpret_cast<State*>(state);
  if (mech < 0 || mech >= MIND_MECH_SLOTS || S.L.group_of_mech[mech] < 0) return fail(MIND_EARG, "mechanism not in model");
  const int g = S.L.group_of_mech[mech];
  const int64_t rows = S.L.groups[g].rows;
  if (n != rows || field < 0 || field >= mind::field_count(mech)) return fail(MIND_EARG, "size or field mismatch");
  HIP_TRY(hipStreamSynchronize(S.stream));
  HIP_TRY(hipMemcpy(out, S.groups_host[g].data + (size_t)field * rows, sizeof(double) * rows, hipMemcpyDeviceToHost));
  return 0;
}

// read_value (state.jl:586-600): one resolved scalar, read from the device after
// the state's stream has drained. Point processes of artificial cells (NetStim)
// are read from their row as well.
int mind_read_value(mind_state* state, const mind_record* record, double* out) {
  if (!state || !record || !out) return fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  if (!S.initialized) return fail(MIND_EARG, "simulation is not initialized");
  const mind::Layout& L = S.L;
  int32_t kind = 0, g = -1, field = -1;
  int64_t index = -1;
  if (record->kind == 1 && record->point >= 0 && record->point < (int32_t)L.point_group.size() &&
      L.point_cell[record->point] < 0) {
    g = L.point_group[record->point];
    if (g < 0) return fail(MIND_EARG, "PatternStim has no fields");
    if (record->field < 0 || record->field >= mind::field_count(L.groups[g].mech)) return fail(MIND_EARG, "unknown field");
    kind = 1;
    field = record->field;
    index = L.point_row[record->point];
  } else {
    std::vector<int32_t> kv, gv, fv, cv;
    std::vector<int64_t> iv;
    const int rc = resolve_records(S, record, 1, kv, gv, fv, iv, cv);
    if (rc) return rc;
    kind = kv[0];
    g = gv[0];
    field = fv[0];
    index = iv[0];
  }
  HIP_TRY(hipSetDevice(S.device));
  HIP_TRY(hipStreamSynchronize(S.stream));
  const double* src = kind == 0   ? S.d.v + index
                      : kind == 3 ? S.d.ion + (size_t)field * L.nion + index
                                  : S.groups_host[g].data + (size_t)field * L.groups[g].rows + index;
  HIP_TRY(hipMemcpy(out, src, sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// connection_weights (state.jl:603-616): the NET_RECEIVE arguments of a
// connection (0-based creation index). The six mechanisms take one argument,
// the weight, and none of them writes it, so it is the built weight.
int mind_connection_weights(const mind_state* state, int64_t connection, double* out, int32_t capacity,
                            int32_t* count) {
  if (!state || !count) return fail(MIND_EARG, "null argument");
  const State& S = *reinterpret_cast<const State*>(state);
  const mind::Layout& L = S.L;
  if (connection < 0 || connection >= (int64_t)L.connection_edge.size()) return fail(MIND_EARG, "unknown connection");
  *count = 1;
  if (capacity < 1 || !out) return fail(MIND_EARG, "weight buffer too small");
  out[0] = L.edge_weight[L.connection_edge[connection]];
  return 0;
}

// inject_event! (state.jl:618-636): one event of connection `connection` (0-based
// creation order) at absolute time `time`, without its delay, ticketed now like
// any enqueued event; it enters the delivery ring before the window that is
// the first to reach the ring's horizon at its step.
int mind_inject_event(mind_state* state, int64_t connection, double time) {
  if (!state) return fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  if (!S.initialized) return fail(MIND_EARG, "simulation is not initialized");
  const mind::Layout& L = S.L;
  if (connection < 0 || connection >= (int64_t)L.connection_edge.size()) return fail(MIND_EARG, "unknown connection");
  const double now = (double)S.step * L.dt;
  if (!(std::isfinite(time) && time >= now)) return fail(MIND_EARG, "event arrival precedes the current time");
  const int64_t edge = L.connection_edge[connection];
  HIP_TRY(hipSetDevice(S.device));
  if (S.serial)  // enqueued now, with the next insertion order
    return mind::serial_inject(S, mind::SEvent{time, 0.0, edge + 1, 0, L.edge_target[edge] + 1, 0});
  HIP_TRY(hipStreamSynchronize(S.stream));
  mind::Clock c;
  HIP_TRY(hipMemcpy(&c, S.d.clock, sizeof(c), hipMemcpyDeviceToHost));
  c.ticket += 1;
  HIP_TRY(hipMemcpy(S.d.clock, &c, sizeof(c), hipMemcpyHostToDevice));
  S.pending_events.push_back(mind::QEvent{time, c.ticket, mind::call_of(time, 2 * S.step, L.dt), (int32_t)edge, 0});
  return 0;
}

// inject_spike! (state.jl:638-659): a spike of external source `source` at
// absolute time `time`; it is ordered with NetStim events by (time, insertion
// order), routed over the source's connections at its delivery call and
// recorded like any spike.
int mind_inject_spike(mind_state* state, int64_t source, double time) {
  if (!state) return fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  if (!S.initialized) return fail(MIND_EARG, "simulation is not initialized");
  const mind::Layout& L = S.L;
  if (source < 0 || source >= (int64_t)L.source_external.size() || !L.source_external[source])
    return fail(MIND_EARG, "source is not a built external spike input");
  const double now = (double)S.step * L.dt;
  if (!(std::isfinite(time) && time >= now)) return fail(MIND_EARG, "spike time precedes the current time");
  HIP_TRY(hipSetDevice(S.device));
  if (S.serial) return mind::serial_inject(S, mind::SEvent{time, 0.0, 0, 0, -(int32_t)(source + 1), 0});
  HIP_TRY(hipStreamSynchronize(S.stream));
  mind::Clock c;
  HIP_TRY(hipMemcpy(&c, S.d.clock, sizeof(c), hipMemcpyDeviceToHost));
  c.ns_seq += 1;
  HIP_TRY(hipMemcpy(S.d.clock, &c, sizeof(c), hipMemcpyHostToDevice));
  S.pending_spikes.push_back(mind::InjSpike{time, c.ns_seq, (int32_t)source, 0});
  return 0;
}

}  // extern "C"

namespace {

// Checkpoint blob: a header, then the state a run continues from, in this order:
// v; each mechanism group's data; detector flags; the NetStim rows' streams,
// pending self-events and flags; the clock; the non-empty delivery-ring buckets
// (index, count, events); injected input not yet on the device. Scratch that
// every step rebuilds (rhs, diag, window lists, emissions) is not saved.
constexpr char kCkptMagic[8] = {'M', 'I', 'N', 'D', 'C', 'K', 'P', '1'};

struct CkptHeader {
  char magic[8];
  uint64_t fingerprint;  // of the packed model and the ring's shape
  int64_t step;
  int64_t bytes;         // the whole blob
};

uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 0x100000001b3ull;
  return h;
}

uint64_t fingerprint(const State& S) {
  const mind::Layout& L = S.L;
  uint64_t h = 0xcbf29ce484222325ull;
  h = fnv(h, &L.nodes, sizeof L.nodes);
  h = fnv(h, &L.dt, sizeof L.dt);
  h = fnv(h, &L.secondorder, sizeof L.secondorder);
  h = fnv(h, L.parent.data(), L.parent.size() * sizeof(int32_t));
  h = fnv(h, L.axial.data(), L.axial.size() * sizeof(double));
  for (const auto& g : L.groups) {
    h = fnv(h, &g.mech, sizeof g.mech);
    h = fnv(h, &g.rows, sizeof g.rows);
  }
  h = fnv(h, L.edge_target.data(), L.edge_target.size() * sizeof(int32_t));
  h = fnv(h, L.edge_delay.data(), L.edge_delay.size() * sizeof(double));
  h = fnv(h, &S.d.ndet, sizeof S.d.ndet);
  h = fnv(h, &L.nion, sizeof L.nion);
  h = fnv(h, &S.d.ring_slots, sizeof S.d.ring_slots);
  h = fnv(h, &S.d.ring_cap, sizeof S.d.ring_cap);
  return h;
}

// The device arrays of a checkpoint as (pointer, bytes) pairs, in blob order.
std::vector<std::pair<void*, size_t>> ckpt_arrays(State& S) {
  mind::Dev& d = S.d;
  std::vector<std::pair<void*, size_t>> a;
  a.push_back({d.v, sizeof(double) * S.L.nodes});
  for (size_t gi = 0; gi < S.groups_host.size(); ++gi)
    a.push_back({S.groups_host[gi].data, sizeof(double) * S.L.groups[gi].data.size()});
  a.push_back({d.det_above, sizeof(uint8_t) * d.ndet});
  a.push_back({d.ion, sizeof(double) * S.L.ion_data.size()});
  a.push_back({d.ns_stream, sizeof(mind::Stream) * d.ns_rows});
  a.push_back({d.ns_t, sizeof(double) * d.ns_rows});
  a.push_back({d.ns_flag, sizeof(double) * d.ns_rows});
  a.push_back({d.ns_seq, sizeof(uint64_t) * d.ns_rows});
  a.push_back({d.ns_edge, sizeof(int32_t) * d.ns_rows});
  a.push_back({d.ns_valid, sizeof(int32_t) * d.ns_rows});
  a.push_back({d.clock, sizeof(mind::Clock)});
  if (S.serial) {  // the event-exact path's queue counters (its heap follows the arrays)
    a.push_back({S.q.sequences, sizeof(int64_t) * S.L.point_group.size()});
    a.push_back({S.q.heap_n, sizeof(int64_t)});
  }
  return a;
}

template <typename T>
void put(std::vector<uint8_t>& b, const T* p, size_t n) {
  const uint8_t* c = reinterpret_cast<const uint8_t*>(p);
  b.insert(b.end(), c, c + sizeof(T) * n);
}

}  // namespace

extern "C" {

// Checkpoint of a micro state between runs (no counterpart in the reference).
int mind_checkpoint(mind_state* state, void* buffer, size_t capacity, size_t* size) {
  if (!state || !size) return fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  if (!S.initialized) return fail(MIND_EARG, "simulation is not initialized");
  if (S.macro) return fail(MIND_EUNSUPPORTED, "checkpoints of micro-macro states are not supported");
  mind::Dev& d = S.d;
  HIP_TRY(hipSetDevice(S.device));
  HIP_TRY(hipStreamSynchronize(S.stream));
  std::vector<uint8_t> blob(sizeof(CkptHeader));
  for (const auto& [p, n] : ckpt_arrays(S)) {
    const size_t at = blob.size();
    blob.resize(at + n);
    if (n) HIP_TRY(hipMemcpy(blob.data() + at, p, n, hipMemcpyDeviceToHost));
  }
  if (S.serial) {
    int64_t hn = 0;
    HIP_TRY(hipMemcpy(&hn, S.q.heap_n, sizeof hn, hipMemcpyDeviceToHost));
    std::vector<mind::SEvent> heap(hn);
    if (hn) HIP_TRY(hipMemcpy(heap.data(), S.q.heap, sizeof(mind::SEvent) * hn, hipMemcpyDeviceToHost));
    put(blob, heap.data(), hn);
  }
  std::vector<unsigned long long> counts(d.ring_slots);
  HIP_TRY(hipMemcpy(counts.data(), d.ring_count, sizeof(unsigned long long) * d.ring_slots, hipMemcpyDeviceToHost));
  std::vector<mind::QEvent> ev;
  for (int64_t b = 0; b < d.ring_slots; ++b) {
    const int64_t n = (int64_t)std::min<unsigned long long>(counts[b], (unsigned long long)d.ring_cap);
    if (!n) continue;
    ev.resize(n);
    HIP_TRY(hipMemcpy(ev.data(), d.ring + b * d.ring_cap, sizeof(mind::QEvent) * n, hipMemcpyDeviceToHost));
    put(blob, &b, 1);
    put(blob, &n, 1);
    put(blob, ev.data(), n);
  }
  const int64_t end = -1, npe = (int64_t)S.pending_events.size(), nps = (int64_t)S.pending_spikes.size();
  put(blob, &end, 1);
  put(blob, &npe, 1);
  put(blob, S.pending_events.data(), npe);
  put(blob, &nps, 1);
  put(blob, S.pending_spikes.data(), nps);
  CkptHeader h;
  std::memcpy(h.magic, kCkptMagic, sizeof h.magic);
  h.fingerprint = fingerprint(S);
  h.step = S.step;
  h.bytes = (int64_t)blob.size();
  std::memcpy(blob.data(), &h, sizeof h);
  *size = blob.size();
  if (!buffer) return 0;
  if (capacity < blob.size()) return fail(MIND_EARG, "checkpoint buffer too small");
  std::memcpy(buffer, blob.data(), blob.size());
  return 0;
}

// Restore into a state built from the same model with the same options.
int mind_restore(mind_state* state, const void* buffer, size_t size) {
  if (!state || !buffer) return fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  if (S.macro) return fail(MIND_EUNSUPPORTED, "checkpoints of micro-macro states are not supported");
  const uint8_t* b = static_cast<const uint8_t*>(buffer);
  CkptHeader h;
  if (size < sizeof h) return fail(MIND_EARG, "not a checkpoint");
  std::memcpy(&h, b, sizeof h);
  if (std::memcmp(h.magic, kCkptMagic, sizeof h.magic) || h.bytes != (int64_t)size)
    return fail(MIND_EARG, "not a checkpoint, or a truncated one");
  if (h.fingerprint != fingerprint(S)) return fail(MIND_EARG, "checkpoint of another model or build");
  mind::Dev& d = S.d;
  HIP_TRY(hipSetDevice(S.device));
  HIP_TRY(hipStreamSynchronize(S.stream));
  size_t at = sizeof h;
  auto take = [&](void* dst, size_t n) -> bool {
    if (at + n > size) return false;
    std::memcpy(dst, b + at, n);
    at += n;
    return true;
  };
  for (const auto& [p, n] : ckpt_arrays(S)) {
    if (at + n > size) return fail(MIND_EARG, "truncated checkpoint");
    if (n) HIP_TRY(hipMemcpy(p, b + at, n, hipMemcpyHostToDevice));
    at += n;
  }
  if (S.serial) {  // heap_n came with the arrays
    int64_t hn = 0;
    HIP_TRY(hipMemcpy(&hn, S.q.heap_n, sizeof hn, hipMemcpyDeviceToHost));
    if (hn < 0 || hn > S.q.heap_cap) return fail(MIND_EARG, "corrupt checkpoint");
    std::vector<mind::SEvent> heap(hn);
    if (!take(heap.data(), sizeof(mind::SEvent) * hn)) return fail(MIND_EARG, "truncated checkpoint");
    if (hn) HIP_TRY(hipMemcpy(S.q.heap, heap.data(), sizeof(mind::SEvent) * hn, hipMemcpyHostToDevice));
  }
  std::vector<unsigned long long> counts(d.ring_slots, 0);
  std::vector<mind::QEvent> ev;
  for (;;) {
    int64_t bucket = 0, n = 0;
    if (!take(&bucket, sizeof bucket)) return fail(MIND_EARG, "truncated checkpoint");
    if (bucket < 0) break;
    if (bucket >= d.ring_slots || !take(&n, sizeof n) || n <= 0 || n > d.ring_cap)
      return fail(MIND_EARG, "corrupt checkpoint");
    ev.resize(n);
    if (!take(ev.data(), sizeof(mind::QEvent) * n)) return fail(MIND_EARG, "truncated checkpoint");
    HIP_TRY(hipMemcpy(d.ring + bucket * d.ring_cap, ev.data(), sizeof(mind::QEvent) * n, hipMemcpyHostToDevice));
    counts[bucket] = (unsigned long long)n;
  }
  int64_t npe = 0, nps = 0;
  std::vector<mind::QEvent> pe;
  std::vector<mind::InjSpike> ps;
  if (!take(&npe, sizeof npe) || npe < 0) return fail(MIND_EARG, "corrupt checkpoint");
  pe.resize(npe);
  if (!take(pe.data(), sizeof(mind::QEvent) * npe) || !take(&nps, sizeof nps) || nps < 0)
    return fail(MIND_EARG, "corrupt checkpoint");
  ps.resize(nps);
  if (!take(ps.data(), sizeof(mind::InjSpike) * nps) || at != size) return fail(MIND_EARG, "corrupt checkpoint");
  HIP_TRY(hipMemcpy(d.ring_count, counts.data(), sizeof(unsigned long long) * d.ring_slots, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(d.inj_count, 0, sizeof(int32_t)));
  HIP_TRY(hipMemset(d.error, 0, sizeof(int32_t)));
  HIP_TRY(hipMemset(d.spike_count, 0, sizeof(unsigned long long)));
  S.pending_events = std::move(pe);
  S.pending_spikes = std::move(ps);
  S.step = h.step;
  S.initialized = true;
  if (S.graph) {
    (void)hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
  }
  return 0;
}

// Connection events routed (ticketed) and gathered for delivery since
// initialize, and spikes recorded by the last run (a device read; synchronises).
int mind_event_counts(mind_state* state, int64_t* routed, int64_t* delivered) {
  if (!state || !routed || !delivered) return fail(MIND_EARG, "null argument");
  State& S = *reinterpret_cast<State*>(state);
  HIP_TRY(hipSetDevice(S.device));
  HIP_TRY(hipStreamSynchronize(S.stream));
  mind::Clock c;
  HIP_TRY(hipMemcpy(&c, S.d.clock, sizeof(c), hipMemcpyDeviceToHost));
  *routed = (int64_t)c.ticket;
  *delivered = c.delivered;
  return 0;
}

int mind_time(const mind_state* state, double* t, int64_t* step) {
  const State& S = *reinterpret_cast<const State*>(state);
  *step = S.step;
  *t = (double)S.step * S.L.dt;
  return 0;
}

int mind_pack_nodes(const mind_model_desc* model, double dt, int64_t* nodes, int64_t* compartments,
                    int32_t* parent, double* area, double* cap, double* axial, double* axial_diag) {
  if (!model || !nodes || !compartments) return fail(MIND_EARG, "null argument");
  mind::Layout L;
  std::string err = mind::pack(*model, dt, 1, L);
  if (!err.empty()) return fail(MIND_EARG, err);
  *nodes = L.nodes;
  *compartments = L.compartments;
  if (parent) {
    std::copy(L.parent.begin(), L.parent.end(), parent);
    std::copy(L.area.begin(), L.area.end(), area);
    std::copy(L.cap.begin(), L.cap.end(), cap);
    std::copy(L.axial.begin(), L.axial.end(), axial);
    std::copy(L.adiag.begin(), L.adiag.end(), axial_diag);
  }
  return 0;
}

int mind_device_count(int32_t* count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return 0;
}

// Plain streaming copy (32 B per lane per iteration): the known-size reference
// for HBM counters and the achievable copy bandwidth of the device.
__global__ void k_stream_copy(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int mind_copy_bandwidth(int32_t device, int64_t bytes, int32_t repeats, double* gbps) {
  if (bytes <= 0 || bytes % 32 != 0 || repeats <= 0 || !gbps) return fail(MIND_EARG, "bytes must be a positive multiple of 32");
  HIP_TRY(hipSetDevice(device));
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return fail(MIND_EOOM, "device allocation failed");
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(MIND_EOOM, "device allocation failed");
  }
  hipEvent_t e0, e1;
  hipError_t err = hipMemset(a, 0, bytes);
  if (err == hipSuccess) err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  const int64_t n = bytes / 32;
  const unsigned grid = 256 * 8 * 4;  // 8 XCDs x 32 CUs x 4 workgroups
  if (err == hipSuccess) {
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, (const double4*)a, (double4*)b, n);
    err = hipEventRecord(e0, 0);
  }
  for (int r = 0; err == hipSuccess && r < repeats; ++r)
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, (const double4*)a, (double4*)b, n);
  if (err == hipSuccess) err = hipEventRecord(e1, 0);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) return fail(MIND_EHIP, std::string("copy bandwidth: ") + hipGetErrorString(err));
  *gbps = 2.0 * (double)bytes * repeats / (ms * 1e-3) / 1e9;
  return 0;
}

int mind_synchronize(mind_state* state) {
  State& S = *reinterpret_cast<State*>(state);
  HIP_TRY(hipSetDevice(S.device));
  HIP_TRY(hipStreamSynchronize(S.stream));
  return 0;
}

int mind_set_profiling(mind_state* state, int32_t on) {
  State& S = *reinterpret_cast<State*>(state);
  S.profiling = on != 0;
  S.kernel_ms = 0.0;
  S.kernel_launches = 0;
  S.kernel_steps = 0;
  S.events_used = 0;
  if (S.graph) {
    (void)hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
  }
  return 0;
}

int mind_kernel_time(const mind_state* state, int64_t* launches, int64_t* steps, double* total_ms) {
  const State& S = *reinterpret_cast<const State*>(state);
  *launches = S.kernel_launches;
  *steps = S.kernel_steps;
  *total_ms = S.kernel_ms;
  return 0;
}

int mind_window_steps(const mind_state* state, int32_t* steps) {
  *steps = reinterpret_cast<const State*>(state)->wmax;
  return 0;
}

}  // extern "C"

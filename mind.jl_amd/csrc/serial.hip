// Event-exact execution for models outside the window path: NetCon delays
// shorter than 1.5 dt (including zero) and connections into NetStims
// (feedback receivers).
//
// The window path needs every event produced inside a window to fall due after
// it, and NetStims that only schedule themselves. When a model breaks either
// rule, the state runs one step per window with the reference's event loop
// kept as it is on the device: one binary heap of pending events keyed by
// (time, order, target) in HBM, one insertion-order counter shared by every
// enqueue, an outbox of net_send / net_event / detector requests flushed in
// (rank, order) order, and the delivery loop that pops due events, routes
// source events at once, applies ordinary receivers and runs a feedback
// receiver alone followed by a flush (src/runtime/events.jl:4-165,
// src/runtime/utils/events.jl:7-62, ranks src/runtime/initialize.jl:629-649).
// Per step:
//   k_serial_step(start): detect(t) -> flush -> deliver(<= t + dt/2)
//   cell kernel, one step, with no detectors and no window events
//   k_serial_step(end):   deliver(<= t + dt), clock to the next step
// The queue work runs on one lane (it is a serial algorithm); threshold
// detection is spread over a wave and compacted by ballot in detector order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "device.h"
#include "state.h"
#include "serial_kernels.h"

namespace mind {

namespace {

template <typename T>
int serial_upload(State& S, T** out, const std::vector<T>& host) {
  return S.upload(out, host);
}

}  // namespace

// Host tables of the event-exact path: ranks (initialize.jl:629-649), sources of
// each point, the heap and the outbox.
int serial_setup(State& S, int64_t capacity) {
  const Layout& L = S.L;
  Dev& d = S.d;
  SerialDev& q = S.q;
  const int64_t npoints = (int64_t)L.point_group.size();
  std::vector<int64_t> target_rank(npoints, 0), source_rank(L.nsources, 0);
  int64_t rank = 0;
  for (size_t gi = 0; gi < L.groups.size(); ++gi) {
    const MechGroup& G = L.groups[gi];
    if (G.mech != MIND_MECH_NETSTIM) continue;
    uint64_t top = 0;
    for (int64_t r = 0; r < G.rows; ++r) {
      if (G.point[r] >= 0) target_rank[G.point[r]] = rank + (int64_t)G.logical[r] + 1;
      top = std::max<uint64_t>(top, G.logical[r]);
    }
    rank += (int64_t)top + 1;
  }
  std::vector<int64_t> src_off(npoints + 1, 0);
  for (int64_t s = 0; s < L.nsources; ++s) {
    const bool event_source = L.source_point[s] >= 0 && !L.source_external[s];
    source_rank[s] = event_source ? target_rank[L.source_point[s]] : rank + s + 1;
    if (event_source) src_off[L.source_point[s] + 1] += 1;
  }
  for (int64_t p = 0; p < npoints; ++p) src_off[p + 1] += src_off[p];
  std::vector<int32_t> src(src_off[npoints]);
  {
    std::vector<int64_t> at(src_off.begin(), src_off.end() - 1);
    for (int64_t s = 0; s < L.nsources; ++s)
      if (L.source_point[s] >= 0 && !L.source_external[s]) src[at[L.source_point[s]]++] = (int32_t)s;
  }
  std::vector<int32_t> ns_point;
  if (d.ns_group >= 0) ns_point.assign(L.groups[d.ns_group].point.begin(), L.groups[d.ns_group].point.end());
  int rc;
  int64_t* p64;
  int32_t* p32;
  if ((rc = serial_upload(S, &p64, target_rank))) return rc;
  q.target_rank = p64;
  if ((rc = serial_upload(S, &p64, source_rank))) return rc;
  q.source_rank = p64;
  if ((rc = serial_upload(S, &p64, src_off))) return rc;
  q.point_src_off = p64;
  if ((rc = serial_upload(S, &p32, src))) return rc;
  q.point_src = p32;
  if ((rc = serial_upload(S, &p32, ns_point))) return rc;
  q.ns_point = p32;
  if ((rc = S.alloc(&q.sequences, npoints))) return rc;
  q.heap_cap = capacity;
  q.outbox_cap = std::max<int64_t>(4096, (int64_t)L.det_node.size() + 2 * npoints + 1024);
  if ((rc = S.alloc(&q.heap, q.heap_cap))) return rc;
  if ((rc = S.alloc(&q.outbox, q.outbox_cap))) return rc;
  if ((rc = S.alloc(&q.heap_n, 1))) return rc;
  if ((rc = S.alloc(&q.outbox_n, 1))) return rc;
  HIP_TRY(hipMemset(q.heap_n, 0, sizeof(int64_t)));
  HIP_TRY(hipMemset(q.outbox_n, 0, sizeof(int64_t)));
  // cells detect nothing in their kernels: k_serial_step detects before the flush
  std::vector<int64_t> no_detectors(L.ncells + 1, 0);
  if ((rc = S.upload(&p64, no_detectors))) return rc;
  d.cell_det_off = p64;
  return 0;
}

int serial_initialize(State& S) {
  const Layout& L = S.L;
  Dev& d = S.d;
  HIP_TRY(hipMemsetAsync(S.q.sequences, 0, sizeof(int64_t) * std::max<size_t>(1, L.point_group.size()), S.stream));
  hipLaunchKernelGGL(k_serial_init, dim3(1), dim3(64), 0, S.stream, d, S.q, S.ns_logical, kNetStimDigest,
                     S.pattern_time, S.pattern_source, (int64_t)L.pattern_time.size());
  HIP_TRY(hipGetLastError());
  return 0;
}

int serial_launch(State& S, int32_t n) {
  for (int32_t s = 0; s < n; ++s) {
    HIP_TRY(launch(S, K_SERIAL_STEP, (const void*)k_serial_step, dim3(1), dim3(64), 0, S.stream, S.d, S.q,
                   (int32_t)0));
    int rc = cells_launch(S, 1);
    if (rc) return rc;
    HIP_TRY(launch(S, K_SERIAL_STEP, (const void*)k_serial_step, dim3(1), dim3(64), 0, S.stream, S.d, S.q,
                   (int32_t)1));
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int serial_inject(State& S, const SEvent& e) {
  hipLaunchKernelGGL(k_serial_inject, dim3(1), dim3(64), 0, S.stream, S.d, S.q, e);
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace mind

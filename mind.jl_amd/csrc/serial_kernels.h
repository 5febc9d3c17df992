// Event-exact path kernels: the device event loop (one binary heap, one
// insertion-order counter, the outbox flushed by rank) and the per-step queue
// kernels. Included by serial.hip, and compiled again with hipRTC, with the
// phases of runtime-compiled mechanisms added, for models that use them.
#pragma once

#include "device.h"

namespace mind {

namespace {

__device__ inline bool before(const SEvent& a, const SEvent& b) {
  return a.time < b.time || (a.time == b.time && (a.order < b.order || (a.order == b.order && a.target < b.target)));
}

struct Engine {  // lane 0's view of the queue
  const Dev& d;
  const SerialDev& q;

  __device__ void push(SEvent e) {
    int64_t i = *q.heap_n;
    if (i >= q.heap_cap) {
      raise_error(d.error, MIND_EQUEUE);
      return;
    }
    *q.heap_n = i + 1;
    while (i > 0) {
      const int64_t p = (i - 1) >> 1;
      const SEvent up = q.heap[p];
      if (!before(e, up)) break;
      q.heap[i] = up;
      i = p;
    }
    q.heap[i] = e;
  }

  __device__ SEvent pop() {
    const SEvent top = q.heap[0];
    const int64_t n = *q.heap_n - 1;
    *q.heap_n = n;
    if (n == 0) return top;
    const SEvent last = q.heap[n];
    int64_t i = 0;
    while (2 * i + 1 < n) {
      int64_t c = 2 * i + 1;
      SEvent ce = q.heap[c];
      if (c + 1 < n) {
        const SEvent cr = q.heap[c + 1];
        if (before(cr, ce)) {
          ++c;
          ce = cr;
        }
      }
      if (!before(ce, last)) break;
      q.heap[i] = ce;
      i = c;
    }
    q.heap[i] = last;
    return top;
  }

  // enqueue_event!: a fresh insertion order from the shared counter
  __device__ void enqueue(SEvent e) {
    e.order = (int64_t)(++d.clock->ticket);
    push(e);
  }

  __device__ void record(int32_t source, double t) {
    if (!d.record_spikes) return;
    const unsigned long long pos = atomicAdd(d.spike_count, 1ull);
    if ((int64_t)pos < d.spike_cap) {
      d.spike_t[pos] = t;
      d.spike_sid[pos] = d.source_sid[source];
    } else {
      raise_error(d.error, MIND_EQUEUE);
    }
  }

  // route_spike!: the source's connections in edge order
  __device__ void route(int32_t source, double t) {
    for (int64_t e = d.fanout[source]; e < d.fanout[source + 1]; ++e)
      enqueue(SEvent{t + d.edge_delay[e], 0.0, e + 1, 0, d.edge_target[e] + 1, 0});
  }

  __device__ void request(SEvent e) {
    const int64_t n = *q.outbox_n;
    if (n >= q.outbox_cap) {
      raise_error(d.error, MIND_EQUEUE);
      return;
    }
    q.outbox[n] = e;
    *q.outbox_n = n + 1;
  }

  __device__ int64_t rank(const SEvent& e) const {
    return e.target > 0 ? q.target_rank[e.target - 1] : e.target < 0 ? q.source_rank[-e.target - 1] : 0;
  }

  // flush_spikes!: requests in (rank, order) order (stable), self-events
  // enqueued, spikes routed
  __device__ void flush() {
    const int64_t n = *q.outbox_n;
    for (int64_t i = 1; i < n; ++i) {
      const SEvent x = q.outbox[i];
      const int64_t rx = rank(x);
      int64_t j = i - 1;
      while (j >= 0) {
        const SEvent y = q.outbox[j];
        const int64_t ry = rank(y);
        if (!(rx < ry || (rx == ry && x.order < y.order))) break;
        q.outbox[j + 1] = y;
        --j;
      }
      q.outbox[j + 1] = x;
    }
    for (int64_t i = 0; i < n; ++i) {
      const SEvent e = q.outbox[i];
      if (e.target > 0) enqueue(e);
      else if (e.target < 0) route(-e.target - 1, e.time);
    }
    *q.outbox_n = 0;
  }

  // net_event at time t from point p: one request per source of the point
  __device__ void net_event(int32_t p, double t, int64_t edge) {
    const int64_t seq = ++q.sequences[p];
    for (int64_t k = q.point_src_off[p]; k < q.point_src_off[p + 1]; ++k) {
      const int32_t source = q.point_src[k];
      record(source, t);
      request(SEvent{t, 0.0, edge, seq, -(source + 1), 0});
    }
  }

  __device__ void net_send(int32_t p, double t, int64_t edge, double delay, double flag) {
    if (!(delay >= 0.0) || !mind_isfinite(delay)) {
      raise_error(d.error, MIND_EEVENT);
      return;
    }
    const int64_t seq = ++q.sequences[p];
    request(SEvent{t + delay, flag, edge, seq, p + 1, 0});
  }

  // One NET_RECEIVE; returns whether the target is a feedback receiver.
  __device__ bool receive(const SEvent& ev) {
    const int32_t p = ev.target - 1;
    const GroupDev& G = d.g[d.point_group[p]];
    const int64_t row = d.point_row[p];
    const double w = ev.edge == 0 ? 0.0 : d.edge_weight[ev.edge - 1];
    d.clock->delivered += 1;
    if (G.mech != MIND_MECH_NETSTIM) {
      receive_event(d, G, row, ev.time, w);
      return false;
    }
    Stream s = d.ns_stream[row];
    const NetStimOut o = phase_netstim_receive(cols(G), row, s, d.seed, ev.time, w, ev.flag);
    d.ns_stream[row] = s;
    if (o.spiked) net_event(p, o.spike_time, ev.edge);
    if (o.send) net_send(p, ev.time, ev.edge, o.send_delay, o.send_flag);
    return true;
  }

  // deliver_events!(cutoff)
  __device__ void deliver(double cutoff) {
    while (*q.heap_n > 0 && q.heap[0].time <= cutoff) {
      const SEvent e = pop();
      if (e.target < 0) {
        const int32_t source = -e.target - 1;
        route(source, e.time);
        record(source, e.time);
      } else if (receive(e)) {
        flush();
      }
      if (*d.error >= MIND_EQUEUE) return;
    }
  }
};

}  // namespace

// initialize!'s queue part: PatternStim spikes first (reset_events!), then
// NetStim INITIAL in row order with its flush, then the time-zero delivery.
__global__ void __launch_bounds__(64) k_serial_init(Dev d, SerialDev q, const uint64_t* logical, uint64_t digest,
                                                    const double* pattern_time, const int32_t* pattern_source,
                                                    int64_t npattern) {
  if (threadIdx.x != 0) return;
  Engine E{d, q};
  *q.heap_n = 0;
  *q.outbox_n = 0;
  d.clock->ticket = 0;
  d.clock->delivered = 0;
  for (int64_t k = 0; k < npattern; ++k) E.enqueue(SEvent{pattern_time[k], 0.0, 0, 0, -(pattern_source[k] + 1), 0});
  if (d.ns_group >= 0) {
    const GroupDev& G = d.g[d.ns_group];
    for (int64_t r = 0; r < d.ns_rows; ++r) {
      Stream s;
      s.counter = 0;
      s.key[0] = (uint32_t)logical[r];
      s.key[1] = logical[r] >> 32;
      s.key[2] = (uint64_t)(uint32_t)digest;
      stream_setseq(s, d.seed, 0.0);
      const NetStimOut o = phase_netstim_initial(cols(G), r, s, d.seed);
      d.ns_stream[r] = s;
      if (o.send) E.net_send(q.ns_point[r], 0.0, 0, o.send_delay, o.send_flag);
    }
    E.flush();
  }
  E.flush();
  E.deliver(0.0);
}

// One half of a step: phase 0 detects at t, flushes and delivers through
// t + dt/2; phase 1 delivers through t + dt and advances the clock.
__global__ void __launch_bounds__(64) k_serial_step(Dev d, SerialDev q, int32_t phase) {
  const int lane = threadIdx.x;
  const int64_t k = d.clock->step;
  const double dt = d.dt, t = (double)k * dt;
  Engine E{d, q};
  if (phase == 0) {
    for (int64_t base = 0; base < d.ndet; base += 64) {
      const int64_t i = base + lane;
      bool cross = false;
      if (i < d.ndet) {
        const bool cur = d.v[d.det_node[i]] > d.det_thr[i];
        cross = cur && !d.det_above[i];
        d.det_above[i] = cur;
      }
      unsigned long long mask = __ballot(cross);
      if (lane == 0) {
        while (mask) {
          const int b = __ffsll((long long)mask) - 1;
          mask &= mask - 1;
          const int32_t source = d.det_source[base + b];
          E.record(source, t + 1.0e-10);
          E.request(SEvent{t + 1.0e-10, 0.0, 0, 0, -(source + 1), 0});
        }
      }
    }
    if (lane == 0) {
      E.flush();
      E.deliver(t + 0.5 * dt);
    }
  } else if (lane == 0) {
    record_artificial(d, k);  // record, then deliver through the end of the step
    E.deliver(t + dt);
    d.clock->step = k + 1;
  }
}

// inject_event! / inject_spike!: enqueued now with the next insertion order.
__global__ void __launch_bounds__(64) k_serial_inject(Dev d, SerialDev q, SEvent e) {
  if (threadIdx.x != 0) return;
  Engine E{d, q};
  E.enqueue(e);
}

}  // namespace mind

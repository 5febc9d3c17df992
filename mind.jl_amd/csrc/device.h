// Device-side data structures shared by the window kernels (see runtime.hip).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include "layout.h"
#endif
#include "mechanisms.h"

namespace mind {

// Event in the delivery ring and in per-cell window lists.
struct QEvent {
  double t;
  uint64_t ticket;
  int64_t call;  // delivery call: 2k = midpoint of step k, 2k+1 = end of step k
  int32_t edge;
  int32_t pad;
};

// An injected source spike (inject_spike!): time, heap insertion order, source.
struct InjSpike {
  double t;
  uint64_t seq;
  int32_t source;
  int32_t pad;
};
constexpr int kInjCap = 1024;  // injected spikes handed to one window, events per staging copy

// A NetStim self-event or injected source spike waiting in the window's heap
// (row < 0: injected spike -row - 1), ordered by (time, insertion order).
struct HeapItem {
  double t;
  uint64_t seq;
  int32_t row;
};

struct Clock {
  int64_t step;
  uint64_t ticket;   // routed events so far
  uint64_t ns_seq;   // NetStim self-event insertion order
  int64_t emitted;   // spike emissions of the last exchange (in emit_all)
  int64_t mstep;     // macro step
  int64_t delivered; // connection events gathered for delivery since initialize
};

struct GroupDev {
  int32_t mech;
  int32_t point;  // 1 for point processes
  int64_t rows;
  double* data;
  const int32_t* node;
  const int64_t* cell_rows;
  const uint8_t* run_start;
};

struct Emit {  // one spike emission: source index and spike time
  int64_t key;
  double t;
  int32_t source;
  int32_t pad;
};

// A pending event of the event-exact path (ScheduledEvent, src/runtime/state.jl:18-24):
// target > 0 is point target-1, target < 0 is source -target-1; edge is the
// 1-based connection edge (0 for none); order is its insertion order.
struct SEvent {
  double time, flag;
  int64_t edge, order;
  int32_t target, pad;
};

// Queue state of the event-exact path (serial.hip).
struct SerialDev {
  SEvent* heap;
  int64_t heap_cap;
  int64_t* heap_n;
  SEvent* outbox;
  int64_t outbox_cap;
  int64_t* outbox_n;
  int64_t* sequences;           // per point: net_send / net_event count
  const int64_t* target_rank;   // per point
  const int64_t* source_rank;   // per source
  const int64_t* point_src_off; // sources of each point (CSR)
  const int32_t* point_src;
  const int32_t* ns_point;      // point of each NetStim row
};

struct Dev {
  double dt, celsius;
  uint64_t seed;
  int32_t ncells;
  int64_t nodes;
  const int64_t* cell_off;
  const int32_t* parent;
  const double* area;
  const double* cdt;  // capacitance / assembly dt (dt, or dt/2 with secondorder)
  double vscale;      // v += vscale * dv: 1, or 2 with secondorder (update_voltage_kernel!)
  const double* axial;
  const double* adiag;
  double* v;
  double* rhs;
  double* diag;
  const int64_t* child_off;
  const int32_t* child;
  int32_t ngroups;
  const GroupDev* g;  // device array of ngroups descriptors
  int32_t ns_group;  // group index of NetStim or -1
  const int32_t* point_group;
  const int32_t* point_row;
  const int32_t* point_cell;
  int64_t ndet;
  const int32_t* det_node;
  const double* det_thr;
  uint8_t* det_above;
  const int64_t* cell_det_off;
  const int32_t* cell_det;
  const int32_t* det_source;
  uint8_t* det_flags;  // [window step][detector]
  const int64_t* fanout;
  const int32_t* edge_target;
  const double* edge_delay;
  const double* edge_weight;
  const int64_t* source_sid;
  // NetStim rows
  int64_t ns_rows;
  Stream* ns_stream;
  double* ns_t;
  double* ns_flag;
  uint64_t* ns_seq;
  int32_t* ns_edge;
  int32_t* ns_valid;
  const int32_t* ns_source;
  HeapItem* ns_heap;  // ns_rows + kInjCap entries: the window's heap when LDS cannot hold it
  // delivery ring
  int64_t ring_slots, ring_cap;
  QEvent* ring;
  unsigned long long* ring_count;
  // window lists, one segment per event bucket (a cell, or a batched cell group)
  int64_t win_cap;
  QEvent* win;
  int64_t nbuckets;
  const int32_t* cell_bucket;
  int64_t* seg_off;                 // nbuckets + 1
  unsigned long long* seg_cursor;   // nbuckets
  // emissions
  int64_t emit_cap;
  Emit* emit_det;
  Emit* emit_ns;
  Emit* emit_all;
  int64_t* emit_base;
  // spike recording
  int64_t spike_cap;
  double* spike_t;
  int64_t* spike_sid;
  unsigned long long* spike_count;
  int32_t record_spikes;
  // value recording
  int64_t nrec;
  const int32_t* rec_kind;   // 0 voltage, 1 field
  const int32_t* rec_group;
  const int32_t* rec_field;
  const int64_t* rec_index;  // node or row
  const int64_t* cell_rec_off;
  const int32_t* cell_rec;
  double* rec_buf;
  int64_t rec_stride;
  int64_t rec_start;
  // records of artificial-cell fields (NetStim): sampled by the event kernels,
  // after the midpoint delivery of each step, before its end-of-step delivery
  const int32_t* art_rec;
  int64_t n_art_rec;
  // soma-voltage snapshots of the MICRO2MACRO outputs at every macro-step end of
  // a window (micro-macro windows longer than one macro step): snap[m * n_snap +
  // det_snap[detector]] after window step (m + 1) * snap_every - 1
  double* snap;
  const int32_t* det_snap;
  int32_t snap_every;
  int64_t n_snap;
  // injected source spikes due in the next window, and their count (the window's
  // exchange consumes them and clears the count)
  InjSpike* inj;
  int32_t* inj_count;
  Clock* clock;
  int32_t* error;
};

__device__ inline void raise_error(int32_t* error, int32_t code) {
  // Keep the first, most specific failure (atomic max like the reference's error word).
  atomicMax(error, code);
}

__host__ __device__ inline double call_cutoff(int64_t j, double dt) {
  const int64_t k = j >= 0 ? j / 2 : (j - 1) / 2;  // floor(j / 2)
  const int64_t phase = j - 2 * k;
  return phase == 0 ? (double)k * dt + 0.5 * dt : (double)(k + 1) * dt;
}

// First delivery call j >= jmin whose cutoff reaches t.
__host__ __device__ inline int64_t call_of(double t, int64_t jmin, double dt) {
  double est = floor(t / dt);
  int64_t j = jmin;
  if (est * 2.0 - 2.0 > (double)jmin) j = (int64_t)est * 2 - 2;
  while (t > call_cutoff(j, dt)) ++j;
  return j;
}

__device__ inline int64_t emit_key(int64_t j) { return j < 0 ? -1 : 3 * (j / 2) + 1 + (j & 1); }

__device__ inline Cols cols(const GroupDev& g) { return Cols{g.data, g.rows}; }

// Samples the artificial-cell records for step k (sample k + 1 of the run).
__device__ inline void record_artificial(const Dev& d, int64_t k) {
  const int64_t sample = k + 1 - d.rec_start;
  for (int64_t q = 0; q < d.n_art_rec; ++q) {
    const int32_t rec = d.art_rec[q];
    const GroupDev& G = d.g[d.rec_group[rec]];
    d.rec_buf[rec * d.rec_stride + sample] = G.data[d.rec_field[rec] * G.rows + d.rec_index[rec]];
  }
}

}  // namespace mind

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
  // ion transfers: ion_row[r * nslot + s] (-1 none); ion_bind: nread loads, nwrite
  // stores, nrecv NET_RECEIVE exchanges
  const int32_t* ion_row;
  const IonBind* ion_bind;
  int16_t nslot, nread, nwrite, nrecv;
  // 1 if nrn_state loads the ion values and stores its writes: some mechanism
  // writes a reversal or a concentration, or this one reads an ion current or
  // di/dv. Otherwise the columns still hold what nrn_cur loaded in the same
  // step, and there is nothing but currents to store.
  int16_t ion_state;
  // 1 for a density mechanism with at most two USEION statements, two reads and
  // two written currents and no other writes (the batched kernels' hh path)
  int16_t ion_fast;
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
  const int32_t* rec_kind;   // 0 voltage, 1 and 2 field, 3 ion column
  const int32_t* rec_group;
  const int32_t* rec_field;
  const int64_t* rec_index;  // node, row or ion row
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
  // ion rows (node, species), sorted by node: columns MIND_ION_* of row q at
  // ion[column * nion + q]
  int64_t nion;
  double* ion;
  const uint8_t* ion_style;
  int32_t ion_styles;            // the OR of every row's style
  const int32_t* ion_species;
  const double* ion_valence;     // per species
  const double* ion_conc0;       // per species: internal, external (cinit)
  const int64_t* node_ion_off;   // nodes + 1
  int32_t ion_ordered;           // a mechanism writes a reversal or concentration
  int32_t ion_correct;           // secondorder 2: currents += di/dv * dv
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

// ------------------------------------------------------------------ ions
// prepare_currents_kernel! (src/runtime/cuda/micro.jl:601-629) for ion row q:
// clears the current and di/dv; the Nernst reversal where style & mask (0x40
// every step, 0x20 at initialization).
__device__ inline void ion_prepare(const Dev& d, int64_t q, uint8_t mask) {
  d.ion[MIND_ION_I * d.nion + q] = 0.0;
  d.ion[MIND_ION_DIDV * d.nion + q] = 0.0;
  if ((d.ion_styles & mask) && (d.ion_style[q] & mask))
    d.ion[MIND_ION_E * d.nion + q] = mind_nernst(d.ion[MIND_ION_CI * d.nion + q], d.ion[MIND_ION_CO * d.nion + q],
                                                 d.ion_valence[d.ion_species[q]], d.celsius);
}

// load_ions! / store_ions! (cuda/micro.jl:376-392): the mechanism fields read from
// the row's ion rows; the fields written back, the current column excepted.
__device__ inline void ion_load(const Dev& d, const GroupDev& G, const Cols& c, int64_t r) {
  for (int k = 0; k < G.nread; ++k) {
    const IonBind b = G.ion_bind[k];
    const int32_t q = G.ion_row[r * G.nslot + b.slot];
    if (q >= 0) c(b.field, r) = d.ion[b.column * d.nion + q];
  }
}
__device__ inline void ion_store(const Dev& d, const GroupDev& G, const Cols& c, int64_t r) {
  for (int k = G.nread; k < G.nread + G.nwrite; ++k) {
    const IonBind b = G.ion_bind[k];
    const int32_t q = G.ion_row[r * G.nslot + b.slot];
    if (q >= 0 && b.column != MIND_ION_I) d.ion[b.column * d.nion + q] = c(b.field, r);
  }
}

// nrn_cur of a row with ion transfers (mechanism_current_instance!, cuda/micro.jl:
// 427-500): loads, both probes (fields stored), then per written ion current
// i += s * i_row and di/dv += s * (i(v + 0.001) - i_row) / 0.001 with s = 1 for
// density rows and 100 / area for point processes; other writes stored.
__device__ inline void ion_current(const Dev& d, const GroupDev& G, const Cols& c, int64_t r, double v, double t,
                                   double area, double& value, double& shifted) {
  ion_load(d, G, c, r);
  double ish[2 * kMaxIonCurrents];
  phase_current(G.mech, c, r, v, t, d.dt, value, shifted, true, d.celsius, ish);
  for (int k = G.nread; k < G.nread + G.nwrite; ++k) {
    const IonBind b = G.ion_bind[k];
    const int32_t q = G.ion_row[r * G.nslot + b.slot];
    if (q < 0) continue;
    if (b.column == MIND_ION_I) {
      const double s = G.point ? 100.0 / area : 1.0;
      const double x = ish[kMaxIonCurrents + b.aux];  // the stored field's value
      d.ion[MIND_ION_I * d.nion + q] += s * x;
      d.ion[MIND_ION_DIDV * d.nion + q] += s * (ish[b.aux] - x) / 0.001;
    } else {
      d.ion[b.column * d.nion + q] = c(b.field, r);
    }
  }
}

// nrn_state of a row: loads, SOLVE statements, stores (the loads and stores only
// where they can change something, GroupDev::ion_state).
__device__ inline void advance_row(const Dev& d, const GroupDev& G, int64_t r, double v, double t) {
  const Cols c = cols(G);
  if (G.ion_state) ion_load(d, G, c, r);
  phase_advance(G.mech, c, r, v, t, d.dt, d.celsius);
  if (G.ion_state) ion_store(d, G, c, r);
}

// NET_RECEIVE of a connection event: the concentrations NET_RECEIVE mentions are
// imported first and the STATE ones published after (receive_ions!,
// cuda/micro.jl:324-337).
__device__ inline void receive_event(const Dev& d, const GroupDev& G, int64_t r, double t, double weight) {
  const Cols c = cols(G);
  const int k0 = G.nread + G.nwrite, k1 = k0 + G.nrecv;
  for (int k = k0; k < k1; ++k) {
    const IonBind b = G.ion_bind[k];
    const int32_t q = G.ion_row[r * G.nslot + b.slot];
    if (q >= 0) c(b.field, r) = d.ion[b.column * d.nion + q];
  }
  phase_receive(G.mech, c, r, t, weight);
  for (int k = k0; k < k1; ++k) {
    const IonBind b = G.ion_bind[k];
    const int32_t q = G.ion_row[r * G.nslot + b.slot];
    if (q >= 0 && b.aux) d.ion[b.column * d.nion + q] = c(b.field, r);
  }
}

// update_voltage_kernel!'s secondorder-2 correction (cuda/micro.jl:660-679): the
// ion currents of node n move by di/dv times the solved increment dv.
__device__ inline void ion_correct_node(const Dev& d, int64_t n, double dv) {
  for (int64_t q = d.node_ion_off[n]; q < d.node_ion_off[n + 1]; ++q)
    d.ion[MIND_ION_I * d.nion + q] += d.ion[MIND_ION_DIDV * d.nion + q] * dv;
}

// A recorded value other than a voltage: a mechanism field (kinds 1, 2) or an
// ion column (3).
__device__ inline double stored_value(const Dev& d, int32_t rec) {
  const int64_t idx = d.rec_index[rec];
  if (d.rec_kind[rec] == 3) return d.ion[d.rec_field[rec] * d.nion + idx];
  const GroupDev& G = d.g[d.rec_group[rec]];
  return G.data[d.rec_field[rec] * G.rows + idx];
}

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

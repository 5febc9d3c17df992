// Cell kernels that run mechanism phases: INITIAL and the time-zero current
// evaluation of a group, and the plain window kernel (one cell per thread).
// Included by runtime.hip for the builtin mechanisms, and compiled again with
// hipRTC, with the phases of runtime-compiled mechanisms added, for models that
// use them.
#pragma once

#include "device.h"

namespace mind {

// INITIAL of a group's rows (initialize_mechanism_kernel!, cuda/micro.jl:352-374),
// with ion loads before and stores after. Point rows sharing a node with ion
// transfers run in row order on one thread (the run's first row).
__global__ void k_init_group(Dev d, int gi) {
  const GroupDev& G = d.g[gi];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= G.rows) return;
  if (G.mech == MIND_MECH_NETSTIM) {
    // initialize_random_kernel! (cuda/micro.jl:79-91): keys from the logical id
    // and the NetStim source digest, then the INITIAL block.
    return;
  }
  const Cols c = cols(G);
  if (!G.nslot) {
    phase_initial(G.mech, c, r, d.v[G.node[r]], d.celsius);
    return;
  }
  if (G.point && !G.run_start[r]) return;
  int64_t q = r;
  do {
    ion_load(d, G, c, q);
    phase_initial(G.mech, c, q, d.v[G.node[q]], d.celsius);
    ion_store(d, G, c, q);
    ++q;
  } while (G.point && q < G.rows && G.node[q] == G.node[r]);
}

// initialize_concentration_reversals_kernel! (cuda/micro.jl:764-784): after the
// INITIAL of a group that writes concentrations, the Nernst reversal of its
// rows' ion rows whose style asks for it at initialization (0x20).
__global__ void k_init_reversals(Dev d, int gi) {
  const GroupDev& G = d.g[gi];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= G.rows) return;
  for (int k = G.nread; k < G.nread + G.nwrite; ++k) {
    const IonBind b = G.ion_bind[k];
    if (b.column != MIND_ION_CI && b.column != MIND_ION_CO) continue;
    const int32_t q = G.ion_row[r * G.nslot + b.slot];
    if (q >= 0 && (d.ion_style[q] & 0x20))
      d.ion[MIND_ION_E * d.nion + q] = mind_nernst(d.ion[MIND_ION_CI * d.nion + q], d.ion[MIND_ION_CO * d.nion + q],
                                                   d.ion_valence[d.ion_species[q]], d.celsius);
  }
}

// initialize_ions_kernel! and prepare_currents_kernel!(Initial) (cuda/micro.jl:
// 753-762, 601-629): concentrations reset to the species' where style & 0x04,
// currents cleared, Nernst reversals where style & 0x20.
__global__ void k_init_ions(Dev d) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.nion) return;
  if (d.ion_style[q] & 0x04) {
    const int32_t sp = d.ion_species[q];
    d.ion[MIND_ION_CI * d.nion + q] = d.ion_conc0[2 * sp];
    d.ion[MIND_ION_CO * d.nion + q] = d.ion_conc0[2 * sp + 1];
  }
  ion_prepare(d, q, 0x20);
}

// micro_currents!(0.0) at the end of initialize!: the assigned current fields and
// the ion currents (cleared first, then summed group by group, rows sharing a
// node in row order).
__global__ void k_currents_fields(Dev d, int gi, double t) {
  const GroupDev& G = d.g[gi];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= G.rows) return;
  double value, shifted;
  if (!G.nslot) {
    phase_current(G.mech, cols(G), r, d.v[G.node[r]], t, d.dt, value, shifted, true, d.celsius, nullptr);
    return;
  }
  if (G.point && !G.run_start[r]) return;
  int64_t q = r;
  do {
    const int32_t node = G.node[q];
    ion_current(d, G, cols(G), q, d.v[node], t, d.area[node], value, shifted);
    ++q;
  } while (G.point && q < G.rows && G.node[q] == G.node[r]);
}

// The step-start part of prepare_currents_kernel! for every ion row (k_currents_fields'
// initialization sweep).
__global__ void k_prepare_ions(Dev d) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < d.nion) ion_prepare(d, q, 0x40);
}

// ------------------------------------------------------------ shared pieces
__device__ inline void record_cell(const Dev& d, int64_t c, int64_t sample) {
  for (int64_t q = d.cell_rec_off[c]; q < d.cell_rec_off[c + 1]; ++q) {
    const int32_t rec = d.cell_rec[q];
    const double x = d.rec_kind[rec] == 0 ? d.v[d.rec_index[rec]] : stored_value(d, rec);
    d.rec_buf[rec * d.rec_stride + sample] = x;
  }
}

// Applies this cell's events of delivery call j; returns the new cursor.
__device__ inline int64_t deliver_cell(const Dev& d, int64_t seg, int64_t end, int64_t j) {
  while (seg < end && d.win[seg].call == j) {
    const int32_t e = d.win[seg].edge;
    const int32_t target = d.edge_target[e];
    receive_event(d, d.g[d.point_group[target]], d.point_row[target], d.win[seg].t, d.edge_weight[e]);
    ++seg;
  }
  return seg;
}

__device__ inline void detect_cell(const Dev& d, int64_t c, int s) {
  for (int64_t q = d.cell_det_off[c]; q < d.cell_det_off[c + 1]; ++q) {
    const int32_t k = d.cell_det[q];
    const bool cur = d.v[d.det_node[k]] > d.det_thr[k];
    d.det_flags[(int64_t)s * d.ndet + k] = (cur && !d.det_above[k]) ? 1 : 0;
    d.det_above[k] = cur;
  }
}

// ------------------------------------------------------------ plain kernel
// One cell per thread; every per-node array lives in HBM. This is the
// reference's one-instance-per-thread execution shape, kept as the baseline
// the batched kernel is measured against.
__global__ void __launch_bounds__(64) k_plain(Dev d, int32_t n) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.ncells) return;
  const int64_t k0 = d.clock->step;
  const int64_t n0 = d.cell_off[c], n1 = d.cell_off[c + 1];
  int64_t seg = d.seg_off[d.cell_bucket[c]];
  const int64_t seg_end = d.seg_off[d.cell_bucket[c] + 1];
  const double dt = d.dt;
  double* v = d.v;
  double* rhs = d.rhs;
  double* diag = d.diag;
  for (int s = 0; s < n; ++s) {
    const int64_t k = k0 + s;
    const double t = (double)k * dt;
    detect_cell(d, c, s);
    seg = deliver_cell(d, seg, seg_end, 2 * k);
    // nrn_cur at t + dt/2
    const double tc = t + 0.5 * dt;
    for (int64_t i = n0; i < n1; ++i) { rhs[i] = 0.0; diag[i] = 0.0; }
    if (d.nion)
      for (int64_t q = d.node_ion_off[n0]; q < d.node_ion_off[n1]; ++q) ion_prepare(d, q, 0x40);
    for (int gi = 0; gi < d.ngroups; ++gi) {
      const GroupDev& G = d.g[gi];
      if (G.mech == MIND_MECH_NETSTIM) continue;
      const Cols col = cols(G);
      for (int64_t r = G.cell_rows[c]; r < G.cell_rows[c + 1]; ++r) {
        const int32_t node = G.node[r];
        double value, shifted;
        if (G.nslot)
          ion_current(d, G, col, r, v[node], tc, d.area[node], value, shifted);
        else
          phase_current(G.mech, col, r, v[node], tc, dt, value, shifted, true, d.celsius, nullptr);
        const double scale = G.point ? 1.0 : d.area[node] * 0.01;
        rhs[node] -= scale * value;
        diag[node] += scale * (shifted - value) / 0.001;
      }
    }
    // cable assembly: own axial current first, children in increasing index
    for (int64_t i = n0; i < n1; ++i) {
      diag[i] = d.cdt[i] + diag[i] + d.adiag[i];
      const int32_t p = d.parent[i];
      if (p >= 0) {
        const double cur = d.axial[i] * (v[p] - v[i]);
        rhs[i] += cur;
        rhs[p] -= cur;
      }
    }
    // Hines elimination (children before parents) and back substitution
    bool bad = false;
    for (int64_t i = n1 - 1; i >= n0; --i) {
      const double piv = diag[i];
      bad |= (piv == 0.0) || !mind_isfinite(piv);
      const int32_t p = d.parent[i];
      if (p >= 0) {
        const double g = d.axial[i];
        const double f = g / piv;
        diag[p] -= f * g;
        rhs[p] += f * rhs[i];
      }
    }
    if (bad) raise_error(d.error, MIND_ESINGULAR);
    for (int64_t i = n0; i < n1; ++i) {
      double x = rhs[i];
      const int32_t p = d.parent[i];
      if (p >= 0) x += d.axial[i] * rhs[p];
      x = x / diag[i];
      rhs[i] = x;
      if (d.ion_correct) ion_correct_node(d, i, x);
      v[i] += d.vscale * x;
    }
    // nrn_state at t + dt
    for (int gi = 0; gi < d.ngroups; ++gi) {
      const GroupDev& G = d.g[gi];
      if (!phase_has_state(G.mech)) continue;
      for (int64_t r = G.cell_rows[c]; r < G.cell_rows[c + 1]; ++r) advance_row(d, G, r, v[G.node[r]], t + dt);
    }
    if (d.nrec) record_cell(d, c, k + 1 - d.rec_start);
    if (d.snap && (s + 1) % d.snap_every == 0) {
      const int64_t m = (s + 1) / d.snap_every - 1;
      for (int64_t q = d.cell_det_off[c]; q < d.cell_det_off[c + 1]; ++q) {
        const int32_t kk = d.cell_det[q];
        if (d.det_snap[kk] >= 0) d.snap[m * d.n_snap + d.det_snap[kk]] = v[d.det_node[kk]];
      }
    }
    seg = deliver_cell(d, seg, seg_end, 2 * k + 1);
  }
}

}  // namespace mind

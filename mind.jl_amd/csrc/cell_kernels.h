// Cell kernels that run mechanism phases: INITIAL and the time-zero current
// evaluation of a group, and the plain window kernel (one cell per thread).
// Included by runtime.hip for the builtin mechanisms, and compiled again with
// hipRTC, with the phases of runtime-compiled mechanisms added, for models that
// use them.
#pragma once

#include "device.h"

namespace mind {

__global__ void k_init_group(Dev d, int gi) {
  const GroupDev& G = d.g[gi];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= G.rows) return;
  if (G.mech == MIND_MECH_NETSTIM) {
    // initialize_random_kernel! (cuda/micro.jl:79-91): keys from the logical id
    // and the NetStim source digest, then the INITIAL block.
    return;
  }
  phase_initial(G.mech, cols(G), r, d.v[G.node[r]], d.celsius);
}

// micro_currents!(0.0) at the end of initialize!: sets the assigned current fields.
__global__ void k_currents_fields(Dev d, int gi, double t) {
  const GroupDev& G = d.g[gi];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= G.rows) return;
  double value, shifted;
  phase_current(G.mech, cols(G), r, d.v[G.node[r]], t, d.dt, value, shifted, true);
}

// ------------------------------------------------------------ shared pieces
__device__ inline void record_cell(const Dev& d, int64_t c, int64_t sample) {
  for (int64_t q = d.cell_rec_off[c]; q < d.cell_rec_off[c + 1]; ++q) {
    const int32_t rec = d.cell_rec[q];
    const int64_t idx = d.rec_index[rec];
    const double x = d.rec_kind[rec] == 0
                         ? d.v[idx]
                         : d.g[d.rec_group[rec]].data[d.rec_field[rec] * d.g[d.rec_group[rec]].rows + idx];
    d.rec_buf[rec * d.rec_stride + sample] = x;
  }
}

// Applies this cell's events of delivery call j; returns the new cursor.
__device__ inline int64_t deliver_cell(const Dev& d, int64_t seg, int64_t end, int64_t j) {
  while (seg < end && d.win[seg].call == j) {
    const int32_t e = d.win[seg].edge;
    const int32_t target = d.edge_target[e];
    const GroupDev& G = d.g[d.point_group[target]];
    phase_receive(G.mech, cols(G), d.point_row[target], d.win[seg].t, d.edge_weight[e]);
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
    for (int gi = 0; gi < d.ngroups; ++gi) {
      const GroupDev& G = d.g[gi];
      if (G.mech == MIND_MECH_NETSTIM) continue;
      const Cols col = cols(G);
      for (int64_t r = G.cell_rows[c]; r < G.cell_rows[c + 1]; ++r) {
        const int32_t node = G.node[r];
        double value, shifted;
        phase_current(G.mech, col, r, v[node], tc, dt, value, shifted, true);
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
      v[i] += d.vscale * x;
    }
    // nrn_state at t + dt
    for (int gi = 0; gi < d.ngroups; ++gi) {
      const GroupDev& G = d.g[gi];
      if (!phase_has_state(G.mech)) continue;
      const Cols col = cols(G);
      for (int64_t r = G.cell_rows[c]; r < G.cell_rows[c + 1]; ++r)
        phase_advance(G.mech, col, r, v[G.node[r]], t + dt, dt, d.celsius);
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

// Batched window kernel: a workgroup advances a group of consecutive cells
// through a whole window (W steps of the micro step) with its node state staged
// in LDS as far as it fits.
//
// The Hines solve of the reference runs one warp per group of cells with
// 32-lane "DHS" rounds and FP64 atomics into parents (src/runtime/cuda/micro.jl:688-749,
// plan src/model/micro/model.jl:906-1077). Here a workgroup solves its cells
// level by level: elimination visits nodes by height (a node's children are all
// lower), each node *pulling* its children's contributions in decreasing child
// index, which reproduces the CPU restatement's elimination order exactly and
// needs no atomics; back substitution visits nodes by depth. One node per lane,
// one barrier per level. The int16 parent / child links and the two node orders
// are built on the host per group. Two layouts of a workgroup's node state:
// all in LDS (mode 0: v, rhs, diag, axial conductances, links and schedule stay
// there for the whole window, so the solve reads no global memory), or lean
// (mode 4: only rhs and diag in LDS, 16 bytes per node, and the solve's per-node
// records streamed from HBM two levels ahead, so that more nodes fit a CU). The
// autotuner chooses the layout, the node budget per workgroup and the block
// size. Cells too large for either keep everything in HBM (mode 3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "device.h"
#include "state.h"

namespace mind {

// Lean layout (mode 4): the solve's per-level node records, read from HBM one
// level ahead. Elimination, in height order: the node, its higher and lower
// child (-1 if none, higher -2 if more than two) and their axial conductances.
// Back substitution, in depth order: the node, its parent and its own axial
// conductance. Indexed by group node base + position.
struct HRec {
  int16_t li, h, l, pad;
  double gh, gl;
};
struct DRec {
  int16_t li, p, pad0, pad1;
  double g;
};

struct BatchDev {
  const int32_t* grp_cell0;
  const int32_t* grp_cell1;
  const HRec* r_hrec;  // mode 4 only
  const DRec* r_drec;
  // per group one int16 block (local node ids): parent, lower and higher child,
  // nodes by height, nodes by depth, and the level offsets of both orders
  const int16_t* r_tab;
  const int32_t* grp_rtab;
  const int32_t* grp_nlh;  // levels by height
  const int32_t* grp_nld;  // levels by depth
};

// Launch lists: 0 groups within the node budget (main stream); 1 single cells
// above the budget, in the plan's layout; 2 cells too large for that, lean; 3
// cells too large for the lean layout, all in HBM (lists 1-3 on the second
// stream).
constexpr int kLists = 4;

struct BatchPlan {
  int64_t budget = 0;  // node budget per workgroup
  int64_t ngroups = 0;
  int threads = 256;
  int layout = 0;      // of the groups within the budget: 0 all in LDS, 4 lean
  std::vector<int32_t> lists[kLists];
  int32_t* dev_lists[kLists] = {};
  size_t lds[kLists] = {};  // dynamic LDS bytes of each list
  // lists 1-3 as one launch: groups and their modes
  std::vector<int32_t> large;
  std::vector<int8_t> large_mode;
  int32_t* dev_large = nullptr;
  int8_t* dev_large_mode = nullptr;
  size_t large_lds = 0;
  BatchDev b{};
  std::vector<void*> allocations;
};

constexpr int kMaxThreads = 512;
constexpr int kMaxGroups = 8;    // mechanism groups cached in LDS (the six of the step)
constexpr int kMaxDetLds = 256;  // detectors of a workgroup cached in LDS
// dynamic LDS of a workgroup, leaving room for its static tables
constexpr size_t kLdsResident = 152 * 1024;

// int16 entries of a group's table block.
__host__ __device__ inline int64_t rtab_len(int64_t nn, int64_t nlh, int64_t nld) { return 5 * nn + nlh + 1 + nld + 1; }
// Dynamic LDS bytes of a workgroup in mode M: mode 0 keeps 4 doubles per node and
// the table block, mode 4 rhs, diag and the level offsets, mode 3 nothing.
inline size_t window_bytes(int m, int64_t nn, int64_t nlh, int64_t nld) {
  if (m == 3) return 0;
  if (m == 4) return (size_t)(16 * nn + (2 * (nlh + 1 + nld + 1) + 7) / 8 * 8);
  return (size_t)(32 * nn + (2 * rtab_len(nn, nlh, nld) + 7) / 8 * 8);
}

// Delivery of the window-list events due at delivery call `call` (thread 0).
__device__ inline int64_t deliver_group(const Dev& d, int64_t seg, int64_t seg_end, int64_t call) {
  while (seg < seg_end && d.win[seg].call == call) {
    const int32_t e = d.win[seg].edge;
    const int32_t target = d.edge_target[e];
    const GroupDev& G = d.g[d.point_group[target]];
    phase_receive(G.mech, cols(G), d.point_row[target], d.edge_weight[e]);
    ++seg;
  }
  return seg;
}

// M says where a workgroup's node state lives: 0 all in LDS (v, rhs, diag, axial
// conductances, links and schedule), 4 rhs, diag and the level offsets in LDS, 3
// everything in HBM (rhs and diag in the global scratch columns, the links and
// schedule read from the plan).
__device__ inline HRec hrec_at(const HRec* hr, const int16_t* lh, int nlh, int L, int tid) {
  if (L < nlh) {
    const int i = lh[L] + tid;
    if (i < lh[L + 1]) return hr[i];
  }
  return HRec{-1, -1, -1, 0, 0.0, 0.0};
}

__device__ inline DRec drec_at(const DRec* dr, const int16_t* ld, int nld, int L, int tid) {
  if (L < nld) {
    const int i = ld[L] + tid;
    if (i < ld[L + 1]) return dr[i];
  }
  return DRec{-1, -1, 0, 0, 0.0};
}

// Tree solve of a lean workgroup: rhs and diag in LDS, each lane's node record
// of a level loaded from HBM two levels ahead. Same operations in the same
// order as the LDS-resident solve.
__device__ void solve_lean(const Dev& d, const BatchDev& b, int64_t n0, int nlh, int nld, const int16_t* lh,
                           const int16_t* ld, double* rhs, double* diag, double* v) {
  const int tid = threadIdx.x, T = blockDim.x;
  const HRec* hr = b.r_hrec + n0;
  bool bad = false;
  HRec cur = hrec_at(hr, lh, nlh, 0, tid), nxt = hrec_at(hr, lh, nlh, 1, tid);
  for (int L = 0; L < nlh; ++L) {
    const HRec ahead = hrec_at(hr, lh, nlh, L + 2, tid);
    const int a = lh[L], z = lh[L + 1];
    for (int i = a + tid; i < z; i += T) {
      const HRec r = i == a + tid ? cur : hr[i];
      const int li = r.li;
      double dd = diag[li], bb = rhs[li];
      if (r.h == -2) {
        const int64_t gi = n0 + li;
        for (int64_t cq = d.child_off[gi + 1] - 1; cq >= d.child_off[gi]; --cq) {
          const int lc = d.child[cq] - (int)n0;
          const double g = d.axial[n0 + lc];
          const double f = g / diag[lc];
          dd -= f * g;
          bb += f * rhs[lc];
        }
      } else {
        if (r.h >= 0) {
          const double f = r.gh / diag[r.h];
          dd -= f * r.gh;
          bb += f * rhs[r.h];
        }
        if (r.l >= 0) {
          const double f = r.gl / diag[r.l];
          dd -= f * r.gl;
          bb += f * rhs[r.l];
        }
      }
      diag[li] = dd;
      rhs[li] = bb;
      bad |= (dd == 0.0) || !mind_isfinite(dd);
    }
    __syncthreads();
    cur = nxt;
    nxt = ahead;
  }
  if (bad) raise_error(d.error, MIND_ESINGULAR);
  const DRec* dr = b.r_drec + n0;
  DRec dc = drec_at(dr, ld, nld, 0, tid), dn = drec_at(dr, ld, nld, 1, tid);
  for (int L = 0; L < nld; ++L) {
    const DRec ahead = drec_at(dr, ld, nld, L + 2, tid);
    const int a = ld[L], z = ld[L + 1];
    for (int i = a + tid; i < z; i += T) {
      const DRec r = i == a + tid ? dc : dr[i];
      const int li = r.li;
      double x = rhs[li];
      if (r.p >= 0) x += r.g * rhs[r.p];
      x = x / diag[li];
      rhs[li] = x;
      v[li] += x;
    }
    __syncthreads();
    dc = dn;
    dn = ahead;
  }
}

// Per-window constants of the step's critical path: mechanism groups, their row
// ranges for this workgroup's cells, and its detectors.
struct WinShared {
  GroupDev sg[kMaxGroups];
  int64_t sr0[kMaxGroups], sr1[kMaxGroups];
  int32_t sdet_k[kMaxDetLds], sdet_n[kMaxDetLds];
  double sdet_thr[kMaxDetLds];
};

template <int M>
__device__ __forceinline__ void window_body(const Dev& d, const BatchDev& b, const int grp, const int32_t n,
                                            WinShared& sh, double* lds) {
  const int tid = threadIdx.x, T = blockDim.x;
  const int64_t c0 = b.grp_cell0[grp], c1 = b.grp_cell1[grp];
  const int64_t n0 = d.cell_off[c0], n1 = d.cell_off[c1];
  const int nn = (int)(n1 - n0);
  const int nlh = b.grp_nlh[grp], nld = b.grp_nld[grp];
  const int ldsd = M == 0 ? 4 : M == 4 ? 2 : 0;  // doubles per node in LDS
  double* rhs = M == 3 ? d.rhs + n0 : lds;
  double* diag = M == 3 ? d.diag + n0 : lds + nn;
  double* v = M == 0 ? lds + 2 * nn : d.v + n0;
  double* axl = lds + 3 * nn;  // LDS copy (M = 0)
  const double* ax = M == 0 ? axl : d.axial + n0;
  const int tlen = (int)rtab_len(nn, nlh, nld);
  const int16_t* rt = b.r_tab + b.grp_rtab[grp];
  int16_t* tabl = reinterpret_cast<int16_t*>(lds + ldsd * nn);  // LDS copy (M = 0; level offsets M = 4)
  const int16_t* tab = M == 0 ? tabl : rt;
  const int16_t* par = tab;          // parent, -1 at a root
  const int16_t* clo = par + nn;     // lower-index child, -1 if none
  const int16_t* chi = clo + nn;     // higher-index child; -1 if fewer than two, -2 if more
  const int16_t* oh = chi + nn;      // nodes by height
  const int16_t* od = oh + nn;       // nodes by depth
  const int16_t* lh = M == 4 ? tabl : od + nn;  // nlh + 1 offsets into oh (LDS copy in mode 4)
  const int16_t* ld = lh + nlh + 1;             // nld + 1 offsets into od
  GroupDev* sg = sh.sg;
  int64_t* sr0 = sh.sr0;
  int64_t* sr1 = sh.sr1;
  int32_t* sdet_k = sh.sdet_k;
  int32_t* sdet_n = sh.sdet_n;
  double* sdet_thr = sh.sdet_thr;
  const int ngr = d.ngroups < kMaxGroups ? d.ngroups : kMaxGroups;
  if (tid < ngr) {
    sg[tid] = d.g[tid];
    sr0[tid] = sg[tid].cell_rows[c0];
    sr1[tid] = sg[tid].cell_rows[c1];
  }
  if (M == 0) {
    for (int i = tid; i < nn; i += T) {
      v[i] = d.v[n0 + i];
      axl[i] = d.axial[n0 + i];
    }
    for (int i = tid; i < tlen; i += T) tabl[i] = rt[i];
  }
  if (M == 4)
    for (int i = tid; i < nlh + 1 + nld + 1; i += T) tabl[i] = rt[5 * nn + i];
  const int64_t k0 = d.clock->step;
  const double dt = d.dt;
  const int32_t bucket = d.cell_bucket[c0];
  int64_t seg = d.seg_off[bucket];
  const int64_t seg_end = d.seg_off[bucket + 1];
  const int64_t det0 = d.cell_det_off[c0], det1 = d.cell_det_off[c1];
  const int ndl = (int)(det1 - det0) <= kMaxDetLds ? (int)(det1 - det0) : 0;  // detectors cached in LDS
  for (int q = tid; q < ndl; q += T) {
    const int32_t kk = d.cell_det[det0 + q];
    sdet_k[q] = kk;
    sdet_n[q] = d.det_node[kk] - (int32_t)n0;
    sdet_thr[q] = d.det_thr[kk];
  }
  const int64_t rec0 = d.nrec ? d.cell_rec_off[c0] : 0, rec1 = d.nrec ? d.cell_rec_off[c1] : 0;
  if (d.ngroups > kMaxGroups) raise_error(d.error, MIND_EUNSUPPORTED);
  __syncthreads();
  for (int s = 0; s < n; ++s) {
    const int64_t k = k0 + s;
    const double t = (double)k * dt;
    // threshold detection from the completed boundary
    if (ndl) {
      for (int q = tid; q < ndl; q += T) {
        const int32_t kk = sdet_k[q];
        const bool cur = v[sdet_n[q]] > sdet_thr[q];
        d.det_flags[(int64_t)s * d.ndet + kk] = (cur && !d.det_above[kk]) ? 1 : 0;
        d.det_above[kk] = cur;
      }
    } else {
      for (int64_t q = det0 + tid; q < det1; q += T) {
        const int32_t kk = d.cell_det[q];
        const bool cur = v[d.det_node[kk] - n0] > d.det_thr[kk];
        d.det_flags[(int64_t)s * d.ndet + kk] = (cur && !d.det_above[kk]) ? 1 : 0;
        d.det_above[kk] = cur;
      }
    }
    if (tid == 0) seg = deliver_group(d, seg, seg_end, 2 * k);  // midpoint delivery
    for (int i = tid; i < nn; i += T) {
      rhs[i] = 0.0;
      diag[i] = 0.0;
    }
    __syncthreads();
    // nrn_cur at t + dt/2, one mechanism group after another
    const double tc = t + 0.5 * dt;
    for (int gi = 0; gi < ngr; ++gi) {
      const GroupDev& G = sg[gi];
      if (G.mech == MIND_MECH_NETSTIM) continue;
      const int64_t r0 = sr0[gi], r1 = sr1[gi];
      if (r0 == r1) continue;
      const Cols col = cols(G);
      for (int64_t r = r0 + tid; r < r1; r += T) {
        if (G.point && !G.run_start[r]) continue;
        const int32_t node = G.node[r];
        const int li = node - (int)n0;
        double acc_r = rhs[li], acc_d = diag[li];
        int64_t q = r;
        do {  // a run of rows on one node, in row order
          double value, shifted;
          phase_current(G.mech, col, q, v[li], tc, value, shifted, true);
          const double scale = G.point ? 1.0 : d.area[node] * 0.01;
          acc_r -= scale * value;
          acc_d += scale * (shifted - value) / 0.001;
          ++q;
        } while (G.point && q < r1 && G.node[q] == node);
        rhs[li] = acc_r;
        diag[li] = acc_d;
      }
      __syncthreads();
    }
    // cable assembly: own axial current first, children in increasing index
    for (int li = tid; li < nn; li += T) {
      const int64_t gi = n0 + li;
      const double dg = d.cdt[gi] + diag[li] + d.adiag[gi];
      double bb = rhs[li];
      const int p = par[li];
      const double vi = v[li];
      if (p >= 0) bb += ax[li] * (v[p] - vi);
      const int h = chi[li];
      if (h == -2) {
        for (int64_t q = d.child_off[gi]; q < d.child_off[gi + 1]; ++q) {
          const int c = d.child[q] - (int)n0;
          bb -= ax[c] * (vi - v[c]);
        }
      } else {
        const int l = clo[li];
        if (l >= 0) bb -= ax[l] * (vi - v[l]);
        if (h >= 0) bb -= ax[h] * (vi - v[h]);
      }
      diag[li] = dg;
      rhs[li] = bb;
    }
    __syncthreads();
    if (M == 4) {
      solve_lean(d, b, n0, nlh, nld, lh, ld, rhs, diag, v);
    } else {
      // elimination by height: a node pulls its children (lower levels) in
      // decreasing index
      bool bad = false;
      for (int L = 0; L < nlh; ++L) {
        const int z = lh[L + 1];
        for (int i = lh[L] + tid; i < z; i += T) {
          const int li = oh[i];
          double dd = diag[li], bb = rhs[li];
          const int h = chi[li];
          if (h == -2) {
            const int64_t gi = n0 + li;
            for (int64_t cq = d.child_off[gi + 1] - 1; cq >= d.child_off[gi]; --cq) {
              const int lc = d.child[cq] - (int)n0;
              const double g = ax[lc];
              const double f = g / diag[lc];
              dd -= f * g;
              bb += f * rhs[lc];
            }
          } else {
            if (h >= 0) {
              const double g = ax[h];
              const double f = g / diag[h];
              dd -= f * g;
              bb += f * rhs[h];
            }
            const int l = clo[li];
            if (l >= 0) {
              const double g = ax[l];
              const double f = g / diag[l];
              dd -= f * g;
              bb += f * rhs[l];
            }
          }
          diag[li] = dd;
          rhs[li] = bb;
          bad |= (dd == 0.0) || !mind_isfinite(dd);
        }
        __syncthreads();
      }
      if (bad) raise_error(d.error, MIND_ESINGULAR);
      // back substitution by depth, root to leaves; v += dv
      for (int L = 0; L < nld; ++L) {
        const int z = ld[L + 1];
        for (int i = ld[L] + tid; i < z; i += T) {
          const int li = od[i];
          const int p = par[li];
          double x = rhs[li];
          if (p >= 0) x += ax[li] * rhs[p];
          x = x / diag[li];
          rhs[li] = x;
          v[li] += x;
        }
        __syncthreads();
      }
    }
    // nrn_state at t + dt
    for (int gi = 0; gi < ngr; ++gi) {
      const GroupDev& G = sg[gi];
      if (G.mech != MIND_MECH_HH && G.mech != MIND_MECH_EXP2SYN && G.mech != MIND_MECH_EXPSYN) continue;
      const Cols col = cols(G);
      const int64_t r0 = sr0[gi], r1 = sr1[gi];
      for (int64_t r = r0 + tid; r < r1; r += T)
        phase_advance(G.mech, col, r, v[G.node[r] - n0], dt, d.celsius);
    }
    if (rec1 > rec0) {
      __syncthreads();
      const int64_t sample = k + 1 - d.rec_start;
      for (int64_t q = rec0 + tid; q < rec1; q += T) {
        const int32_t rec = d.cell_rec[q];
        const int64_t idx = d.rec_index[rec];
        const double x = d.rec_kind[rec] == 0
                             ? v[idx - n0]
                             : d.g[d.rec_group[rec]].data[d.rec_field[rec] * d.g[d.rec_group[rec]].rows + idx];
        d.rec_buf[rec * d.rec_stride + sample] = x;
      }
    }
    __syncthreads();
    if (tid == 0) seg = deliver_group(d, seg, seg_end, 2 * k + 1);  // end-of-step delivery
  }
  if (M == 0) {
    __syncthreads();
    for (int i = tid; i < nn; i += T) d.v[n0 + i] = v[i];
  }
}

// Groups within the node budget: all state in LDS.
__global__ void __launch_bounds__(kMaxThreads) k_window(Dev d, BatchDev b, const int32_t* groups, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<0>(d, b, groups[blockIdx.x], n, sh, lds);
}

// Groups within the node budget in the lean layout.
__global__ void __launch_bounds__(kMaxThreads) k_window_lean(Dev d, BatchDev b, const int32_t* groups, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<4>(d, b, groups[blockIdx.x], n, sh, lds);
}

// Cells above the node budget, one per workgroup, each with the layout that
// fits (modes[k] for groups[k]): one launch, so that they run side by side.
__global__ void __launch_bounds__(kMaxThreads) k_window_large(Dev d, BatchDev b, const int32_t* groups,
                                                              const int8_t* modes, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  const int grp = groups[blockIdx.x];
  switch (modes[blockIdx.x]) {
    case 0: window_body<0>(d, b, grp, n, sh, lds); break;
    case 4: window_body<4>(d, b, grp, n, sh, lds); break;
    default: window_body<3>(d, b, grp, n, sh, lds); break;
  }
}


namespace {

template <typename T>
int plan_upload(State& S, BatchPlan& P, T** out, const std::vector<T>& host) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(1, host.size()) * sizeof(T)) != hipSuccess)
    return fail(MIND_EOOM, "device allocation failed");
  P.allocations.push_back(p);
  if (!host.empty() && hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MIND_EHIP, "plan upload failed");
  *out = static_cast<T*>(p);
  return 0;
}

void free_plan(BatchPlan* P) {
  if (!P) return;
  for (void* p : P->allocations) (void)hipFree(p);
  delete P;
}

// Levels by height and by depth of nodes [a0, a1) (parents precede children).
void node_levels(const Layout& L, int64_t a0, int64_t a1, std::vector<int32_t>& h, std::vector<int32_t>& dp,
                 int64_t& nlh, int64_t& nld) {
  h.assign(a1 - a0, 0);
  dp.assign(a1 - a0, 0);
  for (int64_t i = a1 - 1; i >= a0; --i) {
    const int32_t p = L.parent[i];
    if (p >= 0) h[p - a0] = std::max(h[p - a0], h[i - a0] + 1);
  }
  nlh = nld = 0;
  for (int64_t i = a0; i < a1; ++i) {
    const int32_t p = L.parent[i];
    dp[i - a0] = p >= 0 ? dp[p - a0] + 1 : 0;
    nlh = std::max<int64_t>(nlh, h[i - a0] + 1);
    nld = std::max<int64_t>(nld, dp[i - a0] + 1);
  }
}

// Groups consecutive cells within a node budget and the LDS capacity, and builds
// each group's link and schedule block.
int build_plan(State& S, int64_t budget, int threads, int layout, BatchPlan*& out) {
  const Layout& L = S.L;
  BatchPlan* P = new BatchPlan();
  P->budget = budget;
  P->threads = threads;
  P->layout = layout;
  std::vector<int32_t> cell0, cell1, glist;
  std::vector<int32_t> hh, dd;
  int64_t cur_nodes = 0, cur_lh = 0, cur_ld = 0;
  int32_t start = 0;
  int rc = 0;
  for (int32_t c = 0; c <= L.ncells; ++c) {
    int64_t nc = 0, lh = 0, ld = 0;
    if (c < L.ncells) {
      nc = L.cell_off[c + 1] - L.cell_off[c];
      node_levels(L, L.cell_off[c], L.cell_off[c + 1], hh, dd, lh, ld);
    }
    const bool alone_fits = window_bytes(layout, nc, lh, ld) <= kLdsResident;
    const bool fits =
        window_bytes(layout, cur_nodes + nc, std::max(cur_lh, lh), std::max(cur_ld, ld)) <= kLdsResident;
    const bool flush = c == L.ncells || cur_nodes + nc > budget || !fits || !alone_fits;
    if (flush && c > start) {
      cell0.push_back(start);
      cell1.push_back(c);
      glist.push_back(c - start == 1 && cur_nodes > budget ? 1 : 0);
      start = c;
      cur_nodes = cur_lh = cur_ld = 0;
    }
    if (c == L.ncells) break;
    if (!alone_fits) {  // a cell beyond the LDS capacity: its own workgroup, part in HBM
      if (nc > 32767) {
        free_plan(P);
        return fail(MIND_EUNSUPPORTED, "a cell has more than 32767 nodes");
      }
      // the lean layout if it fits, else everything in HBM
      const int m = window_bytes(4, nc, lh, ld) <= kLdsResident ? 4 : 3;
      cell0.push_back(c);
      cell1.push_back(c + 1);
      glist.push_back(m == 4 ? 2 : 3);
      start = c + 1;
      continue;
    }
    cur_nodes += nc;
    cur_lh = std::max(cur_lh, lh);
    cur_ld = std::max(cur_ld, ld);
  }
  const int ng = (int)cell0.size();
  P->ngroups = ng;
  std::vector<int16_t> r_tab;
  std::vector<int32_t> grp_rtab(ng), grp_nlh(ng), grp_nld(ng);
  bool any_lean = false;
  for (int g = 0; g < ng; ++g) {
    const int64_t n0 = L.cell_off[cell0[g]], n1 = L.cell_off[cell1[g]];
    const int nn = (int)(n1 - n0);
    int64_t nlh = 0, nld = 0;
    node_levels(L, n0, n1, hh, dd, nlh, nld);
    const int list = glist[g];
    P->lists[list].push_back(g);
    const int mode = list == 0 || list == 1 ? layout : list == 2 ? 4 : 3;
    if (mode == 4) any_lean = true;
    P->lds[list] = std::max(P->lds[list], window_bytes(mode, nn, nlh, nld));
    grp_rtab[g] = (int32_t)r_tab.size();
    grp_nlh[g] = (int32_t)nlh;
    grp_nld[g] = (int32_t)nld;
    const size_t base = r_tab.size();
    r_tab.resize(base + (size_t)rtab_len(nn, nlh, nld), 0);
    int16_t* t = r_tab.data() + base;
    int16_t *t_par = t, *t_lo = t + nn, *t_hi = t + 2 * nn, *t_oh = t + 3 * nn, *t_od = t + 4 * nn,
            *t_lh = t + 5 * nn, *t_ld = t_lh + nlh + 1;
    for (int li = 0; li < nn; ++li) {
      const int64_t gi = n0 + li;
      const int32_t p = L.parent[gi];
      t_par[li] = (int16_t)(p < 0 ? -1 : p - (int32_t)n0);
      const int64_t o0 = L.child_off[gi], o1 = L.child_off[gi + 1];
      t_lo[li] = (int16_t)(o1 > o0 ? L.child[o0] - (int32_t)n0 : -1);
      t_hi[li] = (int16_t)(o1 - o0 == 2 ? L.child[o0 + 1] - (int32_t)n0 : (o1 - o0 > 2 ? -2 : -1));
    }
    auto order = [&](const std::vector<int32_t>& lv, int64_t nl, int16_t* ord, int16_t* off) {
      std::vector<int32_t> cnt(nl + 1, 0);
      for (int li = 0; li < nn; ++li) cnt[lv[li] + 1] += 1;
      for (int64_t l = 0; l < nl; ++l) cnt[l + 1] += cnt[l];
      for (int64_t l = 0; l <= nl; ++l) off[l] = (int16_t)cnt[l];
      for (int li = 0; li < nn; ++li) ord[cnt[lv[li]]++] = (int16_t)li;
    };
    order(hh, nlh, t_oh, t_lh);
    order(dd, nld, t_od, t_ld);
  }
  std::vector<HRec> hrec;
  std::vector<DRec> drec;
  if (any_lean) {  // per-node level records of the lean layout, from the int16 block
    hrec.resize(L.nodes);
    drec.resize(L.nodes);
    for (int g = 0; g < ng; ++g) {
      const int64_t n0 = L.cell_off[cell0[g]];
      const int nn = (int)(L.cell_off[cell1[g]] - n0);
      const int16_t* t = r_tab.data() + grp_rtab[g];
      const int16_t *t_par = t, *t_lo = t + nn, *t_hi = t + 2 * nn, *t_oh = t + 3 * nn, *t_od = t + 4 * nn;
      for (int i = 0; i < nn; ++i) {
        const int li = t_oh[i], h = t_hi[li], l = t_lo[li];
        hrec[n0 + i] = HRec{(int16_t)li, (int16_t)h, (int16_t)l, 0, h >= 0 ? L.axial[n0 + h] : 0.0,
                            l >= 0 ? L.axial[n0 + l] : 0.0};
        const int lj = t_od[i];
        drec[n0 + i] = DRec{(int16_t)lj, t_par[lj], 0, 0, L.axial[n0 + lj]};
      }
    }
  }
  BatchDev& b = P->b;
  int32_t* p = nullptr;
  int16_t* p16 = nullptr;
  if ((rc = plan_upload(S, *P, &p, cell0))) goto error; b.grp_cell0 = p;
  if ((rc = plan_upload(S, *P, &p, cell1))) goto error; b.grp_cell1 = p;
  if ((rc = plan_upload(S, *P, &p16, r_tab))) goto error; b.r_tab = p16;
  {
    HRec* ph = nullptr;
    DRec* pd = nullptr;
    if ((rc = plan_upload(S, *P, &ph, hrec))) goto error; b.r_hrec = ph;
    if ((rc = plan_upload(S, *P, &pd, drec))) goto error; b.r_drec = pd;
  }
  if ((rc = plan_upload(S, *P, &p, grp_rtab))) goto error; b.grp_rtab = p;
  if ((rc = plan_upload(S, *P, &p, grp_nlh))) goto error; b.grp_nlh = p;
  if ((rc = plan_upload(S, *P, &p, grp_nld))) goto error; b.grp_nld = p;
  for (int m = 0; m < kLists; ++m)
    if ((rc = plan_upload(S, *P, &P->dev_lists[m], P->lists[m]))) goto error;
  for (int m = 1; m < kLists; ++m) {  // largest layouts first: the longest workgroups start early
    const int list = kLists - m;
    for (int32_t g : P->lists[list]) {
      P->large.push_back(g);
      P->large_mode.push_back((int8_t)(list == 1 ? layout : list == 2 ? 4 : 3));
    }
    P->large_lds = std::max(P->large_lds, P->lds[list]);
  }
  if ((rc = plan_upload(S, *P, &P->dev_large, P->large))) goto error;
  {
    int8_t* p8 = nullptr;
    if ((rc = plan_upload(S, *P, &p8, P->large_mode))) goto error;
    P->dev_large_mode = p8;
  }
  for (int m = 0; m < kLists; ++m)
    if (P->lds[m] > kLdsResident) {
      rc = fail(MIND_EUNSUPPORTED, "batched plan exceeds the LDS capacity");
      goto error;
    }
  // events are bucketed per group
  S.cell_bucket.assign(L.ncells, 0);
  for (int g = 0; g < ng; ++g)
    for (int32_t c = cell0[g]; c < cell1[g]; ++c) S.cell_bucket[c] = g;
  out = P;
  return 0;
error:
  free_plan(P);
  return rc;
}



// The workgroups within the node budget launch on the main stream; larger cells
// (each a long critical path on one workgroup) run concurrently on a second
// stream, joined before the exchange.
int launch_cells(State& S, const BatchPlan& P, int32_t n) {
  const bool side = !P.large.empty();
  if (side) {
    HIP_TRY(hipEventRecord(S.fork, S.stream));
    HIP_TRY(hipStreamWaitEvent(S.aux, S.fork, 0));
    hipLaunchKernelGGL(k_window_large, dim3((unsigned)P.large.size()), dim3(P.threads), P.large_lds, S.aux, S.d,
                       P.b, P.dev_large, P.dev_large_mode, n);
    HIP_TRY(hipEventRecord(S.join, S.aux));
  }
  int rc = 0;
  if (S.profiling && !P.lists[0].empty() && (rc = timer_start(S))) return rc;
  if (!P.lists[0].empty()) {
    if (P.layout == 4)
      hipLaunchKernelGGL(k_window_lean, dim3((unsigned)P.lists[0].size()), dim3(P.threads), P.lds[0], S.stream, S.d,
                         P.b, P.dev_lists[0], n);
    else
      hipLaunchKernelGGL(k_window, dim3((unsigned)P.lists[0].size()), dim3(P.threads), P.lds[0], S.stream, S.d,
                         P.b, P.dev_lists[0], n);
  }
  if (S.profiling && !P.lists[0].empty()) {
    if ((rc = timer_stop(S))) return rc;
    S.kernel_steps += n;
  }
  if (side) HIP_TRY(hipStreamWaitEvent(S.stream, S.join, 0));
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace

int set_buckets(State& S, int64_t nb);
extern "C" int mind_initialize(mind_state* state);

void batched_release(State& S) {
  free_plan(S.batch);
  S.batch = nullptr;
}

int batched_launch(State& S, int32_t n) { return launch_cells(S, *S.batch, n); }

// Builds a plan and makes it the state's; the event buckets follow the groups.
int install_plan(State& S, int64_t budget, int32_t threads, int32_t layout) {
  BatchPlan* P = nullptr;
  int rc = build_plan(S, budget, threads, layout, P);
  if (rc) return rc;
  batched_release(S);
  S.batch = P;
  return set_buckets(S, P->ngroups);
}

// Chooses the node budget and block size by timing the window kernel on this
// model (initialize! must follow, which restores every state the trials touched).
int batched_setup(State& S, int32_t cells_per_group) {
  HIP_TRY(hipStreamCreateWithFlags(&S.aux, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&S.join, hipEventDisableTiming));
  HIP_TRY(hipFuncSetAttribute((const void*)k_window, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kLdsResident));
  HIP_TRY(hipFuncSetAttribute((const void*)k_window_large, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kLdsResident));
  HIP_TRY(hipFuncSetAttribute((const void*)k_window_lean, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kLdsResident));
  int64_t max_cell = 1;
  for (int32_t c = 0; c < S.L.ncells; ++c) max_cell = std::max<int64_t>(max_cell, S.L.cell_off[c + 1] - S.L.cell_off[c]);
  struct Cand { int64_t budget; int threads; int layout; };
  std::vector<Cand> cands;
  if (cells_per_group > 0) {
    cands.push_back({std::max<int64_t>(1, (int64_t)cells_per_group * max_cell), 256, 0});
  } else {
    for (int threads : {256, 512}) {
      for (int64_t bud : {768, 1024, 1400, 1700, 2200, 2750}) cands.push_back({bud, threads, 0});
      for (int64_t bud : {1200, 1800, 2400, 3600, 4800, 7200}) cands.push_back({bud, threads, 4});
    }
  }
  int rc;
  double best = 1e30;
  Cand pick = cands[0];
  const bool tune = cands.size() > 1;
  auto trial = [&](const Cand& c, double& secs) -> int {
    int rc2 = install_plan(S, c.budget, c.threads, c.layout);
    if (rc2) return rc2;
    BatchPlan* P = S.batch;
    if (!tune) return 0;
    if ((rc2 = mind_initialize(reinterpret_cast<mind_state*>(&S)))) return rc2;
    const int32_t n = S.wmax;
    for (int rep = 0; rep < 2; ++rep) {  // first repetition warms up
      HIP_TRY(hipStreamSynchronize(S.stream));
      const auto t0 = std::chrono::steady_clock::now();
      for (int w = 0; w < 3; ++w)
        if ((rc2 = launch_cells(S, *P, n))) return rc2;
      HIP_TRY(hipStreamSynchronize(S.stream));
      secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
  };
  for (const Cand& c : cands) {
    double secs = 0;
    if ((rc = trial(c, secs))) return rc;
    if (tune && secs < best) {
      best = secs;
      pick = c;
    }
  }
  if (tune) {
    double secs = 0;
    if ((rc = trial(pick, secs))) return rc;
  }
  HIP_TRY(hipMemset(S.d.error, 0, sizeof(int32_t)));
  S.initialized = false;
  return 0;
}

// Nodes, pas/hh/Exp2Syn rows and detectors of the cells in the workgroups within
// the node budget: the work of the dominant launch (list 0).
int batched_dominant_work(const State& S, int64_t* counts) {
  if (!S.batch) return fail(MIND_EARG, "state has no batched plan");
  for (int i = 0; i < 6; ++i) counts[i] = 0;
  const Layout& L = S.L;
  std::vector<uint8_t> in0(L.ncells, 0);
  std::vector<int32_t> c0(S.batch->ngroups), c1(S.batch->ngroups);
  hipMemcpy(c0.data(), S.batch->b.grp_cell0, sizeof(int32_t) * c0.size(), hipMemcpyDeviceToHost);
  hipMemcpy(c1.data(), S.batch->b.grp_cell1, sizeof(int32_t) * c1.size(), hipMemcpyDeviceToHost);
  for (int32_t g : S.batch->lists[0])
    for (int32_t c = c0[g]; c < c1[g]; ++c) in0[c] = 1;
  for (int32_t c = 0; c < L.ncells; ++c) {
    if (!in0[c]) continue;
    counts[0] += L.cell_off[c + 1] - L.cell_off[c];
    counts[5] += L.cell_det_off[c + 1] - L.cell_det_off[c];
    const int mechs[4] = {MIND_MECH_PAS, MIND_MECH_HH, MIND_MECH_EXP2SYN, MIND_MECH_ICLAMP};
    for (int k = 0; k < 4; ++k) {
      const int g = L.group_of_mech[mechs[k]];
      if (g >= 0) counts[1 + k] += L.groups[g].cell_rows[c + 1] - L.groups[g].cell_rows[c];
    }
  }
  return 0;
}

int batched_plan_info(const State& S, int64_t* budget, int32_t* threads, int32_t* layout, int64_t* groups) {
  if (!S.batch) return fail(MIND_EARG, "state has no batched plan");
  *budget = S.batch->budget;
  *threads = S.batch->threads;
  *layout = S.batch->layout;
  *groups = S.batch->ngroups;
  return 0;
}

// Replaces the autotuner's choice; the state must be initialized again.
int batched_replan(State& S, int64_t budget, int32_t threads, int32_t layout) {
  if (!S.batch) return fail(MIND_EARG, "state has no batched plan");
  if (budget < 1) return fail(MIND_EARG, "node budget must be positive");
  if (threads < 64 || threads > kMaxThreads || threads % 64) return fail(MIND_EARG, "threads must be 64, 128, ..., 512");
  if (layout != 0 && layout != 4) return fail(MIND_EARG, "layout must be 0 (all in LDS) or 4 (lean)");
  HIP_TRY(hipStreamSynchronize(S.stream));
  if (S.graph) {  // the captured window holds the old plan
    HIP_TRY(hipGraphExecDestroy(S.graph));
    S.graph = nullptr;
  }
  S.initialized = false;
  return install_plan(S, budget, threads, layout);
}

}  // namespace mind

extern "C" int mind_batch_plan(const mind_state* state, int64_t* budget, int32_t* threads, int32_t* layout,
                               int64_t* groups) {
  return mind::batched_plan_info(*reinterpret_cast<const mind::State*>(state), budget, threads, layout, groups);
}

extern "C" int mind_set_batch_plan(mind_state* state, int64_t budget, int32_t threads, int32_t layout) {
  mind::State& S = *reinterpret_cast<mind::State*>(state);
  HIP_TRY(hipSetDevice(S.device));
  return mind::batched_replan(S, budget, threads, layout);
}

extern "C" int mind_batch_work(const mind_state* state, int64_t* counts) {
  return mind::batched_dominant_work(*reinterpret_cast<const mind::State*>(state), counts);
}

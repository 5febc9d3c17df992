// Batched window kernel: a workgroup advances a group of consecutive cells
// through a whole window with their node state (v, rhs, diag) staged in LDS.
//
// The Hines solve of the reference runs one warp per group of cells with
// 32-lane "DHS" rounds and FP64 atomics into parents (src/runtime/cuda/micro.jl:688-749,
// plan src/model/micro/model.jl:906-1077). Here each cell's tree is cut into
// *chain units*: maximal runs of nodes whose parent has exactly one child,
// starting from a node with zero or several children. Elimination processes
// units level by level from the leaves (a unit's level is one more than the
// highest level among its bottom node's child units), each thread walking one
// unit bottom-up; a node *pulls* its children's contributions in decreasing
// child index, which reproduces the CPU restatement's elimination order
// exactly and needs no atomics. Back substitution walks the units top-down
// by depth level. Workgroup barriers separate levels.
//
// Cells too large for LDS run with v (mode 1) or all node arrays (mode 2) in
// HBM; the autotuner chooses the node budget per workgroup and the block size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "device.h"
#include "state.h"

namespace mind {

struct BatchDev {
  const int32_t* grp_cell0;
  const int32_t* grp_cell1;
  const int32_t* unit_start;  // into unit_nodes
  const int32_t* unit_len;
  const int32_t* unit_nodes;  // local node index, bottom first
  const int32_t* tri_order;   // unit ids by (group, level)
  const int32_t* tri_lvl;     // level offsets into tri_order
  const int32_t* grp_tri;     // per group: first entry in tri_lvl
  const int32_t* grp_ntri;    // per group: number of levels
  const int32_t* bk_order;
  const int32_t* bk_lvl;
  const int32_t* grp_bk;
  const int32_t* grp_nbk;
};

struct BatchPlan {
  int64_t budget = 0;  // node budget per workgroup
  int64_t ngroups = 0;
  int threads = 256;
  // launch lists: 0 groups within the budget (LDS), 1 v in HBM, 2 all in HBM,
  // 3 single cells above the budget that still fit in LDS
  std::vector<int32_t> groups_by_mode[4];
  int32_t* dev_groups[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t max_nodes[4] = {0, 0, 0, 0};
  BatchDev b{};
  std::vector<void*> allocations;
};

constexpr int kMaxThreads = 512;
constexpr size_t kLdsLimit = 160 * 1024;

template <int MODE>
__global__ void __launch_bounds__(kMaxThreads) k_batched(Dev d, BatchDev b, const int32_t* groups, int32_t n) {
  extern __shared__ double lds[];
  const int grp = groups[blockIdx.x];
  const int tid = threadIdx.x, T = blockDim.x;
  const int64_t c0 = b.grp_cell0[grp], c1 = b.grp_cell1[grp];
  const int64_t n0 = d.cell_off[c0], n1 = d.cell_off[c1];
  const int nn = (int)(n1 - n0);
  double* rhs = MODE == 2 ? d.rhs + n0 : lds;
  double* diag = MODE == 2 ? d.diag + n0 : lds + nn;
  double* v = MODE == 0 ? lds + 2 * nn : d.v + n0;
  if (MODE == 0)
    for (int i = tid; i < nn; i += T) v[i] = d.v[n0 + i];
  const int64_t k0 = d.clock->step;
  const double dt = d.dt;
  const int32_t bucket = d.cell_bucket[c0];
  int64_t seg = d.seg_off[bucket];
  const int64_t seg_end = d.seg_off[bucket + 1];
  const int64_t det0 = d.cell_det_off[c0], det1 = d.cell_det_off[c1];
  const int64_t rec0 = d.nrec ? d.cell_rec_off[c0] : 0, rec1 = d.nrec ? d.cell_rec_off[c1] : 0;
  const int tri_base = b.grp_tri[grp], ntri = b.grp_ntri[grp];
  const int bk_base = b.grp_bk[grp], nbk = b.grp_nbk[grp];
  __syncthreads();
  for (int s = 0; s < n; ++s) {
    const int64_t k = k0 + s;
    const double t = (double)k * dt;
    // threshold detection from the completed boundary
    for (int64_t q = det0 + tid; q < det1; q += T) {
      const int32_t kk = d.cell_det[q];
      const bool cur = v[d.det_node[kk] - n0] > d.det_thr[kk];
      d.det_flags[(int64_t)s * d.ndet + kk] = (cur && !d.det_above[kk]) ? 1 : 0;
      d.det_above[kk] = cur;
    }
    if (tid == 0) {  // midpoint delivery
      while (seg < seg_end && d.win[seg].call == 2 * k) {
        const int32_t e = d.win[seg].edge;
        const int32_t target = d.edge_target[e];
        const GroupDev& G = d.g[d.point_group[target]];
        phase_receive(G.mech, cols(G), d.point_row[target], d.edge_weight[e]);
        ++seg;
      }
    }
    for (int i = tid; i < nn; i += T) { rhs[i] = 0.0; diag[i] = 0.0; }
    __syncthreads();
    // nrn_cur at t + dt/2, one mechanism group after another
    const double tc = t + 0.5 * dt;
    for (int gi = 0; gi < d.ngroups; ++gi) {
      const GroupDev& G = d.g[gi];
      if (G.mech == MIND_MECH_NETSTIM) continue;
      const int64_t r0 = G.cell_rows[c0], r1 = G.cell_rows[c1];
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
      const int32_t p = d.parent[gi];
      const double vi = v[li];
      if (p >= 0) bb += d.axial[gi] * (v[p - n0] - vi);
      for (int64_t q = d.child_off[gi]; q < d.child_off[gi + 1]; ++q) {
        const int32_t c = d.child[q];
        bb -= d.axial[c] * (vi - v[c - n0]);
      }
      diag[li] = dg;
      rhs[li] = bb;
    }
    __syncthreads();
    // elimination, leaves to root
    bool bad = false;
    for (int L = 0; L < ntri; ++L) {
      const int a = b.tri_lvl[tri_base + L], z = b.tri_lvl[tri_base + L + 1];
      for (int ui = a + tid; ui < z; ui += T) {
        const int u = b.tri_order[ui];
        const int st = b.unit_start[u], len = b.unit_len[u];
        for (int q = 0; q < len; ++q) {
          const int li = b.unit_nodes[st + q];
          const int64_t gi = n0 + li;
          double dd = diag[li], bb = rhs[li];
          for (int64_t cq = d.child_off[gi + 1] - 1; cq >= d.child_off[gi]; --cq) {
            const int32_t c = d.child[cq];
            const int lc = c - (int)n0;
            const double g = d.axial[c];
            const double f = g / diag[lc];
            dd -= f * g;
            bb += f * rhs[lc];
          }
          diag[li] = dd;
          rhs[li] = bb;
          bad |= (dd == 0.0) || !mind_isfinite(dd);
        }
      }
      __syncthreads();
    }
    if (bad) raise_error(d.error, MIND_ESINGULAR);
    // back substitution, root to leaves; v += dv
    for (int L = 0; L < nbk; ++L) {
      const int a = b.bk_lvl[bk_base + L], z = b.bk_lvl[bk_base + L + 1];
      for (int ui = a + tid; ui < z; ui += T) {
        const int u = b.bk_order[ui];
        const int st = b.unit_start[u], len = b.unit_len[u];
        for (int q = len - 1; q >= 0; --q) {
          const int li = b.unit_nodes[st + q];
          const int64_t gi = n0 + li;
          double x = rhs[li];
          const int32_t p = d.parent[gi];
          if (p >= 0) x += d.axial[gi] * rhs[p - n0];
          x = x / diag[li];
          rhs[li] = x;
          v[li] += x;
        }
      }
      __syncthreads();
    }
    // nrn_state at t + dt
    for (int gi = 0; gi < d.ngroups; ++gi) {
      const GroupDev& G = d.g[gi];
      if (G.mech != MIND_MECH_HH && G.mech != MIND_MECH_EXP2SYN && G.mech != MIND_MECH_EXPSYN) continue;
      const Cols col = cols(G);
      const int64_t r0 = G.cell_rows[c0], r1 = G.cell_rows[c1];
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
    if (tid == 0) {  // end-of-step delivery
      while (seg < seg_end && d.win[seg].call == 2 * k + 1) {
        const int32_t e = d.win[seg].edge;
        const int32_t target = d.edge_target[e];
        const GroupDev& G = d.g[d.point_group[target]];
        phase_receive(G.mech, cols(G), d.point_row[target], d.edge_weight[e]);
        ++seg;
      }
    }
  }
  if (MODE == 0) {
    __syncthreads();
    for (int i = tid; i < nn; i += T) d.v[n0 + i] = v[i];
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

// Builds groups, chain units and level schedules for a node budget per workgroup.
int build_plan(State& S, int64_t budget, int threads, BatchPlan*& out) {
  const Layout& L = S.L;
  BatchPlan* P = new BatchPlan();
  P->budget = budget;
  P->threads = threads;
  std::vector<int32_t> cell0, cell1, gmode;
  const int64_t lds_nodes0 = kLdsLimit / 24, lds_nodes1 = kLdsLimit / 16;
  int64_t cur_nodes = 0;
  int32_t start = 0;
  for (int32_t c = 0; c <= L.ncells; ++c) {
    const int64_t nc = c < L.ncells ? L.cell_off[c + 1] - L.cell_off[c] : 0;
    const bool flush = c == L.ncells || cur_nodes + nc > budget || nc > lds_nodes0;
    if (flush && c > start) {
      cell0.push_back(start);
      cell1.push_back(c);
      gmode.push_back(0);
      start = c;
      cur_nodes = 0;
    }
    if (c == L.ncells) break;
    if (nc > lds_nodes0) {  // oversized cell: its own workgroup, v (or all) in HBM
      cell0.push_back(c);
      cell1.push_back(c + 1);
      gmode.push_back(nc <= lds_nodes1 ? 1 : 2);
      start = c + 1;
      cur_nodes = 0;
      continue;
    }
    cur_nodes += nc;
  }
  const int ng = (int)cell0.size();
  P->ngroups = ng;
  std::vector<int32_t> unit_start, unit_len, unit_nodes, tri_order, tri_lvl, grp_tri, grp_ntri, bk_order, bk_lvl,
      grp_bk, grp_nbk;
  std::vector<int32_t> unit_of, height, depth;
  for (int g = 0; g < ng; ++g) {
    const int64_t n0 = L.cell_off[cell0[g]], n1 = L.cell_off[cell1[g]];
    const int nn = (int)(n1 - n0);
    const int list = gmode[g] == 0 && nn > budget ? 3 : gmode[g];
    P->max_nodes[list] = std::max<int64_t>(P->max_nodes[list], nn);
    P->groups_by_mode[list].push_back(g);
    unit_of.assign(nn, -1);
    const int ubase = (int)unit_start.size();
    std::vector<int32_t> uh;  // unit heights (local unit index)
    std::vector<std::vector<int32_t>> members;
    for (int li = nn - 1; li >= 0; --li) {
      const int64_t gi = n0 + li;
      const int64_t nch = L.child_off[gi + 1] - L.child_off[gi];
      if (nch == 1) {
        const int u = unit_of[L.child[L.child_off[gi]] - n0];
        unit_of[li] = u;
        members[u].push_back(li);
      } else {
        int h = 0;
        for (int64_t q = L.child_off[gi]; q < L.child_off[gi + 1]; ++q)
          h = std::max(h, uh[unit_of[L.child[q] - n0]] + 1);
        unit_of[li] = (int)members.size();
        members.push_back({li});
        uh.push_back(h);
      }
    }
    const int nu = (int)members.size();
    std::vector<int32_t> ud(nu, 0);
    // depth by increasing top index (a parent unit's top precedes its children's)
    std::vector<int32_t> by_top(nu);
    for (int u = 0; u < nu; ++u) by_top[u] = u;
    std::sort(by_top.begin(), by_top.end(), [&](int a, int b2) { return members[a].back() < members[b2].back(); });
    for (int u : by_top) {
      const int top = members[u].back();
      const int32_t p = L.parent[n0 + top];
      ud[u] = p < 0 ? 0 : ud[unit_of[p - n0]] + 1;
    }
    for (int u = 0; u < nu; ++u) {
      unit_start.push_back((int32_t)unit_nodes.size());
      unit_len.push_back((int32_t)members[u].size());
      for (int32_t x : members[u]) unit_nodes.push_back(x);
    }
    auto levels = [&](const std::vector<int32_t>& lv, std::vector<int32_t>& order, std::vector<int32_t>& off,
                      std::vector<int32_t>& gbase, std::vector<int32_t>& gn) {
      int nl = 0;
      for (int u = 0; u < nu; ++u) nl = std::max(nl, lv[u] + 1);
      std::vector<int32_t> cnt(nl + 1, 0);
      for (int u = 0; u < nu; ++u) cnt[lv[u] + 1] += 1;
      for (int l = 0; l < nl; ++l) cnt[l + 1] += cnt[l];
      const int obase = (int)order.size();
      order.resize(obase + nu);
      std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
      for (int u = 0; u < nu; ++u) order[obase + pos[lv[u]]++] = ubase + u;
      gbase.push_back((int32_t)off.size());
      gn.push_back(nl);
      for (int l = 0; l <= nl; ++l) off.push_back(obase + cnt[l]);
    };
    levels(uh, tri_order, tri_lvl, grp_tri, grp_ntri);
    levels(ud, bk_order, bk_lvl, grp_bk, grp_nbk);
  }
  int rc = 0;
  BatchDev& b = P->b;
  int32_t* p;
  if ((rc = plan_upload(S, *P, &p, cell0))) goto error; b.grp_cell0 = p;
  if ((rc = plan_upload(S, *P, &p, cell1))) goto error; b.grp_cell1 = p;
  if ((rc = plan_upload(S, *P, &p, unit_start))) goto error; b.unit_start = p;
  if ((rc = plan_upload(S, *P, &p, unit_len))) goto error; b.unit_len = p;
  if ((rc = plan_upload(S, *P, &p, unit_nodes))) goto error; b.unit_nodes = p;
  if ((rc = plan_upload(S, *P, &p, tri_order))) goto error; b.tri_order = p;
  if ((rc = plan_upload(S, *P, &p, tri_lvl))) goto error; b.tri_lvl = p;
  if ((rc = plan_upload(S, *P, &p, grp_tri))) goto error; b.grp_tri = p;
  if ((rc = plan_upload(S, *P, &p, grp_ntri))) goto error; b.grp_ntri = p;
  if ((rc = plan_upload(S, *P, &p, bk_order))) goto error; b.bk_order = p;
  if ((rc = plan_upload(S, *P, &p, bk_lvl))) goto error; b.bk_lvl = p;
  if ((rc = plan_upload(S, *P, &p, grp_bk))) goto error; b.grp_bk = p;
  if ((rc = plan_upload(S, *P, &p, grp_nbk))) goto error; b.grp_nbk = p;
  for (int m = 0; m < 4; ++m)
    if ((rc = plan_upload(S, *P, &P->dev_groups[m], P->groups_by_mode[m]))) goto error;
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

size_t lds_bytes(const BatchPlan& P, int mode) {
  const int64_t nn = P.max_nodes[mode];
  return (mode == 0 || mode == 3) ? (size_t)nn * 24 : mode == 1 ? (size_t)nn * 16 : 0;
}

template <int MODE>
void launch_list(State& S, const BatchPlan& P, int list, int32_t n, hipStream_t st) {
  const unsigned ng = (unsigned)P.groups_by_mode[list].size();
  if (ng)
    hipLaunchKernelGGL(k_batched<MODE>, dim3(ng), dim3(P.threads), lds_bytes(P, list), st, S.d, P.b,
                       P.dev_groups[list], n);
}

// The workgroups within the node budget launch on the main stream; oversized
// cells (each a long critical path on one workgroup) run concurrently on a
// second stream, joined before the exchange.
int launch_cells(State& S, const BatchPlan& P, int32_t n) {
  const bool side = !P.groups_by_mode[1].empty() || !P.groups_by_mode[2].empty() ||
                    !P.groups_by_mode[3].empty();
  if (side) {
    HIP_TRY(hipEventRecord(S.fork, S.stream));
    HIP_TRY(hipStreamWaitEvent(S.aux, S.fork, 0));
    launch_list<2>(S, P, 2, n, S.aux);
    launch_list<1>(S, P, 1, n, S.aux);
    launch_list<0>(S, P, 3, n, S.aux);
    HIP_TRY(hipEventRecord(S.join, S.aux));
  }
  int rc = 0;
  if (S.profiling && !P.groups_by_mode[0].empty() && (rc = timer_start(S))) return rc;
  launch_list<0>(S, P, 0, n, S.stream);
  if (S.profiling && !P.groups_by_mode[0].empty()) {
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

// Chooses the node budget and block size by timing the window kernel on this
// model (initialize! must follow, which restores every state the trials touched).
int batched_setup(State& S, int32_t cells_per_group) {
  HIP_TRY(hipStreamCreateWithFlags(&S.aux, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&S.join, hipEventDisableTiming));
  HIP_TRY(hipFuncSetAttribute((const void*)k_batched<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsLimit));
  HIP_TRY(hipFuncSetAttribute((const void*)k_batched<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsLimit));
  int64_t max_cell = 1;
  for (int32_t c = 0; c < S.L.ncells; ++c) max_cell = std::max<int64_t>(max_cell, S.L.cell_off[c + 1] - S.L.cell_off[c]);
  struct Cand { int64_t budget; int threads; };
  std::vector<Cand> cands;
  if (cells_per_group > 0) {
    cands.push_back({std::max<int64_t>(1, (int64_t)cells_per_group * max_cell), 256});
  } else {
    for (int64_t bud : {512, 1024, 2048, 3072, 4096, 6144}) cands.push_back({bud, 256});
  }
  int rc;
  double best = 1e30;
  Cand pick = cands[0];
  const bool tune = cands.size() > 1;
  auto trial = [&](const Cand& c, double& secs) -> int {
    BatchPlan* P = nullptr;
    int rc2 = build_plan(S, c.budget, c.threads, P);
    if (rc2) return rc2;
    batched_release(S);
    S.batch = P;
    if ((rc2 = set_buckets(S, P->ngroups))) return rc2;
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
    for (int threads : {128, 512}) {
      double secs = 0;
      const Cand c{pick.budget, threads};
      if ((rc = trial(c, secs))) return rc;
      if (secs < best) {
        best = secs;
        pick = c;
      }
    }
    double secs = 0;
    if ((rc = trial(pick, secs))) return rc;
  }
  HIP_TRY(hipMemset(S.d.error, 0, sizeof(int32_t)));
  S.initialized = false;
  return 0;
}

// Nodes, pas/hh/Exp2Syn rows and detectors of the cells in LDS workgroups within
// the node budget: the work of the dominant (mode 0) launch.
int batched_dominant_work(const State& S, int64_t* counts) {
  if (!S.batch) return fail(MIND_EARG, "state has no batched plan");
  for (int i = 0; i < 6; ++i) counts[i] = 0;
  const Layout& L = S.L;
  std::vector<uint8_t> in0(L.ncells, 0);
  std::vector<int32_t> c0(S.batch->ngroups), c1(S.batch->ngroups);
  hipMemcpy(c0.data(), S.batch->b.grp_cell0, sizeof(int32_t) * c0.size(), hipMemcpyDeviceToHost);
  hipMemcpy(c1.data(), S.batch->b.grp_cell1, sizeof(int32_t) * c1.size(), hipMemcpyDeviceToHost);
  for (int32_t g : S.batch->groups_by_mode[0])
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

int batched_plan_info(const State& S, int64_t* budget, int32_t* threads, int64_t* groups) {
  if (!S.batch) return fail(MIND_EARG, "state has no batched plan");
  *budget = S.batch->budget;
  *threads = S.batch->threads;
  *groups = S.batch->ngroups;
  return 0;
}

}  // namespace mind

extern "C" int mind_batch_plan(const mind_state* state, int64_t* budget, int32_t* threads, int64_t* groups) {
  return mind::batched_plan_info(*reinterpret_cast<const mind::State*>(state), budget, threads, groups);
}

extern "C" int mind_batch_work(const mind_state* state, int64_t* counts) {
  return mind::batched_dominant_work(*reinterpret_cast<const mind::State*>(state), counts);
}

// Batched window kernels: a workgroup advances a group of consecutive cells
// through a whole window (W steps of the micro step) with its node state staged
// in LDS as far as it fits.
//
// The Hines solve of the reference runs one warp per group of cells with
// 32-lane "DHS" rounds and FP64 atomics into parents (src/runtime/cuda/micro.jl:688-749,
// plan src/model/micro/model.jl:906-1077). Here each cell of a group belongs to
// one wavefront, which solves its cells pass by pass, a pass being at most 64
// nodes of one depth level: elimination in decreasing depth (a node's children
// are one level deeper), each node *pulling* its children's contributions in
// decreasing child index, which reproduces the CPU restatement's elimination
// order exactly and needs no atomics; back substitution in increasing depth. A
// pass needs no workgroup barrier: the lanes of a wave run in lockstep and LDS
// executes a wave's accesses in order, and the waves' independent chains hide
// each other's latency. The host assigns cells to waves (balancing passes of 64
// lanes) and builds the int16 node links, each wave's two node orders and pass
// offsets per group. Layouts of a workgroup's node state: all in LDS (mode 0: v,
// rhs, pivots, axial conductances, links and schedule, so the solve reads no
// global memory); lean (mode 5: (pivot, rhs) pairs and v in LDS, 24 bytes per
// node; mode 4: the pairs only, 16 bytes per node), where the solve's per-node
// records stream from HBM three passes ahead so that more nodes fit a CU; and
// everything in HBM for cells too large for the others (mode 3). The autotuner
// chooses the layout, the node budget per workgroup and the block size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "device.h"
#include "state.h"
#include "batched_kernels.h"

namespace mind {

// Launch lists: 0 groups within the node budget (main stream); 1 single cells
// above the budget, in the plan's layout; cells too large for that: 2 lean with
// v in LDS, 3 lean with v in HBM, 4 all in HBM (lists 1-4 on the second stream).
constexpr int kLists = 5;
// Mode of the workgroups of a list under a plan layout.
inline int list_mode(int list, int layout) { return list <= 1 ? layout : list == 2 ? 5 : list == 3 ? 4 : 3; }

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
// A batched plan on the host: the groups, their launch lists and LDS sizes, the
// int16 table blocks and the records of the lean and HBM layouts.
struct PlanHost {
  int64_t budget = 0;
  int threads = 256, layout = 0;
  std::vector<int32_t> cell0, cell1, glist;
  std::vector<int16_t> r_tab;
  std::vector<int32_t> grp_rtab, grp_slen, rbase;
  std::vector<HRec> hrec;
  std::vector<DRec> drec;
  std::vector<ARec> arec;
  std::vector<int32_t> lists[kLists];
  size_t lds[kLists] = {};
  std::vector<int32_t> large;
  std::vector<int8_t> large_mode;
  size_t large_lds = 0;
};

// Assigns a group's cells to the W wavefronts of its workgroup for the tree
// solve: cells in decreasing order of their own cost, each to the wave whose
// cost grows least (ties: the lower wave). A wave's cost is its passes of 64
// lanes summed over its height and depth levels (the cells' levels merge).
std::vector<int> assign_waves(const std::vector<std::vector<int32_t>>& hist_h,
                              const std::vector<std::vector<int32_t>>& hist_d, int W) {
  const int nc = (int)hist_h.size();
  auto passes = [](int32_t n) { return (n + 63) / 64; };
  std::vector<int64_t> own(nc, 0);
  for (int c = 0; c < nc; ++c) {
    for (int32_t n : hist_h[c]) own[c] += passes(n);
    for (int32_t n : hist_d[c]) own[c] += passes(n);
  }
  std::vector<int> ord(nc);
  for (int c = 0; c < nc; ++c) ord[c] = c;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return own[a] > own[b]; });
  std::vector<std::vector<int32_t>> wh(W), wd(W);
  std::vector<int64_t> cost(W, 0);
  std::vector<int> wave(nc, 0);
  for (int c : ord) {
    int best = 0;
    int64_t best_cost = -1;
    for (int w = 0; w < W; ++w) {
      int64_t nw = cost[w];
      for (size_t l = 0; l < hist_h[c].size(); ++l) {
        const int32_t have = l < wh[w].size() ? wh[w][l] : 0;
        nw += passes(have + hist_h[c][l]) - passes(have);
      }
      for (size_t l = 0; l < hist_d[c].size(); ++l) {
        const int32_t have = l < wd[w].size() ? wd[w][l] : 0;
        nw += passes(have + hist_d[c][l]) - passes(have);
      }
      if (best_cost < 0 || nw < best_cost) {
        best = w;
        best_cost = nw;
      }
    }
    wave[c] = best;
    cost[best] = best_cost;
    if (wh[best].size() < hist_h[c].size()) wh[best].resize(hist_h[c].size(), 0);
    if (wd[best].size() < hist_d[c].size()) wd[best].resize(hist_d[c].size(), 0);
    for (size_t l = 0; l < hist_h[c].size(); ++l) wh[best][l] += hist_h[c][l];
    for (size_t l = 0; l < hist_d[c].size(); ++l) wd[best][l] += hist_d[c][l];
  }
  return wave;
}

// Units of the lean layouts' solve over nodes [n0, n0 + nn) of one group: pairs
// nodes top-down, a node that is not itself a bottom and has exactly one child
// taking that child as its bottom (with pair = false every node is alone).
// bottom[li] marks bottoms; below[li] is a top's bottom, -1 if it is alone.
void pair_units(const Layout& L, int64_t n0, int nn, bool pair, std::vector<uint8_t>& bottom,
                std::vector<int32_t>& below) {
  bottom.assign(nn, 0);
  below.assign(nn, -1);
  if (!pair) return;
  for (int li = 0; li < nn; ++li) {
    const int32_t pg = L.parent[n0 + li];
    const int p = pg < 0 ? -1 : pg - (int32_t)n0;
    if (p >= 0 && !bottom[p] && L.child_off[n0 + p + 1] - L.child_off[n0 + p] == 1) {
      bottom[li] = 1;
      below[p] = li;
    }
  }
}

// Splits the costliest cell of a lean group across waves when that shortens the
// solve. Its units at unit depth >= D form subtrees, one per unit at depth D,
// that run as stage-1 items beside the other cells; its units above D form the
// top (stage 2, wave 0). Picks the D (or no split) with the smallest estimate of
// 2 x (longest stage-1 wave) + 2 x (the top's passes) + 2 passes for the two
// barriers, a wave's passes being those of its items' entries merged by level,
// items packed as assign_waves packs cells, among the splits whose schedule block
// stays within max_slen int16 words. Updates node_wave and node_stage of the
// scheduled nodes (a unit's bottom is scheduled with its top).
void split_cell(const Layout& L, int64_t n0, int nn, int ncg, const std::vector<int32_t>& cell_of,
                const std::vector<uint8_t>& bottom, const std::vector<int32_t>& dd,
                const std::vector<std::vector<int32_t>>& hist_d, int W, int64_t max_slen,
                std::vector<int32_t>& node_wave, std::vector<uint8_t>& node_stage) {
  auto passes = [](const std::vector<int32_t>& h) {
    int64_t p = 0;
    for (int32_t x : h) p += (x + 63) / 64;
    return p;
  };
  int64_t total = 0;  // passes of all waves, of the last call of longest()
  auto longest = [&](const std::vector<std::vector<int32_t>>& items, const std::vector<int>& wave) {
    std::vector<std::vector<int32_t>> m(W);
    for (size_t k = 0; k < items.size(); ++k) {
      std::vector<int32_t>& b = m[wave[k]];
      if (b.size() < items[k].size()) b.resize(items[k].size(), 0);
      for (size_t l = 0; l < items[k].size(); ++l) b[l] += items[k][l];
    }
    int64_t best = 0;
    total = 0;
    for (int w = 0; w < W; ++w) {
      best = std::max(best, passes(m[w]));
      total += passes(m[w]);
    }
    return best;
  };
  int c = 0;
  for (int k = 1; k < ncg; ++k)
    if (passes(hist_d[k]) > passes(hist_d[c])) c = k;
  const int levels = (int)hist_d[c].size();
  if (passes(hist_d[c]) <= levels + 1) return;  // no level wider than a pass: nothing to gain
  int a0 = 0;
  while (cell_of[a0] != c) ++a0;
  int a1 = a0;
  while (a1 < nn && cell_of[a1] == c) ++a1;
  int64_t best_t = 2 * longest(hist_d, assign_waves(hist_d, hist_d, W));
  int best_d = 0;
  std::vector<int32_t> anchor(a1 - a0, -1);
  auto build = [&](int D, std::vector<std::vector<int32_t>>& items, std::vector<int32_t>& top) {
    items.clear();
    for (int k = 0; k < ncg; ++k)
      if (k != c) items.push_back(hist_d[k]);
    top.assign(D, 0);
    for (int li = a0; li < a1; ++li) {
      const int32_t pg = L.parent[n0 + li];
      const int p = pg < 0 ? -1 : pg - (int32_t)n0;
      if (dd[li] < D) {
        anchor[li - a0] = -1;
        if (!bottom[li]) top[dd[li]] += 1;
        continue;
      }
      if (dd[li] == D && !bottom[li]) {
        anchor[li - a0] = (int32_t)items.size();
        items.emplace_back(levels, 0);
      } else {
        anchor[li - a0] = anchor[p - a0];
      }
      if (!bottom[li]) items[anchor[li - a0]][dd[li]] += 1;
    }
  };
  std::vector<std::vector<int32_t>> items;
  std::vector<int32_t> top;
  int worse = 0;  // consecutive depths without a better estimate: stop at 4
  for (int D = 1; D < levels && D <= 24 && worse < 4; ++D) {
    build(D, items, top);
    const int64_t t = 2 * longest(items, assign_waves(items, items, W)) + 2 * passes(top) + 2;
    // the block: four section headers, each wave's four pass lists with their ends
    const int64_t slen = 8 * (W + 1) + 2 * (total + W) + 2 * (passes(top) + W);
    ++worse;
    if (slen > max_slen) continue;
    if (t < best_t) {
      best_t = t;
      best_d = D;
      worse = 0;
    }
  }
  if (!best_d) return;
  build(best_d, items, top);
  const std::vector<int> wave = assign_waves(items, items, W);
  std::vector<int> item_of_cell(ncg, -1);
  for (int k = 0, q = 0; k < ncg; ++k)
    if (k != c) item_of_cell[k] = q++;
  for (int li = 0; li < nn; ++li) {
    if (cell_of[li] != c) {
      node_wave[li] = wave[item_of_cell[cell_of[li]]];
      node_stage[li] = 1;
    } else if (anchor[li - a0] < 0) {
      node_wave[li] = 0;
      node_stage[li] = 2;
    } else {
      node_wave[li] = wave[anchor[li - a0]];
      node_stage[li] = 1;
    }
  }
}

// Upper bound of a schedule block's length: the headers plus, per wave, one more
// offset than its passes; a wave's levels are at most the sum of its cells', and
// splitting levels into passes of 64 adds at most nodes / 64 per order.
inline int64_t sched_bound(int W, int64_t sum_levels, int64_t nodes) {
  return 8 * (W + 1) + 4 * W + sum_levels + 2 * (nodes / 64);
}

// Groups consecutive cells within a node budget and the LDS capacity, and builds
// each group's link and schedule block.
int plan_host(const Layout& L, int64_t budget, int threads, int layout, PlanHost& H) {
  H = PlanHost();
  H.budget = budget;
  H.threads = threads;
  H.layout = layout;
  const int W0 = threads / 64;                             // waves of a list-0 workgroup
  const int Wl = std::min(threads, kLargeThreads) / 64;    // of the others
  std::vector<int32_t>& cell0 = H.cell0;
  std::vector<int32_t>& cell1 = H.cell1;
  std::vector<int32_t>& glist = H.glist;
  std::vector<int32_t> hh, dd;
  int64_t cur_nodes = 0, cur_levels = 0;
  int32_t start = 0;
  int rc = 0;
  for (int32_t c = 0; c <= L.ncells; ++c) {
    int64_t nc = 0, lh = 0, ld = 0;
    if (c < L.ncells) {
      nc = L.cell_off[c + 1] - L.cell_off[c];
      node_levels(L, L.cell_off[c], L.cell_off[c + 1], hh, dd, lh, ld);
    }
    const bool alone_fits = window_bytes(layout, nc, sched_bound(W0, lh + ld, nc)) <= kLdsResident;
    const bool fits =
        window_bytes(layout, cur_nodes + nc, sched_bound(W0, cur_levels + lh + ld, cur_nodes + nc)) <= kLdsResident;
    const bool flush = c == L.ncells || cur_nodes + nc > budget || !fits || !alone_fits;
    if (flush && c > start) {
      cell0.push_back(start);
      cell1.push_back(c);
      glist.push_back(c - start == 1 && cur_nodes > budget ? 1 : 0);
      start = c;
      cur_nodes = cur_levels = 0;
    }
    if (c == L.ncells) break;
    if (!alone_fits) {  // a cell beyond the LDS capacity: its own workgroup, part in HBM
      if (nc > 32765) return fail(MIND_EUNSUPPORTED, "a cell has more than 32765 nodes");
      // the lean layout if it fits (v in LDS if that fits too), else everything in HBM
      const int64_t sb = sched_bound(Wl, lh + ld, nc);
      cell0.push_back(c);
      cell1.push_back(c + 1);
      glist.push_back(window_bytes(5, nc, sb) <= kLdsResident ? 2 : window_bytes(4, nc, sb) <= kLdsResident ? 3 : 4);
      start = c + 1;
      continue;
    }
    cur_nodes += nc;
    cur_levels += lh + ld;
  }
  const int ng = (int)cell0.size();
  std::vector<int16_t>& r_tab = H.r_tab;
  std::vector<int32_t>& grp_rtab = H.grp_rtab;
  std::vector<int32_t>& grp_slen = H.grp_slen;
  grp_rtab.assign(ng, 0);
  grp_slen.assign(ng, 0);
  bool any_lean = false, any_hbm_links = false;
  std::vector<int64_t> group_cost(ng, 0);
  for (int g = 0; g < ng; ++g) {
    const int64_t n0 = L.cell_off[cell0[g]], n1 = L.cell_off[cell1[g]];
    const int nn = (int)(n1 - n0);
    int64_t nlh = 0, nld = 0;
    node_levels(L, n0, n1, hh, dd, nlh, nld);
    // Both sweeps run by depth: back substitution from the roots, elimination
    // in reverse (a node's children are one level deeper, so any order by
    // decreasing depth eliminates them first; depth levels are narrower than
    // height levels). The lean layouts schedule units (a node and, if it has
    // exactly one child and is not itself a bottom, that child as its bottom) by
    // their depth in the tree of units; the others schedule nodes. Per cell
    // histograms of the scheduled entries by level, then the waves.
    const int list = glist[g];
    const int mode = list_mode(list, layout);
    const bool units = mode == 4 || mode == 5;
    std::vector<uint8_t> bottom;
    std::vector<int32_t> below;
    pair_units(L, n0, nn, units, bottom, below);
    for (int li = 0; li < nn; ++li) {  // depth in the tree of units
      const int32_t pg = L.parent[n0 + li];
      const int p = pg < 0 ? -1 : pg - (int32_t)n0;
      dd[li] = p < 0 ? 0 : bottom[li] ? dd[p] : dd[p] + 1;
    }
    int32_t nsched = 0;  // scheduled entries: nodes, or the tops of units
    for (int li = 0; li < nn; ++li) nsched += bottom[li] ? 0 : 1;
    const int ncg = cell1[g] - cell0[g];
    std::vector<std::vector<int32_t>> hist_d(ncg);
    std::vector<int32_t> cell_of(nn);
    for (int k = 0; k < ncg; ++k) {
      const int64_t a = L.cell_off[cell0[g] + k] - n0, z = L.cell_off[cell0[g] + k + 1] - n0;
      for (int64_t li = a; li < z; ++li) {
        cell_of[li] = k;
        if (bottom[li]) continue;
        if ((int64_t)hist_d[k].size() <= dd[li]) hist_d[k].resize(dd[li] + 1, 0);
        hist_d[k][dd[li]] += 1;
      }
    }
    const int W = glist[g] == 0 ? W0 : Wl;
    // Waves and stages of the scheduled entries. Stage 1: whole cells, and the
    // subtrees below unit depth D of a split cell; stage 2: that cell's top (unit
    // depth < D). A lean group with a split cell eliminates stage 1, then (after a
    // workgroup barrier) eliminates and substitutes stage 2, then (after another)
    // substitutes stage 1; without one, stage 1 is everything and no barrier runs.
    std::vector<int32_t> node_wave(nn, 0);
    std::vector<uint8_t> node_stage(nn, 1);
    {
      const std::vector<int> wave = assign_waves(hist_d, hist_d, W);
      for (int li = 0; li < nn; ++li) node_wave[li] = wave[cell_of[li]];
      if (units && W > 1) {
        // the longest schedule block that keeps the workgroup within the LDS budget
        const int64_t fixed = (int64_t)window_bytes(mode, nn, 0);
        const int64_t max_slen = ((int64_t)kLdsResident - fixed) / 2 - 8;
        split_cell(L, n0, nn, ncg, cell_of, bottom, dd, hist_d, W, max_slen, node_wave, node_stage);
      }
    }
    const int nsec = units ? 4 : 2;  // schedule sections: E1, S1 (, E2, S2)
    int32_t maxd = 0;
    for (int li = 0; li < nn; ++li) maxd = std::max(maxd, dd[li]);
    for (int li = 0; li < nn; ++li) hh[li] = maxd - dd[li];  // elimination level: reversed depth
    // entries per level of each (stage, wave)
    std::vector<std::vector<int32_t>> wnh[2], wnd[2];
    for (int st = 0; st < 2; ++st) {
      wnh[st].assign(W, std::vector<int32_t>(maxd + 1, 0));
      wnd[st].assign(W, std::vector<int32_t>(maxd + 1, 0));
    }
    for (int li = 0; li < nn; ++li) {
      if (bottom[li]) continue;
      wnh[node_stage[li] - 1][node_wave[li]][hh[li]] += 1;
      wnd[node_stage[li] - 1][node_wave[li]][dd[li]] += 1;
    }
    auto npasses = [](const std::vector<int32_t>& n) {
      int64_t p = 0;
      for (int32_t x : n) p += (x + 63) / 64;
      return p;
    };
    const int64_t hdr = 2 * nsec * (W + 1);  // int16 words of the section headers
    int64_t slen = hdr;
    for (int st = 0; st < nsec / 2; ++st)
      for (int w = 0; w < W; ++w) slen += npasses(wnh[st][w]) + 1 + npasses(wnd[st][w]) + 1;
    if (slen > (1 << 30)) return fail(MIND_EUNSUPPORTED, "schedule block too long");
    H.lists[list].push_back(g);
    if (units) any_lean = true;
    if (mode != 0) any_hbm_links = true;
    H.lds[list] = std::max(H.lds[list], window_bytes(mode, nn, slen));
    {  // estimated cycles of one step: the solve's critical passes, and the
       // node-parallel phases (about 450 cycles per pass of nodes, 600 per pass
       // of units, and 5400 per node per thread)
      int64_t crit = 0;
      for (int st = 0; st < 2; ++st) {
        int64_t longest = 0;
        for (int w = 0; w < W; ++w) longest = std::max(longest, npasses(wnh[st][w]) + npasses(wnd[st][w]));
        crit += longest;
      }
      const int Tg = W * 64;
      group_cost[g] = (units ? 600 : 450) * crit + 5400 * ((nn + Tg - 1) / Tg);
    }
    r_tab.resize((r_tab.size() + 3) / 4 * 4, 0);  // 8-byte aligned blocks
    grp_rtab[g] = (int32_t)r_tab.size();
    grp_slen[g] = (int32_t)slen;
    const size_t base = r_tab.size();
    r_tab.resize(base + (size_t)rtab_len(nn, slen), 0);
    int16_t* t = r_tab.data() + base;
    int16_t *t_hn = t, *t_oh = t + 4 * nn, *t_od = t + 5 * nn, *t_s = t + 6 * nn;
    for (int li = 0; li < nn; ++li) {
      const int64_t gi = n0 + li;
      const int32_t p = L.parent[gi];
      const int64_t o0 = L.child_off[gi], o1 = L.child_off[gi + 1];
      t_hn[4 * li + 0] = (int16_t)(o1 - o0 == 2 ? L.child[o0 + 1] - (int32_t)n0 : (o1 - o0 > 2 ? -2 : -1));
      t_hn[4 * li + 1] = (int16_t)(o1 > o0 ? L.child[o0] - (int32_t)n0 : -1);
      t_hn[4 * li + 2] = (int16_t)(p < 0 ? -1 : p - (int32_t)n0);
      t_hn[4 * li + 3] = 0;
    }
    auto put = [&](int k, int64_t value) {
      t_s[2 * k] = (int16_t)(uint16_t)(value & 0xffff);
      t_s[2 * k + 1] = (int16_t)(value >> 16);
    };
    // One section: per wave its entries of the stage by level (increasing node
    // index within a level), appended to the order from position pos, and the
    // offsets of its passes into the order; the header entries of the section.
    auto order = [&](const std::vector<int32_t>& lv, const std::vector<std::vector<int32_t>>& wn, int16_t* ord,
                     int hdr0, int stage, int64_t& cursor, int32_t& pos) {
      for (int w = 0; w < W; ++w) {
        put(hdr0 + w, cursor);
        const int64_t nl = (int64_t)wn[w].size();
        std::vector<int32_t> cnt(nl + 1, 0);
        for (int64_t l = 0; l < nl; ++l) cnt[l + 1] = cnt[l] + wn[w][l];
        for (int64_t l = 0; l < nl; ++l)
          for (int32_t q = cnt[l]; q < cnt[l + 1]; q += 64) t_s[cursor++] = (int16_t)(pos + q);
        t_s[cursor++] = (int16_t)(pos + cnt[nl]);
        for (int li = 0; li < nn; ++li)
          if (node_wave[li] == w && node_stage[li] == stage && !bottom[li]) ord[pos + cnt[lv[li]]++] = (int16_t)li;
        pos += cnt[nl];  // the fill leaves cnt[nl], the wave's entry count, unchanged
      }
      put(hdr0 + W, cursor);
    };
    {
      int64_t cursor = hdr;
      int32_t ph = 0, pd = 0;
      order(hh, wnh[0], t_oh, 0, 1, cursor, ph);
      order(dd, wnd[0], t_od, W + 1, 1, cursor, pd);
      if (units) {
        order(hh, wnh[1], t_oh, 2 * (W + 1), 2, cursor, ph);
        order(dd, wnd[1], t_od, 3 * (W + 1), 2, cursor, pd);
      }
    }
    // the kernels index by these tables without bounds checks: every scheduled
    // node once per order, in its own wave and stage; passes of at most 64
    // entries of one level, in level order
    auto valid = [&](const std::vector<int32_t>& lv, const int16_t* ord, int hdr0, int stage, int32_t& pos,
                     std::vector<uint8_t>& seen) {
      for (int w = 0; w < W; ++w) {
        const int64_t a = sched_at(t_s, hdr0 + w), z = sched_at(t_s, hdr0 + w + 1);
        if (a < hdr || z <= a || z > slen || t_s[a] != pos) return false;
        int32_t level = 0;
        for (int64_t q = 0; q + 1 < z - a; ++q) {
          const int32_t i0 = t_s[a + q], i1 = t_s[a + q + 1];
          if (i1 <= i0 || i1 - i0 > 64 || i1 > nsched) return false;
          const int32_t l0 = lv[ord[i0] >= 0 && ord[i0] < nn ? ord[i0] : 0];
          if (l0 < level) return false;
          level = l0;
          for (int32_t i = i0; i < i1; ++i) {
            const int li = ord[i];
            if (li < 0 || li >= nn || seen[li] || bottom[li] || node_wave[li] != w || node_stage[li] != stage ||
                lv[li] != level)
              return false;
            seen[li] = 1;
          }
          pos = i1;
        }
      }
      return true;
    };
    {
      std::vector<uint8_t> seen_h(nn, 0), seen_d(nn, 0);
      int32_t ph = 0, pd = 0;
      bool ok = valid(hh, t_oh, 0, 1, ph, seen_h) && valid(dd, t_od, W + 1, 1, pd, seen_d);
      if (ok && units) ok = valid(hh, t_oh, 2 * (W + 1), 2, ph, seen_h) && valid(dd, t_od, 3 * (W + 1), 2, pd, seen_d);
      if (!ok || ph != nsched || pd != nsched)
        return fail(MIND_EUNSUPPORTED, "batched plan: inconsistent solve schedule");
    }
  }
  {
    std::vector<HRec>& hrec = H.hrec;
    std::vector<DRec>& drec = H.drec;
    std::vector<int32_t>& rbase = H.rbase;
    rbase.assign(ng, 0);
    if (any_lean) {  // records of the lean layout, from the int16 block; group g's from n0 + g
      hrec.resize(L.nodes + ng);
      drec.resize(L.nodes + ng);
      for (int g = 0; g < ng; ++g) {
        const int64_t n0 = L.cell_off[cell0[g]];
        const int nn = (int)(L.cell_off[cell1[g]] - n0);
        const int64_t rb = n0 + g;
        rbase[g] = (int32_t)rb;
        const int16_t dummy = (int16_t)nn, sink = (int16_t)(nn + 1);
        const int16_t* t = r_tab.data() + grp_rtab[g];
        const int16_t *t_hn = t, *t_oh = t + 4 * nn, *t_od = t + 5 * nn;
        const int mode = list_mode(glist[g], layout);
        int nu = 0;  // units (groups of the other layouts keep sink records only)
        std::vector<uint8_t> bottom;
        std::vector<int32_t> below;
        if (mode == 4 || mode == 5) {
          pair_units(L, n0, nn, true, bottom, below);
          for (int li = 0; li < nn; ++li) nu += bottom[li] ? 0 : 1;
        }
        for (int i = 0; i < nu; ++i) {
          const int u = t_oh[i], c = below[u], x = c >= 0 ? c : u;
          const int h = t_hn[4 * x], l = t_hn[4 * x + 1];
          const int16_t y = c >= 0 ? (int16_t)u : sink;
          const double gy = c >= 0 ? L.axial[n0 + c] : -0.0;
          if (h == -2) {  // more than two children: pulled from the CSR lists
            hrec[rb + i] = HRec{(int16_t)~x, dummy, dummy, y, -0.0, -0.0, gy};
          } else {
            hrec[rb + i] = HRec{(int16_t)x, h >= 0 ? (int16_t)h : dummy, l >= 0 ? (int16_t)l : dummy, y,
                                h >= 0 ? L.axial[n0 + h] : -0.0, l >= 0 ? L.axial[n0 + l] : -0.0, gy};
          }
          const int v = t_od[i], pv = t_hn[4 * v + 2], cv = below[v];
          drec[rb + i] = DRec{(int16_t)v, pv >= 0 ? (int16_t)pv : dummy, cv >= 0 ? (int16_t)cv : sink, 0,
                              pv >= 0 ? L.axial[n0 + v] : -0.0, cv >= 0 ? L.axial[n0 + cv] : -0.0};
        }
        for (int i = nu; i <= nn; ++i) {
          hrec[rb + i] = HRec{sink, dummy, dummy, sink, -0.0, -0.0, -0.0};
          drec[rb + i] = DRec{sink, dummy, sink, 0, -0.0, -0.0};
        }
      }
    }
    std::vector<ARec>& arec = H.arec;
    if (any_hbm_links) {  // assembly records, from the int16 block
      arec.resize(L.nodes);
      for (int g = 0; g < ng; ++g) {
        const int64_t n0 = L.cell_off[cell0[g]];
        const int nn = (int)(L.cell_off[cell1[g]] - n0);
        const int16_t* t_hn = r_tab.data() + grp_rtab[g];
        for (int li = 0; li < nn; ++li) {
          const int h = t_hn[4 * li], l = t_hn[4 * li + 1], p = t_hn[4 * li + 2];
          arec[n0 + li] = ARec{(int16_t)h, (int16_t)l, (int16_t)p, 0, L.axial[n0 + li],
                               h >= 0 ? L.axial[n0 + h] : 0.0, l >= 0 ? L.axial[n0 + l] : 0.0,
                               L.cap[n0 + li] / L.assembly_dt(), L.adiag[n0 + li]};  // as the cdt column
        }
      }
    }
  }
  // each launch list costliest group first, so that the last workgroups to start
  // are short ones
  for (int m = 0; m < kLists; ++m)
    std::stable_sort(H.lists[m].begin(), H.lists[m].end(),
                     [&](int32_t a, int32_t b) { return group_cost[a] > group_cost[b]; });
  for (int m = 1; m < kLists; ++m) {  // largest layouts first: the longest workgroups start early
    const int list = kLists - m;
    for (int32_t g : H.lists[list]) {
      H.large.push_back(g);
      H.large_mode.push_back((int8_t)list_mode(list, layout));
    }
    H.large_lds = std::max(H.large_lds, H.lds[list]);
  }
  for (int m = 0; m < kLists; ++m)
    if (H.lds[m] > kLdsResident) return fail(MIND_EUNSUPPORTED, "batched plan exceeds the LDS capacity");
  return 0;
}

// Runs the lean solve of group g (mode 4 or 5) on the host as the kernel does,
// pass by pass per wave with all lanes of a pass reading before any writes,
// on a seeded system with the group's axial conductances, and compares it bit
// for bit with the canonical elimination (decreasing node index, each parent
// taking its children's terms in that order; substitution in increasing index).
// Empty string if they agree.
std::string emulate_lean(const Layout& L, const PlanHost& H, int g, int W) {
  const int64_t n0 = L.cell_off[H.cell0[g]];
  const int nn = (int)(L.cell_off[H.cell1[g]] - n0);
  const int16_t* sc = H.r_tab.data() + H.grp_rtab[g] + 6 * nn;
  const int64_t rb = H.rbase[g];
  std::vector<double> d0(nn), b0(nn);
  uint64_t z = 0x9e3779b97f4a7c15ull * (uint64_t)(g + 1);
  auto uni = [&]() {  // splitmix64 -> [0, 1)
    z += 0x9e3779b97f4a7c15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return (double)((x ^ (x >> 31)) >> 11) * 0x1.0p-53;
  };
  for (int i = 0; i < nn; ++i) {
    double sum = 0.0;
    for (int64_t q = L.child_off[n0 + i]; q < L.child_off[n0 + i + 1]; ++q) sum += L.axial[L.child[q]];
    d0[i] = 1.0 + uni() + sum + (L.parent[n0 + i] >= 0 ? L.axial[n0 + i] : 0.0);
    b0[i] = uni() - 0.5;
  }
  std::vector<double> dr = d0, br = b0;  // canonical
  for (int i = nn - 1; i >= 0; --i) {
    const int32_t p = L.parent[n0 + i];
    if (p >= 0) {
      const double f = L.axial[n0 + i] / dr[i];
      dr[p - n0] -= f * L.axial[n0 + i];
      br[p - n0] += f * br[i];
    }
  }
  for (int i = 0; i < nn; ++i) {
    const int32_t p = L.parent[n0 + i];
    double value = br[i];
    if (p >= 0) value += L.axial[n0 + i] * br[p - n0];
    br[i] = value / dr[i];
  }
  std::vector<double> dg(nn + 2), bg(nn + 2);  // the LDS pairs, dummy and sink
  for (int i = 0; i < nn; ++i) {
    dg[i] = d0[i];
    bg[i] = b0[i];
  }
  dg[nn] = bg[nn] = 1.0;
  dg[nn + 1] = 1.0;
  bg[nn + 1] = 0.0;
  struct Wr {
    int k;
    double dv, bv;
  };
  std::vector<Wr> out;
  // the kernel's order: stage-1 elimination, stage-2 elimination and
  // substitution, stage-1 substitution (sections 0, 2, 3, 1)
  const int sections[4] = {0, 2, 3, 1};
  // waves of one section run concurrently: a node one wave writes in a section
  // must not be read or written by another wave in that section
  std::vector<int32_t> writer(nn + 2), reader(nn + 2);  // -1 none; reader -2: several waves
  auto write = [&](int k, int w) {
    if (k >= nn) return true;  // dummy (read only) and sink (never read as a node)
    if (writer[k] < 0) writer[k] = w;
    return writer[k] == w;
  };
  auto read = [&](int k, int w) {
    if (k < nn && reader[k] != w) reader[k] = reader[k] == -1 ? w : -2;
    return true;
  };
  for (int sec : sections) {
    std::fill(writer.begin(), writer.end(), -1);
    std::fill(reader.begin(), reader.end(), -1);
    for (int w = 0; w < W; ++w) {
      const int64_t a = sched_at(sc, sec * (W + 1) + w), e = sched_at(sc, sec * (W + 1) + w + 1);
      for (int64_t q = a; q + 1 < e; ++q)
        for (int i = sc[q]; i < sc[q + 1]; ++i) {
          bool ok = true;
          if (sec == 0 || sec == 2) {
            const HRec& r = H.hrec[rb + i];
            const int x = r.x < 0 ? ~r.x : r.x;
            ok = write(x, w) && write(r.y, w) && read(r.h, w) && read(r.l, w);
            if (r.x < 0)
              for (int64_t cq = L.child_off[n0 + x]; cq < L.child_off[n0 + x + 1]; ++cq)
                ok = ok && read(L.child[cq] - (int32_t)n0, w);
          } else {
            const DRec& r = H.drec[rb + i];
            ok = write(r.y, w) && write(r.x, w) && read(r.p, w);
          }
          if (!ok)
            return "lean schedule: two waves of section " + std::to_string(sec) + " share a node (wave " +
                   std::to_string(w) + ", position " + std::to_string(i) + ")";
        }
    }
    for (int k = 0; k < nn; ++k)
      if (writer[k] >= 0 && reader[k] != -1 && reader[k] != writer[k])
        return "lean schedule: a wave reads a node another wave of its stage writes";
  }
  for (int sec : sections)
    for (int w = 0; w < W; ++w) {
      const int64_t a = sched_at(sc, sec * (W + 1) + w), e = sched_at(sc, sec * (W + 1) + w + 1);
      for (int64_t q = a; q + 1 < e; ++q) {
        out.clear();
        for (int i = sc[q]; i < sc[q + 1]; ++i) {
          if (sec == 0 || sec == 2) {
            const HRec& r = H.hrec[rb + i];
            const int x = r.x < 0 ? ~r.x : r.x;
            const double fh = r.gh / dg[r.h], fl = r.gl / dg[r.l];
            double dd = (dg[x] - fh * r.gh) - fl * r.gl;
            double bb = (bg[x] + fh * bg[r.h]) + fl * bg[r.l];
            if (r.x < 0) {
              dd = dg[x];
              bb = bg[x];
              for (int64_t cq = L.child_off[n0 + x + 1] - 1; cq >= L.child_off[n0 + x]; --cq) {
                const int lc = L.child[cq] - (int32_t)n0;
                const double f = L.axial[n0 + lc] / dg[lc];
                dd -= f * L.axial[n0 + lc];
                bb += f * bg[lc];
              }
            }
            const double fy = r.gy / dd;
            out.push_back({x, dd, bb});
            out.push_back({r.y, dg[r.y] - fy * r.gy, bg[r.y] + fy * bb});
          } else {
            const DRec& r = H.drec[rb + i];
            const double xy = (bg[r.y] + r.gy * bg[r.p]) / dg[r.y];
            out.push_back({r.y, dg[r.y], xy});
            out.push_back({r.x, dg[r.x], (bg[r.x] + r.gx * xy) / dg[r.x]});
          }
        }
        for (const Wr& o : out) {
          if (o.k >= nn && o.k != nn + 1) return "lean solve writes the dummy";
          dg[o.k] = o.dv;
          bg[o.k] = o.bv;
        }
      }
    }
  for (int i = 0; i < nn; ++i)
    if (std::memcmp(&dg[i], &dr[i], sizeof(double)) || std::memcmp(&bg[i], &br[i], sizeof(double)))
      return "lean solve differs from the canonical elimination at node " + std::to_string(i);
  return "";
}

// Checks on the host every index the window kernels derive from a plan's tables
// (node links and orders, schedule headers and pass offsets, the records of the
// lean and HBM layouts, launch lists and LDS sizes) against the bounds of what it
// indexes; positions the kernels compute are clamped to these ranges or chosen
// between two of them. Empty string if the plan is consistent.
std::string check_plan(const Layout& L, const PlanHost& H) {
  const int ng = (int)H.cell0.size();
  const int W0 = H.threads / 64, Wl = std::min(H.threads, kLargeThreads) / 64;
  auto bad = [](int g, const char* what) { return "group " + std::to_string(g) + ": " + what; };
  if ((int)H.glist.size() != ng || (int)H.cell1.size() != ng || (int)H.grp_rtab.size() != ng ||
      (int)H.grp_slen.size() != ng || (int)H.rbase.size() != ng)
    return "per-group arrays differ in length";
  std::vector<int> in_list(ng, -1);
  for (int m = 0; m < kLists; ++m)
    for (int32_t g : H.lists[m]) {
      if (g < 0 || g >= ng || in_list[g] >= 0) return "launch list entry out of range or repeated";
      in_list[g] = m;
    }
  int32_t next_cell = 0;
  for (int g = 0; g < ng; ++g) {
    if (in_list[g] != H.glist[g]) return bad(g, "not in its launch list");
    if (H.cell0[g] != next_cell || H.cell1[g] <= H.cell0[g] || H.cell1[g] > L.ncells) return bad(g, "cell range");
    next_cell = H.cell1[g];
    const int list = H.glist[g], mode = list_mode(list, H.layout);
    const int W = list == 0 ? W0 : Wl;
    const int64_t n0 = L.cell_off[H.cell0[g]], nn = L.cell_off[H.cell1[g]] - n0;
    if (nn <= 0 || nn > 32765) return bad(g, "node count");
    const int64_t slen = H.grp_slen[g], base = H.grp_rtab[g];
    if (base < 0 || base % 4 || base + rtab_len(nn, slen) > (int64_t)H.r_tab.size()) return bad(g, "table block");
    if (window_bytes(mode, nn, slen) > H.lds[list] || H.lds[list] > kLdsResident) return bad(g, "LDS size");
    const int16_t* t = H.r_tab.data() + base;
    const int16_t *hn = t, *oh = t + 4 * nn, *od = t + 5 * nn, *sc = t + 6 * nn;
    for (int64_t li = 0; li < nn; ++li) {
      const int h = hn[4 * li], l = hn[4 * li + 1], p = hn[4 * li + 2];
      if (h < -2 || h >= nn || l < -1 || l >= nn || p < -1 || p >= nn) return bad(g, "node links");
      if (oh[li] < 0 || oh[li] >= nn || od[li] < 0 || od[li] >= nn) return bad(g, "node orders");
    }
    const int nsec = mode == 4 || mode == 5 ? 4 : 2;
    if (2 * nsec * (W + 1) > slen) return bad(g, "schedule header");
    for (int sweep = 0; sweep < nsec; ++sweep)
      for (int w = 0; w < W; ++w) {
        const int64_t a = sched_at(sc, sweep * (W + 1) + w), z = sched_at(sc, sweep * (W + 1) + w + 1);
        if (a < 2 * nsec * (W + 1) || z <= a || z > slen) return bad(g, "schedule offsets block");
        for (int64_t q = a; q < z; ++q) {
          if (sc[q] < 0 || sc[q] > nn) return bad(g, "pass offset");
          if (q + 1 < z && (sc[q + 1] < sc[q] || sc[q + 1] - sc[q] > 64)) return bad(g, "pass size");
        }
      }
    if (mode == 4 || mode == 5) {
      // records: every slot index within the nodes, the dummy (read only) and
      // the sink; each node written once per order (as x or y), the dummy
      // never; positions past the units hold sink records up to position nn
      const int64_t rb = H.rbase[g];
      if (rb < 0 || rb + nn >= (int64_t)H.hrec.size() || rb + nn >= (int64_t)H.drec.size()) return bad(g, "records");
      std::vector<uint8_t> done_h(nn, 0), done_d(nn, 0);
      int64_t nu = 0;
      for (int64_t i = 0; i <= nn; ++i) {
        const HRec& r = H.hrec[rb + i];
        const DRec& q = H.drec[rb + i];
        const int x = r.x < 0 ? ~r.x : r.x;
        if (x < 0 || x > nn + 1 || x == nn || r.h < 0 || r.h > nn + 1 || r.l < 0 || r.l > nn + 1 || r.y < 0 ||
            r.y > nn + 1 || r.y == nn || q.y < 0 || q.y > nn + 1 || q.y == nn || q.p < 0 || q.p > nn + 1 ||
            q.x < 0 || q.x > nn + 1 || q.x == nn)
          return bad(g, "lean record");
        const bool sink_h = x == nn + 1, sink_d = q.y == nn + 1;
        if (sink_h != sink_d || (sink_h && (r.y != nn + 1 || q.x != nn + 1))) return bad(g, "lean record pair");
        if (sink_h) continue;
        if (i != nu++) return bad(g, "record after a sink record");
        for (int node : {x, (int)r.y})
          if (node < nn && done_h[node]++) return bad(g, "node eliminated twice");
        for (int node : {(int)q.y, (int)q.x})
          if (node < nn && done_d[node]++) return bad(g, "node substituted twice");
      }
      if (H.hrec[rb + nn].x != nn + 1 || H.drec[rb + nn].y != nn + 1) return bad(g, "sink record");
      for (int64_t li = 0; li < nn; ++li)
        if (!done_h[li] || !done_d[li]) return bad(g, "node missing from the lean records");
      for (int sweep = 0; sweep < 4; ++sweep)
        for (int w = 0; w < W; ++w) {
          const int64_t z = sched_at(sc, sweep * (W + 1) + w + 1);
          if (sc[z - 1] > nu) return bad(g, "pass offset beyond the units");
        }
      const std::string e = emulate_lean(L, H, g, W);
      if (!e.empty()) return bad(g, e.c_str());
    }
    if (mode != 0) {
      if (n0 + nn > (int64_t)H.arec.size()) return bad(g, "assembly records");
      for (int64_t li = 0; li < nn; ++li) {
        const ARec& r = H.arec[n0 + li];
        if (r.h < -2 || r.h >= nn || r.l < -1 || r.l >= nn || r.p < -1 || r.p >= nn) return bad(g, "assembly record");
      }
    }
  }
  if (next_cell != L.ncells) return "groups do not cover the cells";
  for (size_t k = 0; k < H.large.size(); ++k) {
    const int g = H.large[k];
    if (g < 0 || g >= ng || H.glist[g] == 0 || H.large_mode[k] != list_mode(H.glist[g], H.layout))
      return "large launch entry";
    if (H.lds[H.glist[g]] > H.large_lds) return "large launch LDS size";
  }
  return "";
}

// Builds a plan on the host and uploads it.
int build_plan(State& S, int64_t budget, int threads, int layout, BatchPlan*& out) {
  PlanHost H;
  int rc = plan_host(S.L, budget, threads, layout, H);
  if (rc) return rc;
  BatchPlan* P = new BatchPlan();
  P->budget = budget;
  P->threads = threads;
  P->layout = layout;
  P->ngroups = (int64_t)H.cell0.size();
  for (int m = 0; m < kLists; ++m) {
    P->lists[m] = H.lists[m];
    P->lds[m] = H.lds[m];
  }
  P->large = H.large;
  P->large_mode = H.large_mode;
  P->large_lds = H.large_lds;
  BatchDev& b = P->b;
  int32_t* p = nullptr;
  int16_t* p16 = nullptr;
  HRec* ph = nullptr;
  DRec* pd = nullptr;
  ARec* pa = nullptr;
  int8_t* p8 = nullptr;
  if ((rc = plan_upload(S, *P, &p, H.cell0))) goto error; b.grp_cell0 = p;
  if ((rc = plan_upload(S, *P, &p, H.cell1))) goto error; b.grp_cell1 = p;
  if ((rc = plan_upload(S, *P, &p16, H.r_tab))) goto error; b.r_tab = p16;
  if ((rc = plan_upload(S, *P, &p, H.grp_rtab))) goto error; b.grp_rtab = p;
  if ((rc = plan_upload(S, *P, &p, H.grp_slen))) goto error; b.grp_slen = p;
  if ((rc = plan_upload(S, *P, &ph, H.hrec))) goto error; b.r_hrec = ph;
  if ((rc = plan_upload(S, *P, &pd, H.drec))) goto error; b.r_drec = pd;
  if ((rc = plan_upload(S, *P, &p, H.rbase))) goto error; b.grp_rbase = p;
  if ((rc = plan_upload(S, *P, &pa, H.arec))) goto error; b.r_arec = pa;
  for (int m = 0; m < kLists; ++m)
    if ((rc = plan_upload(S, *P, &P->dev_lists[m], P->lists[m]))) goto error;
  if ((rc = plan_upload(S, *P, &P->dev_large, P->large))) goto error;
  if ((rc = plan_upload(S, *P, &p8, P->large_mode))) goto error;
  P->dev_large_mode = p8;
  // events are bucketed per group
  S.cell_bucket.assign(S.L.ncells, 0);
  for (size_t g = 0; g < H.cell0.size(); ++g)
    for (int32_t c = H.cell0[g]; c < H.cell1[g]; ++c) S.cell_bucket[c] = (int32_t)g;
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
    HIP_TRY(launch(S, K_WINDOW_LARGE, S.core ? (const void*)k_window_large_core : (const void*)k_window_large,
                   dim3((unsigned)P.large.size()),
                   dim3(std::min(P.threads, kLargeThreads)), P.large_lds, S.aux, S.d, P.b,
                   (const int32_t*)P.dev_large, (const int8_t*)P.dev_large_mode, n));
    HIP_TRY(hipEventRecord(S.join, S.aux));
  }
  int rc = 0;
  if (S.profiling && !P.lists[0].empty() && (rc = timer_start(S))) return rc;
  if (!P.lists[0].empty()) {
    const int id = P.layout == 5 ? K_WINDOW_LEAN : P.layout == 4 ? K_WINDOW_LEAN_HBM : K_WINDOW;
    const void* fn = P.layout == 5   ? (S.core ? (const void*)k_window_lean_core : (const void*)k_window_lean)
                     : P.layout == 4 ? (S.core ? (const void*)k_window_lean_hbm_core : (const void*)k_window_lean_hbm)
                                     : (S.core ? (const void*)k_window_core : (const void*)k_window);
    HIP_TRY(launch(S, id, fn, dim3((unsigned)P.lists[0].size()), dim3(P.threads), P.lds[0], S.stream, S.d, P.b,
                   (const int32_t*)P.dev_lists[0], n));
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
namespace {

// Autotuned plans by model topology: this process's choices, and optionally a
// file named by MIND_PLAN_CACHE (lines "key budget threads layout").
struct PlanChoice { int64_t budget; int threads; int layout; };
std::mutex g_plan_mu;
std::map<uint64_t, PlanChoice> g_plans;

uint64_t mix(uint64_t h, const void* p, size_t n) {
  const uint8_t* b = static_cast<const uint8_t*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

// What the plan's cost depends on: the cells' trees, the mechanism rows per
// cell, the detectors and the window length.
uint64_t topology_key(const State& S) {
  const Layout& L = S.L;
  uint64_t h = 0xcbf29ce484222325ull;
  const int32_t version = 2;  // of the candidate list
  h = mix(h, &version, sizeof version);
  h = mix(h, L.cell_off.data(), L.cell_off.size() * sizeof(int64_t));
  h = mix(h, L.parent.data(), L.parent.size() * sizeof(int32_t));
  for (const MechGroup& G : L.groups) {
    h = mix(h, &G.mech, sizeof G.mech);
    h = mix(h, G.cell_rows.data(), G.cell_rows.size() * sizeof(int64_t));
  }
  h = mix(h, L.cell_det_off.data(), L.cell_det_off.size() * sizeof(int64_t));
  h = mix(h, &S.wmax, sizeof S.wmax);
  return h;
}

bool cached_plan(uint64_t key, PlanChoice* out) {
  std::lock_guard<std::mutex> lock(g_plan_mu);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    *out = it->second;
    return true;
  }
  const char* path = std::getenv("MIND_PLAN_CACHE");
  if (!path || !*path) return false;
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  unsigned long long k = 0;
  long long b = 0;
  int t = 0, l = 0;
  bool found = false;
  while (std::fscanf(f, "%llx %lld %d %d", &k, &b, &t, &l) == 4)
    if (k == key) {
      *out = PlanChoice{b, t, l};
      found = true;
    }
  std::fclose(f);
  if (found) g_plans[key] = *out;
  return found;
}

void remember_plan(uint64_t key, const PlanChoice& c) {
  std::lock_guard<std::mutex> lock(g_plan_mu);
  g_plans[key] = c;
  const char* path = std::getenv("MIND_PLAN_CACHE");
  if (!path || !*path) return;
  if (FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "%llx %lld %d %d\n", (unsigned long long)key, (long long)c.budget, c.threads, c.layout);
    std::fclose(f);
  }
}

}  // namespace

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
  HIP_TRY(hipFuncSetAttribute((const void*)k_window_lean_hbm, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kLdsResident));
  for (const void* fn : {(const void*)k_window_core, (const void*)k_window_large_core, (const void*)k_window_lean_core,
                         (const void*)k_window_lean_hbm_core})
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsResident));
  int64_t max_cell = 1;
  for (int32_t c = 0; c < S.L.ncells; ++c) max_cell = std::max<int64_t>(max_cell, S.L.cell_off[c + 1] - S.L.cell_off[c]);
  struct Cand { int64_t budget; int threads; int layout; };
  std::vector<Cand> cands;
  const uint64_t key = topology_key(S);
  PlanChoice known{};
  const bool log = std::getenv("MIND_AUTOTUNE_LOG") != nullptr;
  if (cells_per_group > 0) {
    cands.push_back({std::max<int64_t>(1, (int64_t)cells_per_group * max_cell), 256, 0});
  } else if (cached_plan(key, &known)) {
    cands.push_back({known.budget, known.threads, known.layout});
  } else {
    for (int threads : {256, 512}) {
      for (int64_t bud : {1024, 1700, 2750}) cands.push_back({bud, threads, 0});
      for (int64_t bud : {1800, 2400, 3000, 3600, 4800, 6000}) cands.push_back({bud, threads, 5});
      for (int64_t bud : {3000, 3600, 4500, 5400, 7200}) cands.push_back({bud, threads, 4});
    }
    for (int64_t bud : {4000, 5000}) cands.push_back({bud, 512, 4});
    for (int64_t bud : {4800, 6000}) cands.push_back({bud, 1024, 5});
    for (int64_t bud : {7200, 9000}) cands.push_back({bud, 1024, 4});
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
    (void)P;
    // whole windows as a run replays them (gather, cells, exchange captured in
    // one graph): two to warm up, then four timed
    const int32_t n = S.wmax;
    hipGraph_t g;
    hipGraphExec_t exec;
    HIP_TRY(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
    rc2 = launch_window(S, n);
    const hipError_t e = hipStreamEndCapture(S.stream, &g);
    if (rc2) return rc2;
    if (e != hipSuccess) return fail(MIND_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    HIP_TRY(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    for (int rep = 0; rep < 2 && !rc2; ++rep) {
      if (hipStreamSynchronize(S.stream) != hipSuccess) rc2 = fail(MIND_EHIP, "autotune synchronisation");
      const auto t0 = std::chrono::steady_clock::now();
      for (int w = 0; w < 2 + 2 * rep && !rc2; ++w)
        if (hipGraphLaunch(exec, S.stream) != hipSuccess) rc2 = fail(MIND_EHIP, "autotune graph launch");
      if (!rc2 && hipStreamSynchronize(S.stream) != hipSuccess) rc2 = fail(MIND_EHIP, "autotune synchronisation");
      secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    (void)hipGraphExecDestroy(exec);
    return rc2;
  };
  // every candidate once, then the three fastest twice more: the pick has the
  // lowest mean of its three timings
  std::vector<std::pair<double, size_t>> timed;
  for (size_t k = 0; k < cands.size(); ++k) {
    double secs = 0;
    if ((rc = trial(cands[k], secs))) return rc;
    timed.push_back({secs, k});
    if (log && tune)
      std::fprintf(stderr, "autotune: budget %lld threads %d layout %d: %.3f ms per window\n",
                   (long long)cands[k].budget, cands[k].threads, cands[k].layout, 1e3 * secs / 4);
  }
  if (tune) {
    std::sort(timed.begin(), timed.end());
    for (size_t q = 0; q < std::min<size_t>(4, timed.size()); ++q) {
      double total = timed[q].first;
      for (int rep = 0; rep < 2; ++rep) {
        double secs = 0;
        if ((rc = trial(cands[timed[q].second], secs))) return rc;
        total += secs;
      }
      if (total / 3 < best) {
        best = total / 3;
        pick = cands[timed[q].second];
      }
    }
    double secs = 0;
    if ((rc = trial(pick, secs))) return rc;
    if (cells_per_group <= 0) remember_plan(key, PlanChoice{pick.budget, pick.threads, pick.layout});
    if (log)
      std::fprintf(stderr, "autotune: picked budget %lld threads %d layout %d\n", (long long)pick.budget,
                   pick.threads, pick.layout);
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
  if (threads < 64 || threads > kMaxThreads || threads % 64)
    return fail(MIND_EARG, "threads must be 64, 128, ..., 1024");
  if (layout != 0 && layout != 4 && layout != 5)
    return fail(MIND_EARG, "layout must be 0 (all in LDS), 4 (lean) or 5 (lean, v in LDS)");
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

extern "C" int mind_batch_kernel(const mind_state* state, const char** name) {
  const mind::State& S = *reinterpret_cast<const mind::State*>(state);
  if (!S.batch) return mind::fail(MIND_EARG, "state has no batched plan");
  const int lay = S.batch->layout;
  if (S.core)
    *name = lay == 5 ? "k_window_lean_core" : lay == 4 ? "k_window_lean_hbm_core" : "k_window_core";
  else
    *name = lay == 5 ? "k_window_lean" : lay == 4 ? "k_window_lean_hbm" : "k_window";
  return 0;
}

extern "C" int mind_set_batch_plan(mind_state* state, int64_t budget, int32_t threads, int32_t layout) {
  mind::State& S = *reinterpret_cast<mind::State*>(state);
  HIP_TRY(hipSetDevice(S.device));
  return mind::batched_replan(S, budget, threads, layout);
}

extern "C" int mind_batch_work(const mind_state* state, int64_t* counts) {
  return mind::batched_dominant_work(*reinterpret_cast<const mind::State*>(state), counts);
}

// Host-only: builds the batched plan of a model as the autotuner would for
// (budget, threads, layout) and checks every index its kernels would derive from
// it. No device is used.
extern "C" int mind_check_batch_plan(const mind_model_desc* model, double dt, int64_t budget, int32_t threads,
                                     int32_t layout, int64_t* groups) {
  if (!model || budget < 1 || threads < 64 || threads > mind::kMaxThreads || threads % 64 ||
      (layout != 0 && layout != 4 && layout != 5))
    return mind::fail(MIND_EARG, "invalid plan");
  mind::Layout L;
  const std::string err = mind::pack(*model, dt, 1, L);
  if (!err.empty()) return mind::fail(MIND_EARG, err);
  mind::PlanHost H;
  int rc = mind::plan_host(L, budget, threads, layout, H);
  if (rc) return rc;
  const std::string bad = mind::check_plan(L, H);
  if (!bad.empty()) return mind::fail(MIND_EUNSUPPORTED, "batched plan check: " + bad);
  if (groups) *groups = (int64_t)H.cell0.size();
  return 0;
}

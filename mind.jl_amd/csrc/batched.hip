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

namespace mind {

// Lean layouts (modes 4 and 5): the solve works on units, a unit being a node
// whose parent has no other child paired with that parent (the bottom and the
// top of the unit), or a node alone. The host pairs nodes top-down: a node that
// is not itself a bottom and has exactly one child takes that child as its
// bottom. A unit's depth is its top's depth in the tree of units, so one lane
// eliminates a unit's two nodes in one pass (bottom, then top pulling the
// bottom from registers) and substitutes them in one pass (top, then bottom):
// on unbranched stretches of cable half the passes, each one division longer.
// Each node keeps its own operation sequence, so the results do not change.
// The solve's records are read from HBM kAhead passes ahead. A group of nn
// nodes has LDS slots nn (the dummy: pivot 1, rhs 1) and nn + 1 (the sink), and
// nn + 1 records per order: one per unit by position in the order, then sink
// records up to position nn, the sink record of the padding lanes.
// Elimination record (by position in the elimination order): x the unit's
// bottom (its only node if alone), h and l x's higher and lower child (the dummy
// if missing) with their axial conductances gh, gl (-0 for the dummy); y the
// unit's top (the sink if alone) and gy x's own axial conductance (-0 if alone).
// x is stored as ~x (negative) when it has more than two children, which it
// then pulls from the CSR lists. Back-substitution record (by position in the
// order of increasing depth): y the unit's top, p its parent (the dummy at a
// root), gy y's axial conductance (-0 at a root), x the bottom (the sink if
// alone) and gx its axial conductance (-0 if alone). With the dummy's values
// every term of a missing neighbour is an exact identity (x - (+0) = x,
// x + (-0) = x), so no lane branches.
struct HRec {
  int16_t x, h, l, y;
  double gh, gl, gy;
};
struct DRec {
  int16_t y, p, x, pad;
  double gy, gx;
};

// Cable assembly record of a node (modes 3-5, whose links and conductances are
// not in LDS): higher and lower child, parent (-1 if none, higher -2 with more
// than two children), the axial conductances of the node and of those two
// children, and the node's C/dt and axial diagonal. Indexed by global node.
struct ARec {
  int16_t h, l, p, pad;
  double g, gh, gl, cdt, adiag;
};

struct BatchDev {
  const int32_t* grp_cell0;
  const int32_t* grp_cell1;
  const ARec* r_arec;  // modes 3-5
  const HRec* r_hrec;  // modes 4 and 5, group g's from grp_rbase[g]
  const DRec* r_drec;
  const int32_t* grp_rbase;
  // per group one int16 block (local node ids, 8-byte aligned): per node its
  // links {higher child, lower child, parent, 0}, then the nodes by height and
  // by depth (each the concatenation of the waves' orders), then the schedule
  // block of sched_at()
  const int16_t* r_tab;
  const int32_t* grp_rtab;
  const int32_t* grp_slen;  // schedule block length
};

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

constexpr int kMaxThreads = 1024;
constexpr int kLargeThreads = 512;  // block size cap of the launch of cells above the budget
constexpr int kMaxGroups = MIND_MECH_COUNT;  // mechanism groups cached in LDS: one per builtin at most
constexpr int kMaxDetLds = 256;  // detectors of a workgroup cached in LDS
// dynamic LDS of a workgroup, leaving room for its static tables
constexpr size_t kLdsResident = 152 * 1024;

// Schedule block of a group with W waves: entries 0..W give where wave w's
// elimination pass offsets start in the block (entry W: their end), entries
// W+1..2W+1 the same for its back-substitution passes, each entry two int16 words
// (low, high 16 bits); then per wave its pass offsets, positions in oh (od) where
// each pass starts, plus the end. A pass is at most 64 entries of one level (a
// wider level takes several passes), and a wave's passes run in level order.
// Lean groups have two more sections, 2W+2..3W+2 and 3W+3..4W+3: the elimination
// and back-substitution passes of stage 2 (the top of a cell split across waves,
// entries after stage 1's in oh and od); the first two sections then hold stage 1.
__host__ __device__ inline int sched_at(const int16_t* s, int k) {
  return (int)(uint16_t)s[2 * k] | ((int)s[2 * k + 1] << 16);
}
// int16 entries of a group's table block.
__host__ __device__ inline int64_t rtab_len(int64_t nn, int64_t slen) { return 6 * nn + slen; }
// Dynamic LDS bytes of a workgroup in mode M: mode 0 keeps 4 doubles per node and
// the table block, mode 4 the (pivot, rhs) pairs (plus dummy and sink) and the
// schedule block, mode 5 also v, mode 3 nothing.
inline size_t window_bytes(int m, int64_t nn, int64_t slen) {
  if (m == 3) return 0;
  if (m == 4) return (size_t)(16 * (nn + 2) + (2 * slen + 7) / 8 * 8);
  if (m == 5) return (size_t)(16 * (nn + 2) + 8 * nn + (2 * slen + 7) / 8 * 8);
  return (size_t)(32 * nn + (2 * rtab_len(nn, slen) + 7) / 8 * 8);
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
// conductances, links and schedule), 5 (pivot, rhs) pairs, v and the schedule
// in LDS, 4 the pairs and the schedule in LDS, 3 everything in HBM (rhs and diag
// in the global scratch columns, the links and schedule read from the plan).
// Modes 4 and 5 solve with the lean records.
// Orders one wavefront's solve accesses between levels. Its lanes run in
// lockstep, so a fence is all a level needs from the one before: at wavefront
// scope for rhs and diag in LDS (executed in order, no wait), at workgroup scope
// for them in global memory (mode 3: waits for the stores).
template <int M>
__device__ __forceinline__ void wave_sync() {
  if (M == 3) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// A node's links: higher child (-1 if fewer than two children, -2 if more than
// two), lower child (-1 if none), parent (-1 at a root).
struct Links {
  int h, l, p;
};
__device__ __forceinline__ Links links_of(const uint2 w) {
  return Links{(int16_t)(w.x & 0xffffu), (int16_t)(w.x >> 16), (int16_t)(w.y & 0xffffu)};
}

// Elimination of node li: pulls its higher child h, then its lower child l (gh,
// gl their axial conductances), or, with more than two children (h == -2), all
// of them from the CSR lists in decreasing index. Without branches for one or
// two children: a missing child enters as conductance -0, pivot 1 and rhs 1,
// whose terms are exact identities (x - (+0) = x, x + (-0) = x), so both reads
// and both divisions issue at once. Stores only if `store`; returns true if a
// stored pivot is zero or not finite.
// rhs and pivot of a node: interleaved (pivot, rhs) pairs in LDS, one 16-byte
// access per node, or separate global columns (mode 3). S is the stride.
template <int S>
struct Col {
  double* p;
  __device__ __forceinline__ double& operator[](int i) const { return p[(int64_t)i * S]; }
};

// The reads of a one- or two-child elimination, issued before the next level's
// prefetch so that the prefetch does not wait for them.
struct ElimIn {
  double dd, bb, dh, rh, dl, rl, gh, gl;
  bool hv, lv;
};
template <typename C>
__device__ __forceinline__ ElimIn elim_reads(int li, int h, int l, const double* ax, C rhs, C diag) {
  ElimIn e;
  e.hv = h >= 0;
  e.lv = l >= 0;
  const int hs = e.hv ? h : li, ls = e.lv ? l : li;
  e.dd = diag[li];
  e.bb = rhs[li];
  e.dh = diag[hs];
  e.rh = rhs[hs];
  e.dl = diag[ls];
  e.rl = rhs[ls];
  e.gh = ax[hs];
  e.gl = ax[ls];
  return e;
}
// Nothing is scheduled across this point: every read above it is issued before
// the first wait for one of them.
__device__ __forceinline__ void reads_issued() { __builtin_amdgcn_sched_barrier(0); }
template <typename C>
__device__ __forceinline__ bool elim_finish(const ElimIn& e, int li, C rhs, C diag, bool store) {
  const double dh = e.hv ? e.dh : 1.0, rh = e.hv ? e.rh : 1.0, dl = e.lv ? e.dl : 1.0, rl = e.lv ? e.rl : 1.0;
  const double gh = e.hv ? e.gh : -0.0, gl = e.lv ? e.gl : -0.0;
  const double fh = gh / dh, fl = gl / dl;
  const double dd = (e.dd - fh * gh) - fl * gl;
  const double bb = (e.bb + fh * rh) + fl * rl;
  if (!store) return false;
  diag[li] = dd;
  rhs[li] = bb;
  return (dd == 0.0) || !mind_isfinite(dd);
}

template <typename C>
__device__ __forceinline__ bool eliminate(const Dev& d, int64_t n0, int li, int h, int l, double gh, double gl,
                                          const double* ax, C rhs, C diag, bool store) {
  double dd = diag[li], bb = rhs[li];
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
    const bool hv = h >= 0, lv = l >= 0;
    const int hs = hv ? h : li, ls = lv ? l : li;
    const double dh0 = diag[hs], rh0 = rhs[hs], dl0 = diag[ls], rl0 = rhs[ls];
    const double dh = hv ? dh0 : 1.0, rh = hv ? rh0 : 1.0, dl = lv ? dl0 : 1.0, rl = lv ? rl0 : 1.0;
    const double ghs = hv ? gh : -0.0, gls = lv ? gl : -0.0;
    const double fh = ghs / dh, fl = gls / dl;
    dd = (dd - fh * ghs) - fl * gls;
    bb = (bb + fh * rh) + fl * rl;
  }
  if (!store) return false;
  diag[li] = dd;
  rhs[li] = bb;
  return (dd == 0.0) || !mind_isfinite(dd);
}

// A wave's pass offsets (nl + 1 of them at lv) held 64 at a time in one
// register, lane j holding offset base + j: reading one is a cross-lane move, not
// a memory access. at() takes wave-uniform indices j <= nl that grow by at most
// one per call; every call also reads the following 64 offsets, so that moving
// on to them needs no branch and no wait.
struct LevelOffsets {
  const int16_t* lv;
  int nl, base, reg, nxt;
  __device__ LevelOffsets(const int16_t* lv_, int nl_) : lv(lv_), nl(nl_), base(0) {
    const int lane = threadIdx.x & 63;
    reg = lv[min(lane, nl)];
    nxt = lv[min(64 + lane, nl)];
  }
  __device__ __forceinline__ int at(int j) {
    const bool cross = j >= base + 64;
    reg = cross ? nxt : reg;
    base = cross ? base + 64 : base;
    nxt = lv[min(base + 64 + (int)(threadIdx.x & 63), nl)];
    return __builtin_amdgcn_readlane(reg, j - base);
  }
};

// Tree solve of one wavefront's cells (modes 0 and 3): elimination pass by pass
// in its order, back substitution pass by pass in order of depth, a pass being
// at most 64 nodes of one level, one per lane. A lane's node two passes ahead
// and the links of its node one pass ahead are read at the top of a pass, before
// the pass's own reads, so that a pass waits for one round of reads: its
// children's (parent's) values. Lanes past a pass's end compute on a node they
// do not store.
template <int M, typename C>
__device__ __forceinline__ void solve_waves(const Dev& d, const BatchDev& b, int64_t n0, int nn,
                                            const int16_t* sched, const uint2* hn, const int16_t* oh,
                                            const int16_t* od, const double* ax, C rhs, C diag, double* v) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, W = blockDim.x >> 6;
  const int h0 = sched_at(sched, w), nlh = sched_at(sched, w + 1) - h0 - 1;
  const int d0 = sched_at(sched, W + 1 + w), nld = sched_at(sched, W + 2 + w) - d0 - 1;
  const int last = nn - 1;
  bool bad = false;
  if (nlh > 0) {
    LevelOffsets lo(sched + h0, nlh);
    int a = lo.at(0), z = lo.at(1);
    {
      int li = oh[min(a + lane, last)];
      Links k = links_of(hn[li]);
      const int o2 = lo.at(min(2, nlh));
      int z1 = nlh >= 2 ? o2 : z;
      int li1 = oh[min(z + lane, last)];
      for (int L = 0; L < nlh; ++L) {
        const int o3 = lo.at(min(L + 3, nlh));
        const int z2 = L + 2 < nlh ? o3 : z1;
        const int li2 = oh[min(z1 + lane, last)];
        const Links k1 = links_of(hn[li1]);
        if (k.h == -2) {
          bad |= eliminate(d, n0, li, k.h, k.l, 0.0, 0.0, ax, rhs, diag, a + lane < z);
        } else {
          const ElimIn e = elim_reads(li, k.h, k.l, ax, rhs, diag);
          reads_issued();
          bad |= elim_finish(e, li, rhs, diag, a + lane < z);
        }
        wave_sync<M>();
        li = li1;
        li1 = li2;
        k = k1;
        a = z;
        z = z1;
        z1 = z2;
      }
    }
  }
  if (bad) raise_error(d.error, MIND_ESINGULAR);
  if (nld > 0) {
    LevelOffsets lo(sched + d0, nld);
    int a = lo.at(0), z = lo.at(1);
    {
      int li = od[min(a + lane, last)];
      int p = links_of(hn[li]).p;
      const int o2 = lo.at(min(2, nld));
      int z1 = nld >= 2 ? o2 : z;
      int li1 = od[min(z + lane, last)];
      for (int L = 0; L < nld; ++L) {
        const int o3 = lo.at(min(L + 3, nld));
        const int z2 = L + 2 < nld ? o3 : z1;
        const int li2 = od[min(z1 + lane, last)];
        const int p1 = links_of(hn[li1]).p;
        const bool pv = p >= 0;
        const double g = ax[li], rp0 = rhs[pv ? p : li], x0 = rhs[li], dg = diag[li];
        const double vi = v[li];
        reads_issued();
        if (a + lane < z) {
          const double x = (x0 + (pv ? g : -0.0) * (pv ? rp0 : 1.0)) / dg;
          rhs[li] = x;
          v[li] = vi + d.vscale * x;
        }
        wave_sync<M>();
        li = li1;
        li1 = li2;
        p = p1;
        a = z;
        z = z1;
        z1 = z2;
      }
    }
  }
}

// Elimination of node li with more than two children, from its own pivot and
// rhs (dd, bb): children from the CSR lists in decreasing index. Returns true if
// the pivot is zero or not finite.
template <typename C>
__device__ __forceinline__ bool elim_csr(const Dev& d, int64_t n0, int li, double& dd, double& bb, C rhs, C diag) {
  const int64_t gi = n0 + li;
  for (int64_t cq = d.child_off[gi + 1] - 1; cq >= d.child_off[gi]; --cq) {
    const int lc = d.child[cq] - (int)n0;
    const double g = d.axial[n0 + lc];
    const double f = g / diag[lc];
    dd -= f * g;
    bb += f * rhs[lc];
  }
  diag[li] = dd;
  rhs[li] = bb;
  return (dd == 0.0) || !mind_isfinite(dd);
}

// True if a pivot is zero or not finite (no branches).
__device__ __forceinline__ bool bad_pivot(double dd) {
  return !((dd != 0.0) & (__builtin_fabs(dd) < __builtin_huge_val()));
}

// Tree solve of one wave's cells in the lean layouts, by units (rhs and pivots
// in LDS pairs, records from HRec / DRec). The loop runs passes in threes with
// three records in flight. An elimination pass reads the pairs of its bottom x,
// x's children and its top y (four LDS reads), eliminates x (two divisions
// issued together), then y pulling x (a third), and stores both pairs; it then
// loads the record of the pass three ahead into the registers it has just
// freed, so that no register moves and no wait stand between a load and its use
// three passes later. Passes past the wave's last only move the sink. A
// back-substitution pass solves the top from its parent, then the bottom from
// the top. The pass offsets are read one pass ahead (a broadcast LDS read
// alongside the pass's own). Back substitution leaves the node's change in rhs;
// v += change follows for all nodes after the solve.
// Passes of record loads in flight in the lean solve.
constexpr int kAhead = 3;

// Elimination passes of one schedule section (header entries from hdr0) for this
// wave, from the records hr. Returns true if a stored pivot is zero or not finite.
template <typename C>
__device__ __forceinline__ bool elim_section(const Dev& d, const HRec* hr, int64_t n0, int nn, const int16_t* sched,
                                             int hdr0, C rhs, C diag) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int h0 = sched_at(sched, hdr0 + w), nph = sched_at(sched, hdr0 + w + 1) - h0 - 1;
  bool bad = false;
  if (nph <= 0) return false;
  const int16_t* lv = sched + h0;
  auto slot = [&](int a, int z) { return a + lane < z ? a + lane : nn; };
  HRec x[kAhead];
  int next = lv[0];  // start of the next pass to load
#pragma unroll
  for (int k = 0; k < kAhead; ++k) {
    const int e = lv[min(k + 1, nph)];
    x[k] = hr[slot(next, e)];
    next = e;
  }
  auto pass = [&](HRec& r, int q) {
    const int e = lv[min(q + kAhead + 1, nph)];  // end of the pass kAhead ahead
    const int xi = r.x < 0 ? ~r.x : r.x;
    const double od = diag[xi], orr = rhs[xi];
    const double dh = diag[r.h], rh = rhs[r.h], dl = diag[r.l], rl = rhs[r.l];
    const double oy = diag[r.y], ry = rhs[r.y];
    reads_issued();
    const double fh = r.gh / dh, fl = r.gl / dl;
    double dd = (od - fh * r.gh) - fl * r.gl;
    double bb = (orr + fh * rh) + fl * rl;
    if (r.x < 0) {  // more than two children: the CSR lists, from x's own pivot and rhs
      dd = od;
      bb = orr;
      bad |= elim_csr(d, n0, xi, dd, bb, rhs, diag);
    }
    const double fy = r.gy / dd;
    const double dy = oy - fy * r.gy;
    const double by = ry + fy * bb;
    diag[xi] = dd;
    rhs[xi] = bb;
    diag[r.y] = dy;
    rhs[r.y] = by;
    bad |= ((xi < nn) & bad_pivot(dd)) | ((r.y < nn) & bad_pivot(dy));
    wave_sync<4>();
    r = hr[slot(next, e)];
    next = e;
  };
  for (int q = 0; q < nph; q += kAhead) {
#pragma unroll
    for (int k = 0; k < kAhead; ++k) pass(x[k], q + k);
  }
  return bad;
}

// Back-substitution passes of one schedule section for this wave.
template <typename C>
__device__ __forceinline__ void subst_section(const DRec* dr, int nn, const int16_t* sched, int hdr0, C rhs,
                                              C diag) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int d0 = sched_at(sched, hdr0 + w), npd = sched_at(sched, hdr0 + w + 1) - d0 - 1;
  if (npd <= 0) return;
  const int16_t* lv = sched + d0;
  auto slot = [&](int a, int z) { return a + lane < z ? a + lane : nn; };
  DRec x[kAhead];
  int next = lv[0];
#pragma unroll
  for (int k = 0; k < kAhead; ++k) {
    const int e = lv[min(k + 1, npd)];
    x[k] = dr[slot(next, e)];
    next = e;
  }
  auto pass = [&](DRec& r, int q) {
    const int e = lv[min(q + kAhead + 1, npd)];
    const double oy = diag[r.y], ry = rhs[r.y], rp = rhs[r.p];
    const double ox = diag[r.x], rx = rhs[r.x];
    reads_issued();
    const double xy = (ry + r.gy * rp) / oy;
    rhs[r.y] = xy;
    rhs[r.x] = (rx + r.gx * xy) / ox;
    wave_sync<4>();
    r = dr[slot(next, e)];
    next = e;
  };
  for (int q = 0; q < npd; q += kAhead) {
#pragma unroll
    for (int k = 0; k < kAhead; ++k) pass(x[k], q + k);
  }
}

// Tree solve of a group in the lean layouts: stage 1 elimination; if a cell is
// split across waves (stage-2 passes exist), a workgroup barrier, stage 2
// elimination and substitution (the split cell's top), another barrier; then
// stage 1 substitution. Every thread of the workgroup calls this (the barriers),
// and whether a split exists is the same for all of them.
template <typename C>
__device__ __forceinline__ void solve_lean(const Dev& d, const BatchDev& b, int grp, int64_t n0, int nn,
                                           const int16_t* sched, C rhs, C diag) {
  const int W = blockDim.x >> 6;
  const int64_t rb = b.grp_rbase[grp];
  const HRec* hr = b.r_hrec + rb;
  const DRec* dr = b.r_drec + rb;
  const int e2 = 2 * (W + 1), s2 = 3 * (W + 1);
  const bool split = sched_at(sched, e2 + W) - sched_at(sched, e2) > W;
  bool bad = elim_section(d, hr, n0, nn, sched, 0, rhs, diag);
  if (split) {
    __syncthreads();
    bad |= elim_section(d, hr, n0, nn, sched, e2, rhs, diag);
    subst_section(dr, nn, sched, s2, rhs, diag);
    __syncthreads();
  }
  if (bad) raise_error(d.error, MIND_ESINGULAR);
  subst_section(dr, nn, sched, W + 1, rhs, diag);
}

// nrn_cur of a density mechanism's rows [r0, r1) of one workgroup: U rows per
// thread per round, the column reads of all U issued before any row's evaluation
// and stores (each row is its own node, so the rows of a round are independent).
// Same operations per row as the generic loop.
template <int MECH, int U, typename C>
__device__ __forceinline__ void density_currents(const Dev& d, const Cols& col, const int32_t* node_of, int64_t r0,
                                                 int64_t r1, int64_t n0, double tc, const double* v, C rhs, C diag,
                                                 int tid, int T, bool store) {
  using Ops = CurOps<MECH>;
  for (int64_t r = r0 + tid; r < r1; r += (int64_t)U * T) {
    int64_t rr[U];
    int32_t nd[U];
    typename Ops::In in[U];
    double area[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = r + (int64_t)u * T;
      rr[u] = q < r1 ? q : r;
      nd[u] = node_of[rr[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      in[u] = Ops::load(col, rr[u]);
      area[u] = d.area[nd[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (r + (int64_t)u * T < r1) {
        const int li = nd[u] - (int)n0;
        double value, shifted;
        Ops::eval(col, rr[u], in[u], v[li], tc, d.dt, value, shifted, store);
        const double scale = area[u] * 0.01;
        double acc_r = rhs[li], acc_d = diag[li];
        acc_r -= scale * value;
        acc_d += scale * (shifted - value) / 0.001;
        rhs[li] = acc_r;
        diag[li] = acc_d;
      }
    }
  }
}

// Mechanisms whose current-phase stores are outputs only: no other phase and
// no NET_RECEIVE reads them.
__device__ __forceinline__ bool mech_outputs_only(int32_t mech) {
  return mech == MIND_MECH_HH || mech == MIND_MECH_PAS || mech == MIND_MECH_EXPSYN || mech == MIND_MECH_EXP2SYN ||
         mech == MIND_MECH_ICLAMP || mech == MIND_MECH_ALPHASYNAPSE;
}

// Mechanisms with a state update (nrn_state).
__device__ __forceinline__ bool stateful(int32_t mech) {
  return mech_has_state(mech);
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
  const int slen = b.grp_slen[grp];
  const int ldsd = M == 0 ? 4 : M == 4 ? 2 : 0;  // doubles per node in LDS
  constexpr int RS = M == 3 ? 1 : 2;  // (pivot, rhs) pairs in LDS
  const Col<RS> rhs{M == 3 ? d.rhs + n0 : lds + 1};
  const Col<RS> diag{M == 3 ? d.diag + n0 : lds};
  constexpr bool lean = M == 4 || M == 5;
  double* v = M == 0 ? lds + 2 * nn : M == 5 ? lds + 2 * (nn + 2) : d.v + n0;
  double* axl = lds + 3 * nn;  // LDS copy (M = 0)
  const double* ax = M == 0 ? axl : d.axial + n0;
  const int tlen = (int)rtab_len(nn, slen);
  const int16_t* rt = b.r_tab + b.grp_rtab[grp];
  // LDS copy of the table (M = 0) or of the schedule block (M = 4, after the
  // pairs of the nodes, the dummy and the sink)
  int16_t* tabl = reinterpret_cast<int16_t*>(lds + (M == 4 ? 2 * (nn + 2) : M == 5 ? 2 * (nn + 2) + nn : ldsd * nn));
  const int16_t* tab = M == 0 ? tabl : rt;
  const uint2* hn = reinterpret_cast<const uint2*>(tab);  // links of each node
  const int16_t* oh = tab + 4 * nn;                        // nodes by height
  const int16_t* od = oh + nn;                             // nodes by depth
  const int16_t* sched = lean ? tabl : od + nn;            // schedule block (LDS copy in modes 4, 5)
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
  if (M == 5)
    for (int i = tid; i < nn; i += T) v[i] = d.v[n0 + i];
  if (lean) {
    for (int i = tid; i < slen; i += T) tabl[i] = rt[6 * nn + i];
    if (tid == 0) {
      diag[nn] = 1.0;  // the dummy
      rhs[nn] = 1.0;
      diag[nn + 1] = 1.0;  // the sink
      rhs[nn + 1] = 0.0;
    }
  }
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
    // nrn_cur at t + dt/2, one mechanism group after another. The assigned
    // outputs of a current phase that no other phase reads (i, g of hh, pas,
    // ExpSyn, Exp2Syn, IClamp, AlphaSynapse) are stored on the window's last
    // step, which leaves memory as it would be after every step, or on every
    // step when the run records values.
    const double tc = t + 0.5 * dt;
    const bool store_outputs = s == n - 1 || d.nrec > 0;
    for (int gi = 0; gi < ngr; ++gi) {
      const GroupDev& G = sg[gi];
      if (G.mech == MIND_MECH_NETSTIM) continue;
      const int64_t r0 = sr0[gi], r1 = sr1[gi];
      if (r0 == r1) continue;
      const Cols col = cols(G);
      if (G.mech == MIND_MECH_PAS && !G.point) {
        density_currents<MIND_MECH_PAS, 4>(d, col, G.node, r0, r1, n0, tc, v, rhs, diag, tid, T, store_outputs);
        __syncthreads();
        continue;
      }
      if (G.mech == MIND_MECH_HH && !G.point) {
        density_currents<MIND_MECH_HH, 1>(d, col, G.node, r0, r1, n0, tc, v, rhs, diag, tid, T, store_outputs);
        __syncthreads();
        continue;
      }
      for (int64_t r = r0 + tid; r < r1; r += T) {
        if (G.point && !G.run_start[r]) continue;
        const int32_t node = G.node[r];
        const int li = node - (int)n0;
        double acc_r = rhs[li], acc_d = diag[li];
        int64_t q = r;
        do {  // a run of rows on one node, in row order
          double value, shifted;
          phase_current(G.mech, col, q, v[li], tc, dt, value, shifted,
                        store_outputs || !mech_outputs_only(G.mech));
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
    auto assemble = [&](int li, int p, int l, int h, double g, double gl, double gh, double cdt, double adg) {
      const int64_t gi = n0 + li;
      const double dg = cdt + diag[li] + adg;
      double bb = rhs[li];
      const double vi = v[li];
      if (p >= 0) bb += g * (v[p] - vi);
      if (h == -2) {
        for (int64_t q = d.child_off[gi]; q < d.child_off[gi + 1]; ++q) {
          const int c = d.child[q] - (int)n0;
          bb -= d.axial[n0 + c] * (vi - v[c]);
        }
      } else {
        if (l >= 0) bb -= gl * (vi - v[l]);
        if (h >= 0) bb -= gh * (vi - v[h]);
      }
      diag[li] = dg;
      rhs[li] = bb;
    };
    if (M == 0) {
      for (int li = tid; li < nn; li += T) {
        const Links k = links_of(hn[li]);
        assemble(li, k.p, k.l, k.h, ax[li], k.l >= 0 ? ax[k.l] : 0.0, k.h >= 0 ? ax[k.h] : 0.0, d.cdt[n0 + li],
                 d.adiag[n0 + li]);
      }
    } else {  // four nodes per thread per round, their records read first
      constexpr int U = 4;
      const ARec* ar = b.r_arec + n0;
      for (int l0 = tid; l0 < nn; l0 += U * T) {
        ARec r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) r[u] = ar[l0 + u * T < nn ? l0 + u * T : l0];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (l0 + u * T < nn)
            assemble(l0 + u * T, r[u].p, r[u].l, r[u].h, r[u].g, r[u].gl, r[u].gh, r[u].cdt, r[u].adiag);
      }
    }
    __syncthreads();
    if (lean) {
      solve_lean(d, b, grp, n0, nn, sched, rhs, diag);
      __syncthreads();
      for (int i = tid; i < nn; i += T) v[i] += d.vscale * rhs[i];
    } else {
      solve_waves<M>(d, b, n0, nn, sched, hn, oh, od, ax, rhs, diag, v);
    }
    __syncthreads();
    // nrn_state at t + dt: the rows of all stateful groups as one index space
    // (rows are independent), so that the groups' rows run side by side
    {
      int64_t total = 0;
      for (int gi = 0; gi < ngr; ++gi)
        if (stateful(sg[gi].mech)) total += sr1[gi] - sr0[gi];
      for (int64_t f = tid; f < total; f += T) {
        int gi = 0;
        int64_t r = f;
        for (;; ++gi) {
          if (!stateful(sg[gi].mech)) continue;
          const int64_t cnt = sr1[gi] - sr0[gi];
          if (r < cnt) break;
          r -= cnt;
        }
        const GroupDev& G = sg[gi];
        r += sr0[gi];
        phase_advance(G.mech, cols(G), r, v[G.node[r] - n0], t + dt, dt, d.celsius);
      }
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
    if (d.snap && (s + 1) % d.snap_every == 0) {
      const int64_t m = (s + 1) / d.snap_every - 1;
      for (int64_t q = det0 + tid; q < det1; q += T) {
        const int32_t kk = d.cell_det[q];
        const int32_t si = d.det_snap[kk];
        if (si >= 0) d.snap[m * d.n_snap + si] = v[d.det_node[kk] - n0];
      }
    }
    __syncthreads();
    if (tid == 0) seg = deliver_group(d, seg, seg_end, 2 * k + 1);  // end-of-step delivery
  }
  if (M == 0 || M == 5) {
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

// Groups within the node budget in the lean layout, v in LDS.
__global__ void __launch_bounds__(kMaxThreads) k_window_lean(Dev d, BatchDev b, const int32_t* groups, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<5>(d, b, groups[blockIdx.x], n, sh, lds);
}

// Groups within the node budget in the lean layout, v in HBM.
__global__ void __launch_bounds__(kMaxThreads) k_window_lean_hbm(Dev d, BatchDev b, const int32_t* groups,
                                                                 int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<4>(d, b, groups[blockIdx.x], n, sh, lds);
}

// Cells above the node budget, one per workgroup, each with the layout that
// fits (modes[k] for groups[k]): one launch, so that they run side by side.
__global__ void __launch_bounds__(kLargeThreads) k_window_large(Dev d, BatchDev b, const int32_t* groups,
                                                              const int8_t* modes, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  const int grp = groups[blockIdx.x];
  switch (modes[blockIdx.x]) {
    case 0: window_body<0>(d, b, grp, n, sh, lds); break;
    case 4: window_body<4>(d, b, grp, n, sh, lds); break;
    case 5: window_body<5>(d, b, grp, n, sh, lds); break;
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
    hipLaunchKernelGGL(k_window_large, dim3((unsigned)P.large.size()), dim3(std::min(P.threads, kLargeThreads)),
                       P.large_lds, S.aux, S.d,
                       P.b, P.dev_large, P.dev_large_mode, n);
    HIP_TRY(hipEventRecord(S.join, S.aux));
  }
  int rc = 0;
  if (S.profiling && !P.lists[0].empty() && (rc = timer_start(S))) return rc;
  if (!P.lists[0].empty()) {
    if (P.layout == 5)
      hipLaunchKernelGGL(k_window_lean, dim3((unsigned)P.lists[0].size()), dim3(P.threads), P.lds[0], S.stream, S.d,
                         P.b, P.dev_lists[0], n);
    else if (P.layout == 4)
      hipLaunchKernelGGL(k_window_lean_hbm, dim3((unsigned)P.lists[0].size()), dim3(P.threads), P.lds[0], S.stream,
                         S.d, P.b, P.dev_lists[0], n);
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

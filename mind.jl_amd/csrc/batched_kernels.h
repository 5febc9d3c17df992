// Batched window kernels: cells per workgroup, node state staged in LDS, a
// window of steps per launch. Included by batched.hip for the builtin
// mechanisms, and compiled again with hipRTC, with the phases of
// runtime-compiled mechanisms added, for models that use them.
#pragma once

#include "device.h"

// Mechanism groups a workgroup caches in LDS: one per builtin mechanism, plus the
// runtime-compiled mechanisms of a module built with hipRTC.
#ifndef MIND_MAX_GROUPS
#define MIND_MAX_GROUPS MIND_MECH_COUNT
#endif

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


constexpr int kMaxThreads = 1024;
constexpr int kLargeThreads = 512;  // block size cap of the launch of cells above the budget
constexpr int kMaxGroups = MIND_MAX_GROUPS;
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

// Delivery of the window-list events due at delivery call `call` (thread 0), for
// any receiver.
__device__ inline int64_t deliver_group(const Dev& d, int64_t seg, int64_t seg_end, int64_t call) {
  while (seg < seg_end && d.win[seg].call == call) {
    const int32_t e = d.win[seg].edge;
    const int32_t target = d.edge_target[e];
    receive_event(d, d.g[d.point_group[target]], d.point_row[target], d.win[seg].t, d.edge_weight[e]);
    ++seg;
  }
  return seg;
}

// Delivery of the window-list events due at delivery call `call` by the whole
// workgroup, for the core receivers (the builtin synapses without ion exchanges,
// ExpSyn and Exp2Syn; IClamp has no NET_RECEIVE), whose NET_RECEIVE touches its
// own row only, so only their phases are compiled in. Every thread finds the call's end by binary search (the list
// is sorted by call, then target), then each thread delivers whole target runs,
// a run's events in list order, so every row sees the order of the serial
// delivery. Returns the call's end, the same in every thread.
__device__ inline int64_t deliver_core_runs(const Dev& d, int64_t seg, int64_t seg_end, int64_t call, int tid,
                                            int T) {
  int64_t lo = seg, hi = seg_end;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (d.win[mid].call <= call) lo = mid + 1;
    else hi = mid;
  }
  for (int64_t i = seg + tid; i < lo; i += T) {
    const int32_t target = d.edge_target[d.win[i].edge];
    if (i > seg && d.edge_target[d.win[i - 1].edge] == target) continue;  // inside a run
    const GroupDev& G = d.g[d.point_group[target]];
    const int64_t row = d.point_row[target];
    for (int64_t j = i; j < lo; ++j) {
      const int32_t e = d.win[j].edge;
      if (d.edge_target[e] != target) break;
      if (G.mech == MIND_MECH_EXP2SYN) phase_receive(MIND_MECH_EXP2SYN, cols(G), row, d.win[j].t, d.edge_weight[e]);
      else if (G.mech == MIND_MECH_EXPSYN) phase_receive(MIND_MECH_EXPSYN, cols(G), row, d.win[j].t, d.edge_weight[e]);
    }
  }
  return lo;
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
        Ops::eval(col, rr[u], in[u], v[li], tc, d.dt, value, shifted, store, d.celsius);
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

// nrn_cur of a density mechanism whose ion transfers are at most two USEION
// statements, two reads and two written currents (GroupDev::ion_fast; hh's ena,
// ek, ina, ik): a row's ion rows, columns, area and voltage are read together;
// then the ion values its reads name and the current and di/dv sums of its ion
// rows (a density mechanism has one row per node, so no other row of the group
// adds to them); then both probes and the sums. The same operations as
// ion_current: the read values are stored in their columns and replace the
// loaded ones, and each sum adds s * i with s = 1.
template <int MECH, typename C>
__device__ __forceinline__ void density_ion_currents(const Dev& d, const GroupDev& G, const Cols& col, int64_t r0,
                                                     int64_t r1, int64_t n0, double tc, const double* v, C rhs,
                                                     C diag, int tid, int T) {
  using Ops = CurOps<MECH>;
  const int ns = G.nslot, nr = G.nread, nw = G.nwrite;
  IonBind br[2], bw[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    br[k] = k < nr ? G.ion_bind[k] : IonBind{0, 0, 0, 0};
    bw[k] = k < nw ? G.ion_bind[nr + k] : IonBind{0, 0, 0, 0};
  }
  const int64_t nion = d.nion;
  for (int64_t r = r0 + tid; r < r1; r += T) {
    const int32_t node = G.node[r];
    const int32_t q0 = G.ion_row[r * ns], q1 = G.ion_row[r * ns + (ns > 1 ? 1 : 0)];
    typename Ops::In in = Ops::load(col, r);
    const double area = d.area[node];
    const int li = node - (int)n0;
    const double vn = v[li];
    int32_t qr[2], qw[2];
    double x[2], ci[2], cd[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      qr[k] = k < nr ? (br[k].slot ? q1 : q0) : -1;
      qw[k] = k < nw ? (bw[k].slot ? q1 : q0) : -1;
      x[k] = qr[k] >= 0 ? d.ion[br[k].column * nion + qr[k]] : 0.0;
      ci[k] = qw[k] >= 0 ? d.ion[MIND_ION_I * nion + qw[k]] : 0.0;
      cd[k] = qw[k] >= 0 ? d.ion[MIND_ION_DIDV * nion + qw[k]] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (qr[k] >= 0) {
        col(br[k].field, r) = x[k];
        Ops::set(in, br[k].field, x[k]);
      }
    double value, shifted, ish[2 * kMaxIonCurrents];
    Ops::eval_ish(col, r, in, vn, tc, d.dt, value, shifted, true, d.celsius, ish);
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (qw[k] >= 0) {
        const double xs = bw[k].aux ? ish[1] : ish[0];
        const double xi = bw[k].aux ? ish[kMaxIonCurrents + 1] : ish[kMaxIonCurrents];
        d.ion[MIND_ION_I * nion + qw[k]] = ci[k] + 1.0 * xi;
        d.ion[MIND_ION_DIDV * nion + qw[k]] = cd[k] + 1.0 * (xs - xi) / 0.001;
      }
    const double scale = area * 0.01;
    double acc_r = rhs[li], acc_d = diag[li];
    acc_r -= scale * value;
    acc_d += scale * (shifted - value) / 0.001;
    rhs[li] = acc_r;
    diag[li] = acc_d;
  }
}

// Mechanisms whose current-phase stores are outputs only: no other phase and
// no NET_RECEIVE reads them.
__device__ __forceinline__ bool mech_outputs_only(int32_t mech) {
  return mech == MIND_MECH_HH || mech == MIND_MECH_PAS || mech == MIND_MECH_EXPSYN || mech == MIND_MECH_EXP2SYN ||
         mech == MIND_MECH_ICLAMP || mech == MIND_MECH_ALPHASYNAPSE;
}

// nrn_cur of any group's rows [r0, r1) of a workgroup: runs of point rows on one
// node in row order, ion transfers through ion_current.
template <typename C>
__device__ __forceinline__ void group_currents(const Dev& d, const GroupDev& G, int64_t r0, int64_t r1, int64_t n0,
                                            double tc, const double* v, C rhs, C diag, bool store_outputs) {
  const int tid = threadIdx.x, T = blockDim.x;
  const Cols col = cols(G);
  for (int64_t r = r0 + tid; r < r1; r += T) {
    if (G.point && !G.run_start[r]) continue;
    const int32_t node = G.node[r];
    const int li = node - (int)n0;
    double acc_r = rhs[li], acc_d = diag[li];
    int64_t q = r;
    do {  // a run of rows on one node, in row order
      double value, shifted;
      if (G.nslot)
        ion_current(d, G, col, q, v[li], tc, d.area[node], value, shifted);
      else
        phase_current(G.mech, col, q, v[li], tc, d.dt, value, shifted, store_outputs || !mech_outputs_only(G.mech),
                      d.celsius, nullptr);
      const double scale = G.point ? 1.0 : d.area[node] * 0.01;
      acc_r -= scale * value;
      acc_d += scale * (shifted - value) / 0.001;
      ++q;
    } while (G.point && q < r1 && G.node[q] == node);
    rhs[li] = acc_r;
    diag[li] = acc_d;
  }
}

// nrn_state at t of one group's rows [r0, r1) of a workgroup, rows sharing a
// node in row order (the ion-ordered path).
__device__ __forceinline__ void group_states(const Dev& d, const GroupDev& G, int64_t r0, int64_t r1, int64_t n0,
                                          double t, const double* v) {
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    if (G.point && !G.run_start[r]) continue;
    int64_t q = r;
    do {
      advance_row(d, G, q, v[G.node[q] - n0], t);
      ++q;
    } while (G.point && q < r1 && G.node[q] == G.node[r]);
  }
}

// nrn_cur of a point-process group of builtin mechanism MECH without ion
// transfers (the core kernels' ExpSyn, Exp2Syn, IClamp): the same runs of rows on
// one node and operations as group_currents, with MECH's phase alone compiled in.
template <int MECH, typename C>
__device__ __forceinline__ void point_currents(const Dev& d, const GroupDev& G, const Cols& col, int64_t r0,
                                               int64_t r1, int64_t n0, double tc, const double* v, C rhs, C diag,
                                               int tid, int T, bool store_outputs) {
  using Ops = CurOps<MECH>;
  const bool store = store_outputs || !mech_outputs_only(MECH);
  for (int64_t r = r0 + tid; r < r1; r += T) {
    if (!G.run_start[r]) continue;
    const int32_t node = G.node[r];
    const int li = node - (int)n0;
    double acc_r = rhs[li], acc_d = diag[li];
    const double vn = v[li];
    int64_t q = r;
    do {  // a run of rows on one node, in row order
      double value, shifted;
      Ops::eval(col, q, Ops::load(col, q), vn, tc, d.dt, value, shifted, store, d.celsius);
      acc_r -= 1.0 * value;
      acc_d += 1.0 * (shifted - value) / 0.001;
      ++q;
    } while (q < r1 && G.node[q] == node);
    rhs[li] = acc_r;
    diag[li] = acc_d;
  }
}

// nrn_state of row r of a core-kernel group (hh, ExpSyn, Exp2Syn; none of them
// loads ion values here, GroupDev::ion_state being 0): the mechanism's phase
// alone, as advance_row runs it.
__device__ __forceinline__ void core_advance(const Dev& d, const GroupDev& G, int64_t r, double v, double t) {
  const Cols c = cols(G);
  if (G.mech == MIND_MECH_HH)
    phase_advance(MIND_MECH_HH, c, r, v, t, d.dt, d.celsius);
  else if (G.mech == MIND_MECH_EXP2SYN)
    phase_advance(MIND_MECH_EXP2SYN, c, r, v, t, d.dt, d.celsius);
  else if (G.mech == MIND_MECH_EXPSYN)
    phase_advance(MIND_MECH_EXPSYN, c, r, v, t, d.dt, d.celsius);
}

// Mechanisms with a state update (nrn_state).
__device__ __forceinline__ bool stateful(int32_t mech) {
  return phase_has_state(mech);
}

// Per-window constants of the step's critical path: mechanism groups, their row
// ranges for this workgroup's cells, and its detectors.
struct WinShared {
  GroupDev sg[kMaxGroups];
  int64_t sr0[kMaxGroups], sr1[kMaxGroups];
  int32_t sdet_k[kMaxDetLds], sdet_n[kMaxDetLds];
  double sdet_thr[kMaxDetLds];
};

// One workgroup's window. CORE (the host's State::core) compiles in only the
// inlined mechanism paths: pas and hh density currents (hh with its own ion
// currents), ExpSyn / Exp2Syn / IClamp point currents, their state updates and
// receivers, no ion orders or secondorder-2 correction. The general kernels run
// every other model, with the same operations for these mechanisms.
template <int M, bool CORE>
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
  const int64_t iq0 = d.nion ? d.node_ion_off[n0] : 0, iq1 = d.nion ? d.node_ion_off[n1] : 0;  // ion rows
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
    // midpoint delivery (nrn_cur follows a barrier)
    if (CORE) seg = deliver_core_runs(d, seg, seg_end, 2 * k, tid, T);
    else if (tid == 0) seg = deliver_group(d, seg, seg_end, 2 * k);
    for (int i = tid; i < nn; i += T) {
      rhs[i] = 0.0;
      diag[i] = 0.0;
    }
    for (int64_t q = iq0 + tid; q < iq1; q += T) ion_prepare(d, q, 0x40);
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
      if (G.mech == MIND_MECH_PAS && !G.point && !G.nslot) {
        density_currents<MIND_MECH_PAS, 4>(d, col, G.node, r0, r1, n0, tc, v, rhs, diag, tid, T, store_outputs);
        __syncthreads();
        continue;
      }
      if (G.mech == MIND_MECH_HH && !G.point && !G.nslot) {
        density_currents<MIND_MECH_HH, 1>(d, col, G.node, r0, r1, n0, tc, v, rhs, diag, tid, T, store_outputs);
        __syncthreads();
        continue;
      }
      if (G.mech == MIND_MECH_HH && G.ion_fast) {
        density_ion_currents<MIND_MECH_HH>(d, G, col, r0, r1, n0, tc, v, rhs, diag, tid, T);
        __syncthreads();
        continue;
      }
      if (!CORE)
        group_currents(d, G, r0, r1, n0, tc, v, rhs, diag, store_outputs);
      else if (G.mech == MIND_MECH_EXP2SYN)
        point_currents<MIND_MECH_EXP2SYN>(d, G, col, r0, r1, n0, tc, v, rhs, diag, tid, T, store_outputs);
      else if (G.mech == MIND_MECH_EXPSYN)
        point_currents<MIND_MECH_EXPSYN>(d, G, col, r0, r1, n0, tc, v, rhs, diag, tid, T, store_outputs);
      else if (G.mech == MIND_MECH_ICLAMP)
        point_currents<MIND_MECH_ICLAMP>(d, G, col, r0, r1, n0, tc, v, rhs, diag, tid, T, store_outputs);
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
    if (!CORE && d.ion_correct && d.nion) {  // secondorder 2: ion currents += di/dv * dv
      for (int i = tid; i < nn; i += T) ion_correct_node(d, n0 + i, rhs[i]);
      __syncthreads();
    }
    if (!CORE && d.ion_ordered) {
      // nrn_state at t + dt group by group, as the reference's launches run: a
      // mechanism reads what an earlier one wrote to the ion rows
      for (int gi = 0; gi < ngr; ++gi) {
        if (!stateful(sg[gi].mech)) continue;
        group_states(d, sg[gi], sr0[gi], sr1[gi], n0, t + dt, v);
        __syncthreads();
      }
    } else {
    // nrn_state at t + dt: the rows of all stateful groups as one index space
    // (rows are independent), so that the groups' rows run side by side; each
    // group's rows start at a multiple of 64, so that no wavefront runs two
    // mechanisms' updates one after the other under divergent masks
    {
      auto span = [](int64_t cnt) { return (cnt + 63) & ~(int64_t)63; };
      int64_t total = 0;
      for (int gi = 0; gi < ngr; ++gi)
        if (stateful(sg[gi].mech)) total += span(sr1[gi] - sr0[gi]);
      for (int64_t f = tid; f < total; f += T) {
        int gi = 0;
        int64_t r = f;
        for (;; ++gi) {
          if (!stateful(sg[gi].mech)) continue;
          const int64_t cnt = span(sr1[gi] - sr0[gi]);
          if (r < cnt) break;
          r -= cnt;
        }
        if (r >= sr1[gi] - sr0[gi]) continue;  // the padding of the group's last wavefront
        const GroupDev& G = sg[gi];
        r += sr0[gi];
        if (CORE)
          core_advance(d, G, r, v[G.node[r] - n0], t + dt);
        else
          advance_row(d, G, r, v[G.node[r] - n0], t + dt);
      }
    }
    }
    if (rec1 > rec0) {
      __syncthreads();
      const int64_t sample = k + 1 - d.rec_start;
      for (int64_t q = rec0 + tid; q < rec1; q += T) {
        const int32_t rec = d.cell_rec[q];
        const double x = d.rec_kind[rec] == 0 ? v[d.rec_index[rec] - n0] : stored_value(d, rec);
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
    // end-of-step delivery; with events, a barrier before the next step's midpoint
    // delivery, which may deliver to the same rows
    if (CORE) {
      const int64_t next = deliver_core_runs(d, seg, seg_end, 2 * k + 1, tid, T);
      if (next != seg) __syncthreads();
      seg = next;
    } else if (tid == 0) {
      seg = deliver_group(d, seg, seg_end, 2 * k + 1);
    }
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
  window_body<0, false>(d, b, groups[blockIdx.x], n, sh, lds);
}

// Groups within the node budget in the lean layout, v in LDS.
__global__ void __launch_bounds__(kMaxThreads) k_window_lean(Dev d, BatchDev b, const int32_t* groups, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<5, false>(d, b, groups[blockIdx.x], n, sh, lds);
}

// Groups within the node budget in the lean layout, v in HBM.
__global__ void __launch_bounds__(kMaxThreads) k_window_lean_hbm(Dev d, BatchDev b, const int32_t* groups,
                                                                 int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<4, false>(d, b, groups[blockIdx.x], n, sh, lds);
}

// Cells above the node budget, one per workgroup, each with the layout that
// fits (modes[k] for groups[k]): one launch, so that they run side by side.
__global__ void __launch_bounds__(kLargeThreads) k_window_large(Dev d, BatchDev b, const int32_t* groups,
                                                              const int8_t* modes, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  const int grp = groups[blockIdx.x];
  switch (modes[blockIdx.x]) {
    case 0: window_body<0, false>(d, b, grp, n, sh, lds); break;
    case 4: window_body<4, false>(d, b, grp, n, sh, lds); break;
    case 5: window_body<5, false>(d, b, grp, n, sh, lds); break;
    default: window_body<3, false>(d, b, grp, n, sh, lds); break;
  }
}

#ifndef __HIPCC_RTC__
// The same kernels for models the core paths cover (State::core). Models with
// runtime-compiled mechanisms never take them, so the hipRTC build leaves them out.
__global__ void __launch_bounds__(kMaxThreads) k_window_core(Dev d, BatchDev b, const int32_t* groups, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<0, true>(d, b, groups[blockIdx.x], n, sh, lds);
}
__global__ void __launch_bounds__(kMaxThreads) k_window_lean_core(Dev d, BatchDev b, const int32_t* groups,
                                                                  int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<5, true>(d, b, groups[blockIdx.x], n, sh, lds);
}
__global__ void __launch_bounds__(kMaxThreads) k_window_lean_hbm_core(Dev d, BatchDev b, const int32_t* groups,
                                                                      int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  window_body<4, true>(d, b, groups[blockIdx.x], n, sh, lds);
}
__global__ void __launch_bounds__(kLargeThreads) k_window_large_core(Dev d, BatchDev b, const int32_t* groups,
                                                                   const int8_t* modes, int32_t n) {
  extern __shared__ double lds[];
  __shared__ WinShared sh;
  const int grp = groups[blockIdx.x];
  switch (modes[blockIdx.x]) {
    case 0: window_body<0, true>(d, b, grp, n, sh, lds); break;
    case 4: window_body<4, true>(d, b, grp, n, sh, lds); break;
    case 5: window_body<5, true>(d, b, grp, n, sh, lds); break;
    default: window_body<3, true>(d, b, grp, n, sh, lds); break;
  }
}
#endif

}  // namespace mind

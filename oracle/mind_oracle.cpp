// CPU oracle for the MIND simulation step — TEST INFRASTRUCTURE ONLY.
//
// A scalar FP64 restatement of the reference's micro step, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker of
// the HIP path. The product library never links or calls this file.
//
// Each part cites the reference function it restates (paths relative to the
// reference repository):
//   geometry      src/model/micro/morphology.jl:406-484, utils/morphology.jl:3-119
//   packing       src/model/micro/model.jl:420-902 (rows, ranks, fanout)
//   mechanisms    mod/{hh,pas,ExpSyn,Exp2Syn,IClamp,NetStim}.mod lowered as in
//                 src/compiler/codegen.jl:362-485 (phases), :569-641 (cnexp)
//   currents      src/runtime/cuda/micro.jl:427-500, 601-629
//   cable         src/runtime/cuda/micro.jl:633-749
//   events        src/runtime/events.jl:4-165, utils/events.jl:7-62,
//                 src/runtime/initialize.jl:604-649 (ranks)
//   Philox        src/runtime/cuda/micro.jl:3-91
//   step / init   src/runtime/step.jl:4-151, src/runtime/initialize.jl:270-429
//   recording     src/runtime/recording.jl:256-295
//
// Canonical orders. The reference sums through FP64 atomics whose order is not
// defined (cuda/micro.jl:471-472, 652-653, 722-723). This restatement fixes one
// order and the device path follows it:
//   rhs/diag at a node: density groups in group order, then point groups in
//   group order (rows in row order), then the node's own axial current, then
//   its children's in increasing node index; Hines elimination visits nodes
//   in decreasing index order (children before parents within a cell).
//
// Parity status (how each part is pinned, see tests/test_oracle.py):
//   geometry       closed forms for cylinders
//   Hines solve    dense numpy solve on random trees
//   hh / pas / syn closed forms, textbook HH rates, mpmath
//   Philox4x32-10  published Random123 known-answer vectors
//   event order    hand-computed orders on tiny networks
//   exp/log        shared with the kernels (csrc/mind_math.h), pinned against mpmath
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <set>
#include <tuple>
#include <string>
#include <vector>

#include "../include/mind.h"
#include "../mind.jl_amd/csrc/mind_math.h"

namespace {

// ---------------------------------------------------------------- Philox ----
// Restates refresh_random_words!/uniform_random!/set_random_sequence!
// (src/runtime/cuda/micro.jl:3-67).
void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
  uint32_t a = ctr[0], b = ctr[1], c = ctr[2], d = ctr[3];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xd2511f53u * a;
    const uint64_t p1 = (uint64_t)0xcd9e8d57u * c;
    const uint32_t na = (uint32_t)(p1 >> 32) ^ b ^ k0;
    const uint32_t nb = (uint32_t)p1;
    const uint32_t nc = (uint32_t)(p0 >> 32) ^ d ^ k1;
    const uint32_t nd = (uint32_t)p0;
    a = na; b = nb; c = nc; d = nd;
    k0 += 0x9e3779b9u;
    k1 += 0xbb67ae85u;
  }
  ctr[0] = a; ctr[1] = b; ctr[2] = c; ctr[3] = d;
}

struct Rng {  // one Random123 stream of one mechanism row
  uint64_t counter = 0;
  uint64_t keys[3] = {0, 0, 0};  // id lo32, id hi32, digest ^ (stream-1)
  uint32_t words[4] = {0, 0, 0, 0};
};

void rng_refresh(Rng& r, uint64_t seed) {
  uint32_t ctr[4] = {(uint32_t)(r.counter >> 2), (uint32_t)r.keys[2], (uint32_t)r.keys[0],
                     (uint32_t)r.keys[1]};
  philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
  for (int i = 0; i < 4; ++i) r.words[i] = ctr[i];
}

double rng_uniform(Rng& r, uint64_t seed) {
  const uint64_t counter = r.counter;
  const int which = (int)(counter & 3u);
  const uint32_t word = r.words[which];
  const uint64_t next = (counter + 1) & 0x3ffffffffull;
  r.counter = next;
  if (which == 3) {
    if (next == 0) {
      const int order[3] = {2, 0, 1};
      for (int o : order) {
        const uint32_t v = (uint32_t)(r.keys[o] + 1);
        r.keys[o] = v;
        if (v != 0) break;
      }
    }
    rng_refresh(r, seed);
  }
  return ((double)word + 1.0) * (1.0 / 4294967297.0);
}

void rng_setseq(Rng& r, uint64_t seed, double sequence) {
  const double bounded = (sequence >= 0.0 && sequence <= 1099511627775.0) ? sequence : 0.0;
  r.counter = ((uint64_t)trunc(bounded)) & 0x3ffffffffull;
  rng_refresh(r, seed);
}

// -------------------------------------------------------------- Geometry ----
struct Pt { double x, y, z, d; };

struct Template {  // one discretized morphology (MorphologyLayout)
  std::vector<int32_t> parent;  // 0-based, -1 = root
  std::vector<double> area, diameter, axial, cap;
  std::vector<std::vector<int32_t>> sec_nodes;  // per section: nseg+1 nodes
  std::vector<int32_t> sec_start;
};

struct SectionIn {
  int32_t parent; double parent_x, child_x; int32_t nseg; double length, diameter, Ra, cm;
  std::vector<Pt> points;
};

// integrate_geometry (utils/morphology.jl:28-63)
void integrate(const std::vector<Pt>& p, const std::vector<double>& arcs, double boundary,
               double Ra, int& index, double& prev_arc, double& prev_d, double& area_out,
               double& res_out, double& dint_out) {
  double area = 0.0, resistance = 0.0, dint = 0.0;
  const int n = (int)p.size();
  while (true) {
    const int next = index + 1;
    const bool at_boundary = arcs[next] > boundary || next == n - 1;
    double arc, diameter;
    if (at_boundary) {
      const double span = arcs[next] - arcs[index];
      const double fraction = fabs(span) < 1.0e-10 ? 1.0 : (boundary - arcs[index]) / span;
      arc = boundary;
      diameter = (1.0 - fraction) * p[index].d + fraction * p[next].d;
    } else {
      arc = arcs[next];
      diameter = p[next].d;
    }
    const double delta = arc - prev_arc;
    dint += (prev_d + diameter) * delta;
    const double distance = delta > 1.0e-15 ? delta : 1.0e-15;
    const double conductance = prev_d * diameter / distance;
    resistance += 1.0 / (conductance == 0.0 ? 1.0e-15 : conductance);
    area += (prev_d + diameter) *
            sqrt(distance * distance + 0.25 * ((diameter - prev_d) * (diameter - prev_d)));
    prev_arc = arc;
    prev_d = diameter;
    if (at_boundary) break;
    index = next;
  }
  area_out = 0.5 * M_PI * area;
  res_out = 0.04 * Ra / M_PI * resistance;
  dint_out = 0.5 * dint;
}

int32_t location_node(const Template& t, int sec, double x) {  // utils/morphology.jl:110-119
  if (x == 0.0) return t.sec_start[sec];
  const std::vector<int32_t>& nodes = t.sec_nodes[sec];
  const int len = (int)nodes.size();
  if (x == 1.0) return nodes[len - 1];
  int index = (int)floor(x * (len - 1));
  if (index > len - 2) index = len - 2;
  return nodes[index];
}

int32_t range_node(const Template& t, int sec, double x) {  // utils/morphology.jl:68-75
  const std::vector<int32_t>& nodes = t.sec_nodes[sec];
  const int count = (int)nodes.size() - 1;
  int index = (int)floor(x * count);
  if (index > count - 1) index = count - 1;
  return nodes[index];
}

// section_order (utils/morphology.jl:77-104): parents before children, input order otherwise
std::vector<int> section_order(const std::vector<SectionIn>& s) {
  const int n = (int)s.size();
  std::vector<uint8_t> marks(n, 0);
  std::vector<int> order, pending;
  for (int start = 0; start < n; ++start) {
    int index = start;
    while (marks[index] != 2) {
      marks[index] = 1;
      pending.push_back(index);
      if (s[index].parent < 0) break;
      index = s[index].parent;
    }
    while (!pending.empty()) {
      const int i = pending.back();
      pending.pop_back();
      marks[i] = 2;
      order.push_back(i);
    }
  }
  return order;
}

// discretize (morphology.jl:406-484)
Template discretize(const std::vector<SectionIn>& secs) {
  Template t;
  t.sec_nodes.resize(secs.size());
  t.sec_start.resize(secs.size());
  auto add = [&](int32_t parent, double area, double diam, double g, double c) {
    t.parent.push_back(parent);
    t.area.push_back(area);
    t.diameter.push_back(diam);
    t.axial.push_back(g);
    t.cap.push_back(c);
    return (int32_t)t.parent.size() - 1;
  };
  for (int si : section_order(secs)) {
    const SectionIn& s = secs[si];
    std::vector<Pt> pts = s.points;
    if (pts.size() < 2) pts = {{0.0, 0.0, 0.0, s.diameter}, {s.length, 0.0, 0.0, s.diameter}};
    std::vector<double> arcs(pts.size(), 0.0);
    for (size_t i = 1; i < pts.size(); ++i) {
      const double dx = pts[i].x - pts[i - 1].x, dy = pts[i].y - pts[i - 1].y,
                   dz = pts[i].z - pts[i - 1].z;
      arcs[i] = arcs[i - 1] + sqrt(dx * dx + dy * dy + dz * dz);
    }
    int32_t connection;
    if (s.parent < 0) connection = add(-1, 100.0, 0.0, 0.0, 0.0);
    else connection = location_node(t, s.parent, s.parent_x);
    std::vector<int32_t> nodes(s.nseg + 1);
    int32_t previous = connection;
    double previous_right = 0.0;
    const double step = arcs.back() / s.nseg;
    int index = 0;
    double prev_arc = arcs[0], prev_d = pts[0].d;
    const bool reversed = s.child_x == 1.0;
    for (int ordinal = 1; ordinal <= s.nseg; ++ordinal) {
      const double middle = (ordinal - 0.5) * step;
      const double right = ordinal * step;
      double al, rl, dl, ar, rr, dr;
      integrate(pts, arcs, middle, s.Ra, index, prev_arc, prev_d, al, rl, dl);
      integrate(pts, arcs, right, s.Ra, index, prev_arc, prev_d, ar, rr, dr);
      const double area = al + ar;
      const int segment = reversed ? s.nseg + 1 - ordinal : ordinal;
      const int32_t node = add(previous, area, (dl + dr) / step, 1.0 / (previous_right + rl),
                               s.cm * area * 1.0e-5);
      nodes[segment - 1] = node;
      previous = node;
      previous_right = rr;
    }
    const int32_t endpoint = add(previous, 100.0, 0.0, 1.0 / previous_right, 0.0);
    t.sec_start[si] = reversed ? endpoint : connection;
    nodes[s.nseg] = reversed ? connection : endpoint;
    t.sec_nodes[si] = nodes;
  }
  return t;
}

// ------------------------------------------------------------ Mechanisms ----
const int kFields[MIND_MECH_COUNT] = {HH_NFIELDS,     PAS_NFIELDS,    EXPSYN_NFIELDS,
                                      EXP2SYN_NFIELDS, ICLAMP_NFIELDS, NETSTIM_NFIELDS,
                                      ALPHASYNAPSE_NFIELDS, PATTERNSTIM_NFIELDS, SECLAMP_NFIELDS,
                                      VCLAMP_NFIELDS};
const bool kPoint[MIND_MECH_COUNT] = {false, false, true, true, true, true, true, true, true, true};

// A runtime-compiled mechanism: its phases come from a host build (g++) of the
// library's lowering of its MOD text, registered by oracle_register_mech.
using UserCurrent = void (*)(double* data, int64_t rows, int64_t r, double v, double t, double dt, double celsius,
                             double* value, double* shifted, int32_t store, double* ish);
using UserAdvance = void (*)(double* data, int64_t rows, int64_t r, double v, double t, double dt, double celsius);
using UserInitial = void (*)(double* data, int64_t rows, int64_t r, double v, double celsius);
using UserReceive = void (*)(double* data, int64_t rows, int64_t r, double t, double weight);
// One USEION statement: per ion column (0 reversal, 1 internal, 2 external
// concentration, 3 current) the field read into / written from, -1 if none.
struct IonDecl {
  std::string name;
  bool has_valence = false;
  double valence = 0.0;
  int read[4] = {-1, -1, -1, -1};
  int write[4] = {-1, -1, -1, -1};
};
struct UserMech {
  int nfields = 0;
  bool point = false, has_state = false;
  std::vector<double> defaults;
  std::vector<IonDecl> ions;
  std::vector<uint8_t> state_col;
  std::vector<int> recv_fields;  // columns NET_RECEIVE touches
  UserCurrent current = nullptr;
  UserAdvance advance = nullptr;
  UserInitial initial = nullptr;
  UserReceive receive = nullptr;
};
UserMech g_user[MIND_MECH_SLOTS];

int field_count(int mech) { return mech < MIND_MECH_COUNT ? kFields[mech] : g_user[mech].nfields; }

// USEION statements (hh.mod: USEION na READ ena WRITE ina, USEION k READ ek WRITE ik).
const std::vector<IonDecl>& ion_decls(int mech) {
  static const std::vector<IonDecl> none;
  static const std::vector<IonDecl> hh = [] {
    IonDecl na, k;
    na.name = "na"; na.read[0] = HH_ENA; na.write[3] = HH_INA;
    k.name = "k"; k.read[0] = HH_EK; k.write[3] = HH_IK;
    return std::vector<IonDecl>{na, k};
  }();
  if (mech == MIND_MECH_HH) return hh;
  if (mech >= MIND_MECH_COUNT && mech < MIND_MECH_SLOTS) return g_user[mech].ions;
  return none;
}
bool is_point(int mech) { return mech < MIND_MECH_COUNT ? kPoint[mech] : g_user[mech].point; }

void defaults(int mech, double* f) {  // PARAMETER defaults of the MOD files
  if (mech >= MIND_MECH_COUNT) {
    for (int k = 0; k < g_user[mech].nfields; ++k) f[k] = g_user[mech].defaults[k];
    return;
  }
  switch (mech) {
    case MIND_MECH_HH:
      f[HH_GNABAR] = 0.12; f[HH_GKBAR] = 0.036; f[HH_GL] = 0.0003; f[HH_EL] = -54.3;
      f[HH_ENA] = 50.0; f[HH_EK] = -77.0;  // default_ions (utils/biophysics.jl:31-41)
      break;
    case MIND_MECH_PAS: f[PAS_G] = 0.001; f[PAS_E] = -70.0; break;
    case MIND_MECH_EXPSYN: f[EXPSYN_TAU] = 0.1; f[EXPSYN_E] = 0.0; break;
    case MIND_MECH_EXP2SYN: f[EXP2SYN_TAU1] = 0.1; f[EXP2SYN_TAU2] = 10.0; f[EXP2SYN_E] = 0.0; break;
    case MIND_MECH_ICLAMP: break;
    case MIND_MECH_NETSTIM:
      f[NETSTIM_INTERVAL] = 10.0; f[NETSTIM_NUMBER] = 10.0; f[NETSTIM_START] = 50.0;
      f[NETSTIM_NOISE] = 0.0;
      break;
    case MIND_MECH_ALPHASYNAPSE: f[ALPHASYNAPSE_TAU] = 0.1; break;  // AlphaSynapse.mod PARAMETER
    case MIND_MECH_SECLAMP: f[SECLAMP_RS] = 1.0; break;              // SEClamp.mod PARAMETER
    case MIND_MECH_VCLAMP:                                             // VClamp.mod PARAMETER
      f[VCLAMP_GAIN] = 1e5; f[VCLAMP_RSTIM] = 1.0; f[VCLAMP_TAU1] = 0.001;
      break;
  }
}

// hh.mod PROCEDURE rates / FUNCTION vtrap (hh.mod:90-122)
double vtrap(double x, double y) {
  if (fabs(x / y) < 1e-6) return y * (1.0 - x / y / 2.0);
  return x / (mind_exp(x / y) - 1.0);
}
void hh_rates(double* f, double v, double celsius) {
  const double q10 = mind_pow(3.0, (celsius - 6.3) / 10.0);
  double alpha = 0.1 * vtrap(-(v + 40.0), 10.0);
  double beta = 4.0 * mind_exp(-(v + 65.0) / 18.0);
  double sum = alpha + beta;
  f[HH_MTAU] = 1.0 / (q10 * sum);
  f[HH_MINF] = alpha / sum;
  alpha = 0.07 * mind_exp(-(v + 65.0) / 20.0);
  beta = 1.0 / (mind_exp(-(v + 35.0) / 10.0) + 1.0);
  sum = alpha + beta;
  f[HH_HTAU] = 1.0 / (q10 * sum);
  f[HH_HINF] = alpha / sum;
  alpha = 0.01 * vtrap(-(v + 55.0), 10.0);
  beta = 0.125 * mind_exp(-(v + 65.0) / 80.0);
  sum = alpha + beta;
  f[HH_NTAU] = 1.0 / (q10 * sum);
  f[HH_NINF] = alpha / sum;
}

// cnexp relaxation x += (1 - exp(dt*(-1/tau))) * (xinf - x)   (codegen.jl:625-627)
inline double relax(double x, double xinf, double tau, double dt) {
  return x + (1.0 - mind_exp(dt * (-1.0 / tau))) * (xinf - x);
}

// Nernst potential and IonStyle flags (src/model/micro/utils/biophysics.jl:3-70).
double nernst(double ci, double co, double z, double celsius) {
  if (z == 0.0) return 0.0;
  if (ci <= 0.0) return 1.0e6;
  if (co <= 0.0) return -1.0e6;
  const double r = 1.380649e-23 * 6.02214076e23, f = 1.602176634e-19 * 6.02214076e23;
  const double rt_f = 1000.0 * r * (celsius + 273.15) / f;
  return rt_f / z * mind_log(co / ci);
}
uint8_t style_bits(int c_style, int e_style, bool einit, bool eadvance, bool cinit) {
  return (uint8_t)(c_style | (cinit ? 0x04 : 0) | (e_style << 3) | (einit ? 0x20 : 0) | (eadvance ? 0x40 : 0));
}
uint8_t promote(uint8_t style, int c_style, int e_style) {
  const int c = std::max<int>(style & 0x03, c_style);
  int e = std::max<int>((style >> 3) & 0x03, e_style);
  if (c > 0 && e < 2) e = 2;
  return style_bits(c, e, c > 0 && e == 2, c == 3 && e == 2, c == 3);
}

// ------------------------------------------------------------------ Model ---
struct Group {
  int mech;
  std::vector<int32_t> node;     // packed node (-1 for artificial)
  std::vector<int32_t> target;   // point index + 1 (0 for density rows)
  std::vector<uint64_t> logical;
  std::vector<double> data;      // field-major: data[field*rows + row]
  std::vector<double> initial;   // data snapshot after parameters (restored by initialize)
  std::vector<Rng> rng;
  std::vector<int64_t> cell_rows;  // ncells+1: rows of cell c are [cell_rows[c], cell_rows[c+1])
  std::vector<int32_t> slot_species;  // species of each USEION statement
  std::vector<int32_t> ion_row;       // ion row of (row, statement), row-major; -1 none
  int rows() const { return (int)node.size(); }
  double& at(int field, int row) { return data[(size_t)field * node.size() + row]; }
};

struct Event {  // ScheduledEvent (src/runtime/state.jl:18-24)
  double time, flag;
  int64_t edge, order;
  int32_t target;  // > 0 point index + 1, < 0 -(source index + 1)
};

bool event_before(const Event& a, const Event& b) {  // utils/events.jl:7-10
  return a.time < b.time ||
         (a.time == b.time && (a.order < b.order || (a.order == b.order && a.target < b.target)));
}

struct Sim {
  // configuration
  double dt = 0.025, celsius = 6.3;
  uint64_t seed = 1;
  int32_t secondorder = 0;  // build! keyword: 1, 2 Crank-Nicolson (step.jl:112, cuda/micro.jl:660-679)
  int threads = 1;          // OpenMP threads of the membrane phases
  // nodes
  int64_t total = 0, compartments = 0;
  std::vector<int32_t> parent;
  std::vector<double> area, cap, axial, axial_diag, v, v_init, rhs, diag;
  std::vector<std::vector<int32_t>> children;
  std::vector<std::vector<int32_t>> cell_sec_nodes_offset;
  std::vector<int64_t> cell_offset;
  std::vector<int32_t> cell_morph;
  std::vector<Template> templates;
  std::vector<std::vector<SectionIn>> morph_secs;
  // ions (model.jl:672-855): rows per (node, species), sorted by node then name
  struct Species { std::string name; double z, ci, co, e; bool nernst; };
  std::vector<Species> species;
  int64_t nion = 0;
  std::vector<int32_t> ion_sp;
  std::vector<uint8_t> ion_style;
  std::vector<double> ion[5], ion_packed[5];  // e, ci, co, i, di/dv
  std::vector<int64_t> ion_off;               // per node (CSR)
  // mechanisms
  std::vector<Group> groups;
  std::vector<int> group_of_mech;       // mech -> group index or -1
  std::vector<int32_t> point_group, point_row;
  // sources / detectors
  int64_t n_sources = 0;
  std::vector<int64_t> source_sid;
  std::vector<int32_t> source_target;   // point index + 1 for event sources, 0 for detectors
  std::vector<int32_t> det_node, det_source;
  std::vector<double> det_threshold;
  std::vector<uint8_t> det_above;
  std::vector<std::vector<int32_t>> target_sources;  // point -> sources with that point
  // connections (CSR by source)
  std::vector<int64_t> fanout;          // n_sources+1
  std::vector<int32_t> edge_target;     // point index + 1
  std::vector<double> edge_delay, edge_weight;
  // event machinery
  std::vector<Event> heap, outbox;
  std::vector<int64_t> sequences;       // per point target
  std::vector<int64_t> target_rank, source_rank;
  std::vector<uint8_t> feedback_group;
  int64_t next_order = 0;
  std::vector<int64_t> connection_edge;  // edge of each connection in creation order
  std::vector<double> pattern_time;      // PatternStim spikes (time, source) in enqueue order
  std::vector<int32_t> pattern_source;
  std::vector<int32_t> source_external;  // external spike inputs
  int64_t step = 0;
  int error = 0;
  bool initialized = false;
  // recording
  bool recording_spikes = false;
  std::vector<double> spike_t;
  std::vector<int64_t> spike_src;
  // micro-macro coupling: spikes of the current macro interval (source, time)
  int64_t rank_total = 0;
  bool collect_window = false;
  std::vector<int32_t> window_src;
  std::vector<double> window_t;
  void* macro = nullptr;
  // delivery log for tests: (point, time, order) per applied event
  bool log_deliveries = false;
  std::vector<double> log_time;
  std::vector<int64_t> log_order;
  std::vector<int64_t> log_edge;
  std::vector<int32_t> log_target;
};

thread_local std::string g_error;

int fail(int code, const std::string& message) {
  g_error = message;
  return code;
}

// ------------------------------------------------------------------ Ions ----
// prepare_currents_kernel! for ion row q (cuda/micro.jl:601-629).
void ion_prepare(Sim& s, int64_t q, uint8_t mask) {
  s.ion[3][q] = 0.0;
  s.ion[4][q] = 0.0;
  if (s.ion_style[q] & mask) s.ion[0][q] = nernst(s.ion[1][q], s.ion[2][q], s.species[s.ion_sp[q]].z, s.celsius);
}
// load_ions! / store_ions! (cuda/micro.jl:376-392)
void ion_load(Sim& s, Group& g, int r) {
  const std::vector<IonDecl>& decl = ion_decls(g.mech);
  for (size_t k = 0; k < decl.size(); ++k) {
    const int32_t q = g.ion_row[(size_t)r * decl.size() + k];
    if (q < 0) continue;
    for (int c = 0; c < 4; ++c)
      if (decl[k].read[c] >= 0) g.at(decl[k].read[c], r) = s.ion[c][q];
  }
}
void ion_store(Sim& s, Group& g, int r) {
  const std::vector<IonDecl>& decl = ion_decls(g.mech);
  for (size_t k = 0; k < decl.size(); ++k) {
    const int32_t q = g.ion_row[(size_t)r * decl.size() + k];
    if (q < 0) continue;
    for (int c = 0; c < 3; ++c)
      if (decl[k].write[c] >= 0) s.ion[c][q] = g.at(decl[k].write[c], r);
  }
}
// The ion writes of nrn_cur (mechanism_current_instance!, cuda/micro.jl:474-497):
// currents summed with their derivative from the v + 0.001 probe `ish`, scaled
// by 100/area for point processes; other writes stored.
void ion_current_writes(Sim& s, Group& g, int r, const double* ish) {
  const std::vector<IonDecl>& decl = ion_decls(g.mech);
  int probe = 0;
  for (size_t k = 0; k < decl.size(); ++k) {
    const int32_t q = g.ion_row[(size_t)r * decl.size() + k];
    const bool cur = decl[k].write[3] >= 0;
    if (q >= 0) {
      for (int c = 0; c < 3; ++c)
        if (decl[k].write[c] >= 0) s.ion[c][q] = g.at(decl[k].write[c], r);
      if (cur) {
        const double scale = is_point(g.mech) ? 100.0 / s.area[g.node[r]] : 1.0;
        const double x = g.at(decl[k].write[3], r);
        s.ion[3][q] += scale * x;
        s.ion[4][q] += scale * (ish[probe] - x) / 0.001;
      }
    }
    probe += cur ? 1 : 0;
  }
}
// receive_ions! (cuda/micro.jl:324-337): concentrations NET_RECEIVE mentions.
void ion_receive(Sim& s, Group& g, int r, bool publish) {
  if (g.mech < MIND_MECH_COUNT) return;
  const UserMech& u = g_user[g.mech];
  const std::vector<IonDecl>& decl = u.ions;
  for (size_t k = 0; k < decl.size(); ++k) {
    const int32_t q = g.ion_row[(size_t)r * decl.size() + k];
    if (q < 0) continue;
    for (int c = 1; c < 3; ++c) {
      const int f = decl[k].read[c];
      if (f < 0 || std::find(u.recv_fields.begin(), u.recv_fields.end(), f) == u.recv_fields.end()) continue;
      if (!publish) g.at(f, r) = s.ion[c][q];
      else if (u.state_col[f]) s.ion[c][q] = g.at(f, r);
    }
  }
}

// --------------------------------------------------------------- Events -----
void push_event(Sim& s, const Event& e) {  // push_event! (utils/events.jl:12-24)
  s.heap.push_back(e);
  size_t child = s.heap.size() - 1;
  while (child > 0) {
    const size_t parent = (child - 1) / 2;
    if (!event_before(e, s.heap[parent])) break;
    s.heap[child] = s.heap[parent];
    child = parent;
  }
  s.heap[child] = e;
}

Event pop_event(Sim& s) {  // pop_event! (utils/events.jl:26-40)
  const Event top = s.heap[0];
  const Event replacement = s.heap.back();
  s.heap.pop_back();
  const size_t count = s.heap.size();
  if (count == 0) return top;
  size_t parent = 0;
  while (2 * parent + 1 < count) {
    size_t child = 2 * parent + 1;
    if (child + 1 < count && event_before(s.heap[child + 1], s.heap[child])) child += 1;
    if (!event_before(s.heap[child], replacement)) break;
    s.heap[parent] = s.heap[child];
    parent = child;
  }
  s.heap[parent] = replacement;
  return top;
}

void enqueue(Sim& s, Event e) {  // enqueue_event! (utils/events.jl:42-50)
  s.next_order += 1;
  e.order = s.next_order;
  push_event(s, e);
}

void record_spike(Sim& s, int32_t source, double time) {
  if (s.collect_window) {
    s.window_src.push_back(source);
    s.window_t.push_back(time);
  }
  if (!s.recording_spikes) return;
  s.spike_t.push_back(time);
  s.spike_src.push_back(source);
}

void route_spike(Sim& s, int32_t source, double time) {  // route_spike! (utils/events.jl:52-62)
  for (int64_t e = s.fanout[source]; e < s.fanout[source + 1]; ++e)
    enqueue(s, Event{time + s.edge_delay[e], 0.0, e + 1, e + 1, s.edge_target[e]});
}

void flush(Sim& s) {  // flush_spikes! (events.jl:4-54)
  if (s.outbox.empty()) return;
  std::vector<Event> req;
  req.swap(s.outbox);
  auto rank = [&](const Event& e) -> int64_t {
    return e.target > 0 ? s.target_rank[e.target - 1]
                        : e.target < 0 ? s.source_rank[-e.target - 1] : 0;
  };
  std::stable_sort(req.begin(), req.end(), [&](const Event& a, const Event& b) {
    const int64_t ra = rank(a), rb = rank(b);
    return ra < rb || (ra == rb && a.order < b.order);
  });
  for (const Event& e : req) {
    if (e.target > 0) enqueue(s, e);
    else if (e.target < 0) route_spike(s, -e.target - 1, e.time);
  }
}

// net_send / net_event lowering (cuda/micro.jl:104-118, 149-167)
void net_send(Sim& s, int point, double t, int64_t edge, double delay, double flag) {
  if (delay < 0.0 || !mind_isfinite(delay)) {
    if (s.error == 0 || s.error < MIND_EEVENT) s.error = MIND_EEVENT;
    return;
  }
  const int64_t seq = ++s.sequences[point];
  s.outbox.push_back(Event{t + delay, flag, edge, seq, point + 1});
}

void net_event(Sim& s, int point, double t, int64_t edge) {
  const int64_t seq = ++s.sequences[point];
  for (int32_t source : s.target_sources[point]) {
    record_spike(s, source, t);
    s.outbox.push_back(Event{t, 0.0, edge, seq, -(source + 1)});
  }
}

// NetStim (NetStim.mod:1-105)
double netstim_invl(Sim& s, Group& g, int row, double mean) {
  if (mean <= 0.0) mean = 0.01;
  const double noise = g.at(NETSTIM_NOISE, row);
  if (noise == 0.0) return mean;
  const double erand = -mind_log(rng_uniform(g.rng[row], s.seed)) * 1.0;
  return (1.0 - noise) * mean + noise * mean * erand;
}

void netstim_init_sequence(Group& g, int row) {
  if (g.at(NETSTIM_NUMBER, row) > 0.0) {
    g.at(NETSTIM_ON, row) = 1.0;
    g.at(NETSTIM_EVENT, row) = 0.0;
    g.at(NETSTIM_ISPIKE, row) = 0.0;
  }
}

void netstim_next_invl(Sim& s, Group& g, int row) {
  if (g.at(NETSTIM_NUMBER, row) > 0.0)
    g.at(NETSTIM_EVENT, row) = netstim_invl(s, g, row, g.at(NETSTIM_INTERVAL, row));
  if (g.at(NETSTIM_ISPIKE, row) >= g.at(NETSTIM_NUMBER, row)) g.at(NETSTIM_ON, row) = 0.0;
}

void netstim_receive(Sim& s, Group& g, int row, double t, int64_t edge, double w, double flag) {
  const int point = g.target[row] - 1;
  if (flag == 0.0) {
    if (w > 0.0 && g.at(NETSTIM_ON, row) == 0.0) {
      netstim_init_sequence(g, row);
      netstim_next_invl(s, g, row);
      g.at(NETSTIM_EVENT, row) =
          g.at(NETSTIM_EVENT, row) - g.at(NETSTIM_INTERVAL, row) * (1.0 - g.at(NETSTIM_NOISE, row));
      net_send(s, point, t, edge, g.at(NETSTIM_EVENT, row), 1.0);
    } else if (w < 0.0) {
      g.at(NETSTIM_ON, row) = 0.0;
    }
  }
  if (flag == 3.0) {
    if (g.at(NETSTIM_ON, row) == 1.0) {
      netstim_init_sequence(g, row);
      net_send(s, point, t, edge, 0.0, 1.0);
    }
  }
  if (flag == 1.0 && g.at(NETSTIM_ON, row) == 1.0) {
    g.at(NETSTIM_ISPIKE, row) = g.at(NETSTIM_ISPIKE, row) + 1.0;
    net_event(s, point, t, edge);
    netstim_next_invl(s, g, row);
    if (g.at(NETSTIM_ON, row) == 1.0) net_send(s, point, t, edge, g.at(NETSTIM_EVENT, row), 1.0);
  }
}

void netstim_initial(Sim& s, Group& g, int row) {
  rng_setseq(g.rng[row], s.seed, 0.0);
  g.at(NETSTIM_ON, row) = 0.0;
  g.at(NETSTIM_ISPIKE, row) = 0.0;
  if (g.at(NETSTIM_NOISE, row) < 0.0) g.at(NETSTIM_NOISE, row) = 0.0;
  if (g.at(NETSTIM_NOISE, row) > 1.0) g.at(NETSTIM_NOISE, row) = 1.0;
  if (g.at(NETSTIM_START, row) >= 0.0 && g.at(NETSTIM_NUMBER, row) > 0.0) {
    g.at(NETSTIM_ON, row) = 1.0;
    const double interval = g.at(NETSTIM_INTERVAL, row);
    double event = g.at(NETSTIM_START, row) + netstim_invl(s, g, row, interval) -
                   interval * (1.0 - g.at(NETSTIM_NOISE, row));
    if (event < 0.0) event = 0.0;
    g.at(NETSTIM_EVENT, row) = event;
    net_send(s, g.target[row] - 1, 0.0, 0, event, 3.0);
  }
}

void receive(Sim& s, int gi, int row, const Event& ev) {
  Group& g = s.groups[gi];
  const double t = ev.time, flag = ev.flag;
  const int64_t edge = ev.edge;
  const double w = edge == 0 ? 0.0 : s.edge_weight[edge - 1];
  if (s.log_deliveries) {
    s.log_time.push_back(t);
    s.log_order.push_back(ev.order);
    s.log_edge.push_back(ev.edge);
    s.log_target.push_back(g.target[row] - 1);
  }
  switch (g.mech) {
    case MIND_MECH_EXP2SYN: {
      const double f = g.at(EXP2SYN_FACTOR, row);
      g.at(EXP2SYN_A, row) = g.at(EXP2SYN_A, row) + w * f;
      g.at(EXP2SYN_B, row) = g.at(EXP2SYN_B, row) + w * f;
      break;
    }
    case MIND_MECH_EXPSYN: g.at(EXPSYN_G, row) = g.at(EXPSYN_G, row) + w; break;
    case MIND_MECH_NETSTIM: netstim_receive(s, g, row, t, edge, w, flag); break;
    default:
      if (g.mech >= MIND_MECH_COUNT) {
        ion_receive(s, g, row, false);
        g_user[g.mech].receive(g.data.data(), g.rows(), row, t, w);
        ion_receive(s, g, row, true);
      }
      break;
  }
}

void deliver(Sim& s, double cutoff) {  // deliver_events! (events.jl:57-165)
  std::vector<Event> pending;
  while (!s.heap.empty() && s.heap[0].time <= cutoff) {
    pending.clear();
    if (s.heap[0].target < 0) {  // explicit source events
      while (!s.heap.empty() && s.heap[0].time <= cutoff && s.heap[0].target < 0) {
        const Event e = pop_event(s);
        pending.push_back(e);
        route_spike(s, -e.target - 1, e.time);
      }
      std::stable_sort(pending.begin(), pending.end(), [](const Event& a, const Event& b) {
        return a.target < b.target || (a.target == b.target && a.time < b.time);
      });
      for (const Event& e : pending) record_spike(s, -e.target - 1, e.time);
      continue;
    }
    bool feedback = false;
    while (!s.heap.empty() && s.heap[0].time <= cutoff && s.heap[0].target > 0) {
      const int gi = s.point_group[s.heap[0].target - 1];
      const bool needs = s.feedback_group[gi] != 0;
      if (needs && !pending.empty()) break;
      pending.push_back(pop_event(s));
      feedback = needs;
      if (feedback) break;
    }
    std::stable_sort(pending.begin(), pending.end(), [&](const Event& a, const Event& b) {
      const int ga = s.point_group[a.target - 1], gb = s.point_group[b.target - 1];
      const int ra = s.point_row[a.target - 1], rb = s.point_row[b.target - 1];
      return ga < gb || (ga == gb && ra < rb);
    });
    for (const Event& e : pending)
      receive(s, s.point_group[e.target - 1], s.point_row[e.target - 1], e);
    if (feedback) flush(s);
  }
}

// ----------------------------------------------------------- Membrane -------
// VClamp.mod PROCEDURE clamp: the amplifier equations eliminated for e, vo, vi.
void vclamp_clamp(Group& g, int r, double v, double dt, double stim, double e0, double vo0, double& e,
                  double& vo, double& vi) {
  const double t1 = g.at(VCLAMP_TAU1, r) / dt;
  const double t2 = g.at(VCLAMP_TAU2, r) / dt;
  const double fac = g.at(VCLAMP_FAC, r);
  const double gain = g.at(VCLAMP_GAIN, r);
  const double num = (1.0 + t1) * ((1.0 - fac) * v + fac * t2 * vo0 / (1.0 + t2)) - stim -
                     t1 * (g.at(VCLAMP_VI0, r) - e0);
  e = num / (2.0 + t1 + (1.0 + t1) * fac * gain / (1.0 + t2));
  vo = (t2 * vo0 - gain * e) / (1.0 + t2);
  vi = v + fac * vo - fac * v;
}

// The membrane phases below work on a range of cells [c0, c1): a node's rhs and
// diag receive contributions only from rows and nodes of its own cell, so cells
// may run on separate threads and every node still sees the canonical order.
void currents(Sim& s, double t, int64_t c0, int64_t c1) {  // micro_currents! (step.jl:208-224)
  std::fill(s.rhs.begin() + s.cell_offset[c0], s.rhs.begin() + s.cell_offset[c1], 0.0);
  std::fill(s.diag.begin() + s.cell_offset[c0], s.diag.begin() + s.cell_offset[c1], 0.0);
  for (int64_t q = s.ion_off[s.cell_offset[c0]]; q < s.ion_off[s.cell_offset[c1]]; ++q) ion_prepare(s, q, 0x40);
  for (int pass = 0; pass < 2; ++pass) {  // density groups, then point groups
    for (Group& g : s.groups) {
      if (is_point(g.mech) != (pass == 1)) continue;
      if (g.mech == MIND_MECH_NETSTIM) continue;
      for (int64_t r = g.cell_rows[c0]; r < g.cell_rows[c1]; ++r) {
        const int32_t node = g.node[r];
        if (node < 0) continue;
        const double v = s.v[node];
        double shifted = 0.0, value = 0.0;
        // written ion currents at v + 0.001 (ish[k]); compiled phases also write
        // them at v (ish[8 + k])
        double ish[16] = {0};
        const bool ions = !ion_decls(g.mech).empty();
        if (ions) ion_load(s, g, (int)r);
        if (g.mech >= MIND_MECH_COUNT) {  // both probes inside the compiled phase
          g_user[g.mech].current(g.data.data(), g.rows(), r, v, t, s.dt, s.celsius, &value, &shifted, 1, ish);
          const double scale = is_point(g.mech) ? 1.0 : s.area[node] * 0.01;
          s.rhs[node] -= scale * value;
          s.diag[node] += scale * (shifted - value) / 0.001;
          if (ions) ion_current_writes(s, g, (int)r, ish);
          continue;
        }
        for (int probe = 0; probe < 2; ++probe) {
          const double vv = probe == 0 ? v + 0.001 : v;
          double sum = 0.0;
          switch (g.mech) {
            case MIND_MECH_HH: {
              const double gna = g.at(HH_GNABAR, r) * g.at(HH_M, r) * g.at(HH_M, r) *
                                 g.at(HH_M, r) * g.at(HH_H, r);
              const double ina = gna * (vv - g.at(HH_ENA, r));
              const double gk = g.at(HH_GKBAR, r) * g.at(HH_N, r) * g.at(HH_N, r) *
                                g.at(HH_N, r) * g.at(HH_N, r);
              const double ik = gk * (vv - g.at(HH_EK, r));
              const double il = g.at(HH_GL, r) * (vv - g.at(HH_EL, r));
              g.at(HH_GNA, r) = gna; g.at(HH_INA, r) = ina;
              g.at(HH_GK, r) = gk; g.at(HH_IK, r) = ik; g.at(HH_IL, r) = il;
              sum = 0.0; sum += il; sum += ina; sum += ik;  // currents, then ion writes
              if (probe == 0) {
                ish[0] = ina;
                ish[1] = ik;
              }
              break;
            }
            case MIND_MECH_PAS: {
              const double i = g.at(PAS_G, r) * (vv - g.at(PAS_E, r));
              g.at(PAS_I, r) = i;
              sum = 0.0; sum += i;
              break;
            }
            case MIND_MECH_EXPSYN: {
              const double i = g.at(EXPSYN_G, r) * (vv - g.at(EXPSYN_E, r));
              g.at(EXPSYN_I, r) = i;
              sum = 0.0; sum += i;
              break;
            }
            case MIND_MECH_EXP2SYN: {
              const double gg = g.at(EXP2SYN_B, r) - g.at(EXP2SYN_A, r);
              const double i = gg * (vv - g.at(EXP2SYN_E, r));
              g.at(EXP2SYN_G, r) = gg; g.at(EXP2SYN_I, r) = i;
              sum = 0.0; sum += i;
              break;
            }
            case MIND_MECH_ICLAMP: {
              const double del = g.at(ICLAMP_DEL, r), dur = g.at(ICLAMP_DUR, r);
              const double i = (t < del + dur && t >= del) ? g.at(ICLAMP_AMP, r) : 0.0;
              g.at(ICLAMP_I, r) = i;
              sum = 0.0; sum -= i;  // electrode current
              break;
            }
            case MIND_MECH_ALPHASYNAPSE: {  // AlphaSynapse.mod: g = gmax alpha((t - onset)/tau)
              const double x = (t - g.at(ALPHASYNAPSE_ONSET, r)) / g.at(ALPHASYNAPSE_TAU, r);
              const double a = (x < 0.0 || x > 10.0) ? 0.0 : x * mind_exp(1.0 - x);  // FUNCTION alpha
              const double gg = g.at(ALPHASYNAPSE_GMAX, r) * a;
              const double i = gg * (vv - g.at(ALPHASYNAPSE_E, r));
              g.at(ALPHASYNAPSE_G, r) = gg; g.at(ALPHASYNAPSE_I, r) = i;
              sum = 0.0; sum += i;
              break;
            }
            case MIND_MECH_VCLAMP: {  // VClamp.mod vstim(), then i = icur()
              double tc = 0.0, stim = g.at(VCLAMP_STIM, r);
              for (int j = 0; j < 3; ++j) {  // the first level whose end lies after t
                stim = g.at(VCLAMP_AMP + j, r);
                tc = tc + g.at(VCLAMP_DUR + j, r);
                if (t < tc) {
                  tc = tc + 100.0;
                  break;
                }
              }
              double e0 = g.at(VCLAMP_E0, r), vo0 = g.at(VCLAMP_VO0, r);
              double e = g.at(VCLAMP_E, r), vo = g.at(VCLAMP_VO, r), vi = g.at(VCLAMP_VI, r), i = 0.0;
              if (t > tc) {
                e0 = 0.0;
                vo0 = 0.0;
              } else {
                vclamp_clamp(g, r, vv, s.dt, stim, e0, vo0, e, vo, vi);
                i = (vo - vv) / g.at(VCLAMP_RSTIM, r);
              }
              g.at(VCLAMP_TC, r) = tc; g.at(VCLAMP_STIM, r) = stim; g.at(VCLAMP_I, r) = i;
              g.at(VCLAMP_E0, r) = e0; g.at(VCLAMP_VO0, r) = vo0;
              g.at(VCLAMP_E, r) = e; g.at(VCLAMP_VO, r) = vo; g.at(VCLAMP_VI, r) = vi;
              sum = 0.0; sum -= i;  // electrode current
              break;
            }
            case MIND_MECH_SECLAMP: {  // SEClamp.mod vstim() and icur(); at_time has no effect
              double on = 1.0, vc;
              if (t < g.at(SECLAMP_DUR1, r)) vc = g.at(SECLAMP_AMP1, r);
              else if (t < g.at(SECLAMP_TC2, r)) vc = g.at(SECLAMP_AMP2, r);
              else if (t < g.at(SECLAMP_TC3, r)) vc = g.at(SECLAMP_AMP3, r);
              else { vc = 0.0; on = 0.0; }
              double i = 0.0;
              if (on != 0.0) i = (vc - vv) / g.at(SECLAMP_RS, r);
              g.at(SECLAMP_ON, r) = on; g.at(SECLAMP_VC, r) = vc; g.at(SECLAMP_I, r) = i;
              sum = 0.0; sum -= i;  // electrode current
              break;
            }
          }
          if (probe == 0) shifted = sum; else value = sum;
        }
        const double scale = is_point(g.mech) ? 1.0 : s.area[node] * 0.01;
        s.rhs[node] -= scale * value;
        s.diag[node] += scale * (shifted - value) / 0.001;
        if (ions) ion_current_writes(s, g, (int)r, ish);
      }
    }
  }
}

// assemble_cable_kernel!, solve_dhs_kernel!, update_voltage_kernel!; returns
// MIND_ESINGULAR on a zero or non-finite pivot
int cable(Sim& s, int64_t c0, int64_t c1) {
  // assembly at dt, or dt/2 with secondorder (step.jl:112)
  const double dt = s.secondorder ? 0.5 * s.dt : s.dt;
  const int64_t a = s.cell_offset[c0], n = s.cell_offset[c1];
  int err = 0;
  for (int64_t i = a; i < n; ++i) {  // each node pulls its own and its children's currents
    s.diag[i] = s.cap[i] / dt + s.diag[i] + s.axial_diag[i];
    double b = s.rhs[i];
    if (s.parent[i] >= 0) b += s.axial[i] * (s.v[s.parent[i]] - s.v[i]);
    for (int32_t c : s.children[i]) b -= s.axial[c] * (s.v[i] - s.v[c]);
    s.rhs[i] = b;
  }
  for (int64_t i = n - 1; i >= a; --i) {  // triang: children before parents
    const double pivot = s.diag[i];
    if (pivot == 0.0 || !mind_isfinite(pivot)) err = MIND_ESINGULAR;
    const int32_t p = s.parent[i];
    if (p >= 0) {
      const double factor = s.axial[i] / pivot;
      s.diag[p] -= factor * s.axial[i];
      s.rhs[p] += factor * s.rhs[i];
    }
  }
  for (int64_t i = a; i < n; ++i) {  // bksub: parents before children
    double value = s.rhs[i];
    const int32_t p = s.parent[i];
    if (p >= 0) value += s.axial[i] * s.rhs[p];
    s.rhs[i] = value / s.diag[i];
  }
  // v += dv, or 2 dv with secondorder (update_voltage_kernel!, cuda/micro.jl:660-679; the
  // secondorder 2 ion-current correction writes per-node ion current columns that none of
  // the six mechanisms reads)
  if (s.secondorder == 2)  // ion currents follow the increment (update_voltage_kernel!)
    for (int64_t i = a; i < n; ++i)
      for (int64_t q = s.ion_off[i]; q < s.ion_off[i + 1]; ++q) s.ion[3][q] += s.ion[4][q] * s.rhs[i];
  if (s.secondorder) {
    for (int64_t i = a; i < n; ++i) s.v[i] += 2.0 * s.rhs[i];
  } else {
    for (int64_t i = a; i < n; ++i) s.v[i] += s.rhs[i];
  }
  return err;
}


void advance_states(Sim& s, int64_t c0, int64_t c1) {  // nrn_state at t + dt (step.jl:140)
  const double dt = s.dt;
  for (Group& g : s.groups) {
    const int64_t r0 = g.cell_rows[c0], rows = g.cell_rows[c1];
    if (g.mech >= MIND_MECH_COUNT) {
      if (!g_user[g.mech].has_state) continue;
      const double t = (double)s.step * dt + dt;
      for (int64_t r = r0; r < rows; ++r) {
        ion_load(s, g, (int)r);
        g_user[g.mech].advance(g.data.data(), g.rows(), r, s.v[g.node[r]], t, dt, s.celsius);
        ion_store(s, g, (int)r);
      }
    } else if (g.mech == MIND_MECH_HH) {
      for (int64_t r = r0; r < rows; ++r) {
        ion_load(s, g, (int)r);
        double f[HH_NFIELDS];
        for (int k = 0; k < HH_NFIELDS; ++k) f[k] = g.at(k, r);
        hh_rates(f, s.v[g.node[r]], s.celsius);
        f[HH_M] = relax(f[HH_M], f[HH_MINF], f[HH_MTAU], dt);
        f[HH_H] = relax(f[HH_H], f[HH_HINF], f[HH_HTAU], dt);
        f[HH_N] = relax(f[HH_N], f[HH_NINF], f[HH_NTAU], dt);
        for (int k = 0; k < HH_NFIELDS; ++k) g.at(k, r) = f[k];
        ion_store(s, g, (int)r);
      }
    } else if (g.mech == MIND_MECH_EXP2SYN) {
      for (int64_t r = r0; r < rows; ++r) {
        g.at(EXP2SYN_A, r) = relax(g.at(EXP2SYN_A, r), 0.0, g.at(EXP2SYN_TAU1, r), dt);
        g.at(EXP2SYN_B, r) = relax(g.at(EXP2SYN_B, r), 0.0, g.at(EXP2SYN_TAU2, r), dt);
      }
    } else if (g.mech == MIND_MECH_EXPSYN) {
      for (int64_t r = r0; r < rows; ++r)
        g.at(EXPSYN_G, r) = relax(g.at(EXPSYN_G, r), 0.0, g.at(EXPSYN_TAU, r), dt);
    } else if (g.mech == MIND_MECH_VCLAMP) {  // SOLVE update METHOD after_cvode (VClamp.mod update)
      const double t = (double)s.step * dt + dt;  // time + dt
      for (int64_t r = r0; r < rows; ++r) {
        const double v = s.v[g.node[r]];
        double i = 0.0;  // FUNCTION icur
        if (t > g.at(VCLAMP_TC, r)) {
          g.at(VCLAMP_E0, r) = 0.0;
          g.at(VCLAMP_VO0, r) = 0.0;
        } else {
          vclamp_clamp(g, r, v, dt, g.at(VCLAMP_STIM, r), g.at(VCLAMP_E0, r), g.at(VCLAMP_VO0, r),
                       g.at(VCLAMP_E, r), g.at(VCLAMP_VO, r), g.at(VCLAMP_VI, r));
          i = (g.at(VCLAMP_VO, r) - v) / g.at(VCLAMP_RSTIM, r);
        }
        g.at(VCLAMP_I, r) = i;
        g.at(VCLAMP_E0, r) = g.at(VCLAMP_E, r);
        g.at(VCLAMP_VO0, r) = g.at(VCLAMP_VO, r);
        g.at(VCLAMP_VI0, r) = g.at(VCLAMP_VI, r);
      }
    } else if (g.mech == MIND_MECH_SECLAMP) {  // SOLVE icur METHOD after_cvode: i at the new v
      for (int64_t r = r0; r < rows; ++r) {
        double i = 0.0;
        if (g.at(SECLAMP_ON, r) != 0.0) i = (g.at(SECLAMP_VC, r) - s.v[g.node[r]]) / g.at(SECLAMP_RS, r);
        g.at(SECLAMP_I, r) = i;
      }
    }
  }
}

void detect(Sim& s, double time) {  // detect_spikes_kernel! (cuda/micro.jl:285-302)
  for (size_t d = 0; d < s.det_node.size(); ++d) {
    const bool current = s.v[s.det_node[d]] > s.det_threshold[d];
    if (current && !s.det_above[d]) {
      const int32_t source = s.det_source[d];
      record_spike(s, source, time);
      s.outbox.push_back(Event{time, 0.0, 0, 0, -(source + 1)});
    }
    s.det_above[d] = current;
  }
}

// currents at t + dt/2, cable, nrn_state for all cells; cells are independent
// in these phases, so they are split over s.threads OpenMP threads (chunks of
// consecutive cells), with results identical to one thread
void membrane(Sim& s, double t_mid) {
  const int64_t ncells = (int64_t)s.cell_offset.size() - 1;
  const int64_t chunk = 16;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(s.threads) reduction(max : err) if (s.threads > 1)
  for (int64_t c0 = 0; c0 < ncells; c0 += chunk) {
    const int64_t c1 = std::min(ncells, c0 + chunk);
    currents(s, t_mid, c0, c1);
    err = std::max(err, cable(s, c0, c1));
    advance_states(s, c0, c1);
  }
  if (err && s.error < err) s.error = err;
}

void advance_micro(Sim& s) {  // advance_micro! (step.jl:91-151)
  const double time = (double)s.step * s.dt;
  const double dt = s.dt;
  detect(s, time + 1.0e-10);
  flush(s);
  deliver(s, time + 0.5 * dt);
  membrane(s, time + 0.5 * dt);
  s.step += 1;
}

void initial_blocks(Sim& s) {  // INITIAL per group (initialize.jl:342-374)
  for (size_t gi = 0; gi < s.groups.size(); ++gi) {
    Group& g = s.groups[gi];
    const std::vector<IonDecl>& decl = ion_decls(g.mech);
    for (int r = 0; r < g.rows(); ++r) {
      if (!decl.empty()) ion_load(s, g, r);
      if (g.mech >= MIND_MECH_COUNT) {
        g_user[g.mech].initial(g.data.data(), g.rows(), r, s.v[g.node[r]], s.celsius);
        ion_store(s, g, r);
        continue;
      }
      switch (g.mech) {
        case MIND_MECH_HH: {
          double f[HH_NFIELDS];
          for (int k = 0; k < HH_NFIELDS; ++k) f[k] = g.at(k, r);
          hh_rates(f, s.v[g.node[r]], s.celsius);
          f[HH_M] = f[HH_MINF]; f[HH_H] = f[HH_HINF]; f[HH_N] = f[HH_NINF];
          for (int k = 0; k < HH_NFIELDS; ++k) g.at(k, r) = f[k];
          break;
        }
        case MIND_MECH_EXP2SYN: {
          double tau1 = g.at(EXP2SYN_TAU1, r);
          const double tau2 = g.at(EXP2SYN_TAU2, r);
          if (tau1 / tau2 > 0.9999) tau1 = 0.9999 * tau2;
          if (tau1 / tau2 < 1e-9) tau1 = tau2 * 1e-9;
          g.at(EXP2SYN_TAU1, r) = tau1;
          g.at(EXP2SYN_A, r) = 0.0;
          g.at(EXP2SYN_B, r) = 0.0;
          const double tp = (tau1 * tau2) / (tau2 - tau1) * mind_log(tau2 / tau1);
          double factor = -mind_exp(-tp / tau1) + mind_exp(-tp / tau2);
          factor = 1.0 / factor;
          g.at(EXP2SYN_FACTOR, r) = factor;
          break;
        }
        case MIND_MECH_EXPSYN: g.at(EXPSYN_G, r) = 0.0; break;
        case MIND_MECH_ICLAMP: g.at(ICLAMP_I, r) = 0.0; break;
        case MIND_MECH_VCLAMP: {  // VClamp.mod INITIAL
          const double v = s.v[g.node[r]];
          g.at(VCLAMP_E0, r) = 0.0;
          g.at(VCLAMP_VO, r) = v;
          g.at(VCLAMP_VO0, r) = v;
          g.at(VCLAMP_VI, r) = v;
          g.at(VCLAMP_VI0, r) = v;
          g.at(VCLAMP_E, r) = 0.0;
          break;
        }
        case MIND_MECH_SECLAMP:  // SEClamp.mod INITIAL
          g.at(SECLAMP_TC2, r) = g.at(SECLAMP_DUR1, r) + g.at(SECLAMP_DUR2, r);
          g.at(SECLAMP_TC3, r) = g.at(SECLAMP_TC2, r) + g.at(SECLAMP_DUR3, r);
          g.at(SECLAMP_ON, r) = 0.0;
          break;
        case MIND_MECH_NETSTIM: netstim_initial(s, g, r); break;
        default: break;
      }
      if (!decl.empty()) ion_store(s, g, r);
    }
    if (g.mech == MIND_MECH_NETSTIM) flush(s);
    // initialize_concentration_reversals_kernel! (cuda/micro.jl:764-784)
    for (size_t k = 0; k < decl.size(); ++k) {
      if (decl[k].write[1] < 0 && decl[k].write[2] < 0) continue;
      for (int r = 0; r < g.rows(); ++r) {
        const int32_t q = g.ion_row[(size_t)r * decl.size() + k];
        if (q >= 0 && (s.ion_style[q] & 0x20))
          s.ion[0][q] = nernst(s.ion[1][q], s.ion[2][q], s.species[s.ion_sp[q]].z, s.celsius);
      }
    }
  }
}

}  // namespace

// ------------------------------------------------------------------ API -----
extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

void oracle_free(void* p) { delete static_cast<Sim*>(p); }

// Registers the host build of a runtime-compiled mechanism's phases under its
// slot id (before oracle_build of a model that uses it).
int oracle_register_mech(int32_t mech, int32_t nfields, int32_t point, int32_t has_state, const double* defaults,
                         void* current, void* advance, void* initial, void* receive) {
  if (mech < MIND_MECH_COUNT || mech >= MIND_MECH_SLOTS || nfields < 0) return fail(MIND_EARG, "bad mechanism slot");
  UserMech& u = g_user[mech];
  u.nfields = nfields;
  u.point = point != 0;
  u.has_state = has_state != 0;
  u.defaults.assign(defaults, defaults + nfields);
  u.current = (UserCurrent)current;
  u.advance = (UserAdvance)advance;
  u.initial = (UserInitial)initial;
  u.receive = (UserReceive)receive;
  return 0;
}

// The USEION statements of a registered mechanism: per statement its ion name,
// VALENCE and the read / write field of each ion column (4 per statement); the
// STATE flag of each column and the columns NET_RECEIVE touches.
int oracle_register_ions(int32_t mech, int32_t n, const char* const* names, const int32_t* has_valence,
                         const double* valence, const int32_t* reads, const int32_t* writes, const uint8_t* state_col,
                         int32_t nrecv, const int32_t* recv_fields) {
  if (mech < MIND_MECH_COUNT || mech >= MIND_MECH_SLOTS) return fail(MIND_EARG, "bad mechanism slot");
  UserMech& u = g_user[mech];
  u.ions.clear();
  for (int32_t k = 0; k < n; ++k) {
    IonDecl decl;
    decl.name = names[k];
    decl.has_valence = has_valence[k] != 0;
    decl.valence = valence[k];
    for (int c = 0; c < 4; ++c) {
      decl.read[c] = reads[4 * k + c];
      decl.write[c] = writes[4 * k + c];
    }
    u.ions.push_back(decl);
  }
  u.state_col.assign(state_col, state_col + u.nfields);
  u.recv_fields.assign(recv_fields, recv_fields + nrecv);
  return 0;
}

// inject_event! (state.jl:618-636): the connection's event at `time` without its
// delay, enqueued (ticketed) now.
int oracle_inject_event(void* p, int64_t connection, double time) {
  Sim& s = *static_cast<Sim*>(p);
  if (connection < 0 || connection >= (int64_t)s.connection_edge.size()) return fail(MIND_EARG, "unknown connection");
  if (!(mind_isfinite(time) && time >= (double)s.step * s.dt))
    return fail(MIND_EARG, "event arrival precedes the current time");
  const int64_t e = s.connection_edge[connection];
  enqueue(s, Event{time, 0.0, e + 1, e + 1, s.edge_target[e]});
  return 0;
}

// inject_spike! (state.jl:638-659): a source event of an external input at `time`.
int oracle_inject_spike(void* p, int64_t source, double time) {
  Sim& s = *static_cast<Sim*>(p);
  if (std::find(s.source_external.begin(), s.source_external.end(), (int32_t)source) == s.source_external.end())
    return fail(MIND_EARG, "source is not a built external spike input");
  if (!(mind_isfinite(time) && time >= (double)s.step * s.dt))
    return fail(MIND_EARG, "spike time precedes the current time");
  enqueue(s, Event{time, 0.0, 0, 0, -(int32_t)(source + 1)});
  return 0;
}

// OpenMP threads of the membrane phases (currents, cable, nrn_state); the event
// queue stays serial. Results do not depend on the count.
int oracle_set_threads(void* p, int32_t threads) {
  if (threads < 1) return fail(MIND_EARG, "threads must be positive");
  static_cast<Sim*>(p)->threads = threads;
  return 0;
}

// build!(...; secondorder) (build.jl:120-131); before oracle_initialize.
int oracle_set_secondorder(void* p, int32_t secondorder) {
  if (secondorder < 0 || secondorder > 2) return fail(MIND_EARG, "secondorder must be 0, 1 or 2");
  static_cast<Sim*>(p)->secondorder = secondorder;
  return 0;
}

// Build (pack_micro, model.jl:420-902, in canonical per-cell contiguous order).
void* oracle_build(const mind_model_desc* d, double dt, uint64_t seed) {
  Sim* sp = new Sim();
  Sim& s = *sp;
  s.dt = dt;
  s.seed = seed;
  s.celsius = d->celsius;
  // 1. templates
  for (int m = 0; m < d->n_morphologies; ++m) {
    std::vector<SectionIn> secs;
    for (int64_t i = d->morph_offsets[m]; i < d->morph_offsets[m + 1]; ++i) {
      SectionIn si{d->sec_parent[i], d->sec_parent_x[i], d->sec_child_x[i], d->sec_nseg[i],
                   d->sec_length[i], d->sec_diameter[i], d->sec_Ra[i], d->sec_cm[i], {}};
      if (d->sec_point_offsets)
        for (int64_t p = d->sec_point_offsets[i]; p < d->sec_point_offsets[i + 1]; ++p)
          si.points.push_back(Pt{d->points[4 * p], d->points[4 * p + 1], d->points[4 * p + 2],
                                 d->points[4 * p + 3]});
      secs.push_back(si);
    }
    s.morph_secs.push_back(secs);
    s.templates.push_back(discretize(secs));
  }
  // 2. cells, contiguous in cell order
  s.cell_offset.assign(d->n_cells + 1, 0);
  for (int c = 0; c < d->n_cells; ++c)
    s.cell_offset[c + 1] = s.cell_offset[c] + (int64_t)s.templates[d->cell_morphology[c]].parent.size();
  s.total = s.cell_offset[d->n_cells];
  s.parent.resize(s.total); s.area.resize(s.total); s.cap.resize(s.total);
  s.axial.resize(s.total); s.v_init.resize(s.total);
  s.cell_morph.assign(d->cell_morphology, d->cell_morphology + d->n_cells);
  for (int c = 0; c < d->n_cells; ++c) {
    const Template& t = s.templates[d->cell_morphology[c]];
    const int64_t off = s.cell_offset[c];
    for (size_t i = 0; i < t.parent.size(); ++i) {
      s.parent[off + i] = t.parent[i] < 0 ? -1 : (int32_t)(off + t.parent[i]);
      s.area[off + i] = t.area[i];
      s.cap[off + i] = t.cap[i];
      s.axial[off + i] = t.axial[i];
      s.v_init[off + i] = d->cell_voltage[c];
    }
  }
  for (int c = 0; c < d->n_cells; ++c) {
    const Template& t = s.templates[d->cell_morphology[c]];
    for (const std::vector<int32_t>& nodes : t.sec_nodes) s.compartments += (int64_t)nodes.size() - 1;
  }
  s.children.assign(s.total, {});
  for (int64_t i = 0; i < s.total; ++i)
    if (s.parent[i] >= 0) s.children[s.parent[i]].push_back((int32_t)i);
  // axial_diagonal in template order (model.jl:857-863)
  s.axial_diag.assign(s.total, 0.0);
  for (int64_t i = 0; i < s.total; ++i) {
    if (s.parent[i] < 0) continue;
    s.axial_diag[i] += s.axial[i];
    s.axial_diag[s.parent[i]] += s.axial[i];
  }
  auto cell_node = [&](int cell, int sec, double x, bool range) -> int64_t {
    const Template& t = s.templates[d->cell_morphology[cell]];
    return s.cell_offset[cell] + (range ? range_node(t, sec, x) : location_node(t, sec, x));
  };
  // 3. mechanism groups in first-use order: density insertions, then points
  s.group_of_mech.assign(MIND_MECH_SLOTS, -1);
  auto group_for = [&](int mech) {
    if (s.group_of_mech[mech] < 0) {
      s.group_of_mech[mech] = (int)s.groups.size();
      Group g;
      g.mech = mech;
      s.groups.push_back(g);
    }
    return s.group_of_mech[mech];
  };
  struct RowInit { int64_t poff0, poff1; };
  std::vector<std::vector<RowInit>> row_params;
  std::vector<uint8_t> has_density(MIND_MECH_SLOTS * (size_t)s.total, 0);
  for (int64_t k = 0; k < d->n_density; ++k) {
    const int mech = d->dens_mech[k];
    const int gi = group_for(mech);
    Group& g = s.groups[gi];
    if ((int)row_params.size() <= gi) row_params.resize(gi + 1);
    const Template& t = s.templates[d->cell_morphology[d->dens_cell[k]]];
    const std::vector<int32_t>& nodes = t.sec_nodes[d->dens_section[k]];
    for (size_t seg = 0; seg + 1 < nodes.size(); ++seg) {
      const int64_t node = s.cell_offset[d->dens_cell[k]] + nodes[seg];
      if (has_density[(size_t)mech * s.total + node]) continue;  // insert! keeps the first
      has_density[(size_t)mech * s.total + node] = 1;
      g.node.push_back((int32_t)node);
      g.target.push_back(0);
      g.logical.push_back((uint64_t)node + 1);
      row_params[gi].push_back(RowInit{d->dens_param_offsets[k], d->dens_param_offsets[k + 1]});
    }
  }
  s.point_group.resize(d->n_points);
  s.point_row.resize(d->n_points);
  std::vector<std::pair<double, int32_t>> pattern;  // PatternStim (time, point)
  for (int64_t p = 0; p < d->n_points; ++p) {
    if (d->pt_mech[p] == MIND_MECH_PATTERNSTIM) {  // native: its spike times, no mechanism data
      s.point_group[p] = -1;
      s.point_row[p] = -1;
      if (d->pt_spike_offsets)
        for (int64_t k = d->pt_spike_offsets[p]; k < d->pt_spike_offsets[p + 1]; ++k)
          pattern.push_back({d->pt_spike_times[k], (int32_t)p});
      continue;
    }
    const int gi = group_for(d->pt_mech[p]);
    Group& g = s.groups[gi];
    if ((int)row_params.size() <= gi) row_params.resize(gi + 1);
    const int64_t node = d->pt_cell[p] < 0 ? -1 : cell_node(d->pt_cell[p], d->pt_section[p], d->pt_x[p], false);
    g.node.push_back((int32_t)node);
    g.target.push_back((int32_t)p + 1);
    g.logical.push_back((uint64_t)(p + 1) + (uint64_t)s.total);
    s.point_group[p] = gi;
    s.point_row[p] = g.rows() - 1;
    row_params[gi].push_back(RowInit{d->pt_param_offsets[p], d->pt_param_offsets[p + 1]});
  }
  // rows sorted by node (stable) except artificial cells, which keep logical order
  struct IonParam { size_t gi; int row, field; double value; };
  std::vector<IonParam> ion_params;  // parameters of mechanism fields, in group and row order
  for (size_t gi = 0; gi < s.groups.size(); ++gi) {
    Group& g = s.groups[gi];
    const int rows = g.rows();
    std::vector<int> order(rows);
    for (int r = 0; r < rows; ++r) order[r] = r;
    if (g.mech != MIND_MECH_NETSTIM)
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return g.node[a] < g.node[b]; });
    Group h;
    h.mech = g.mech;
    std::vector<RowInit> rp(rows);
    std::vector<int> inverse(rows);
    for (int r = 0; r < rows; ++r) {
      h.node.push_back(g.node[order[r]]);
      h.target.push_back(g.target[order[r]]);
      h.logical.push_back(g.logical[order[r]]);
      rp[r] = row_params[gi][order[r]];
      inverse[order[r]] = r;
    }
    for (int r = 0; r < rows; ++r)
      if (h.target[r] > 0) s.point_row[h.target[r] - 1] = r;
    const int nf = field_count(g.mech);
    h.data.assign((size_t)nf * rows, 0.0);
    std::vector<double> def(std::max(nf, 32), 0.0);
    defaults(g.mech, def.data());
    for (int f = 0; f < nf; ++f)
      for (int r = 0; r < rows; ++r) h.data[(size_t)f * rows + r] = def[f];
    for (int r = 0; r < rows; ++r)
      for (int64_t q = rp[r].poff0; q < rp[r].poff1; ++q) {
        h.data[(size_t)d->param_field[q] * rows + r] = d->param_value[q];
        ion_params.push_back({gi, r, d->param_field[q], d->param_value[q]});
      }
    h.initial = h.data;
    h.rng.resize(rows);
    h.cell_rows.assign(d->n_cells + 1, 0);  // NetStim rows (logical order) belong to no cell phase
    if (g.mech != MIND_MECH_NETSTIM)
      for (int c = 0; c <= d->n_cells; ++c)  // rows are sorted by node (artificial rows, -1, first)
        h.cell_rows[c] = std::lower_bound(h.node.begin(), h.node.end(), (int32_t)s.cell_offset[c]) - h.node.begin();
    s.groups[gi] = h;
  }
  // 3b. ion species, rows, values and styles (model.jl:672-855)
  if (d->n_species == 0) {
    s.species = {{"na", 1.0, 10.0, 140.0, 50.0, false},
                 {"k", 1.0, 54.4, 2.5, -77.0, false},
                 {"ca", 2.0, 5.0e-5, 2.0, 132.4579341637009, false}};
  } else {
    for (int32_t k = 0; k < d->n_species; ++k)
      s.species.push_back({d->species_name[k], d->species_valence[k], d->species_internal[k],
                           d->species_external[k], d->species_reversal[k], d->species_nernst[k] != 0});
  }
  auto species_index = [&](const std::string& name) {
    for (size_t k = 0; k < s.species.size(); ++k)
      if (s.species[k].name == name) return (int)k;
    return -1;
  };
  for (Group& g : s.groups)
    for (const IonDecl& u : ion_decls(g.mech)) {
      int k = species_index(u.name);
      if (k < 0) {
        if (!u.has_valence) {
          delete sp;
          fail(MIND_EARG, "USEION " + u.name + " needs a VALENCE declaration");
          return nullptr;
        }
        s.species.push_back({u.name, u.valence, 1.0, 1.0, 0.0, false});
        k = (int)s.species.size() - 1;
      }
      g.slot_species.push_back(k);
    }
  auto section_segments = [&](int cell, int sec) {
    std::vector<int64_t> out;
    const std::vector<int32_t>& nodes = s.templates[d->cell_morphology[cell]].sec_nodes[sec];
    for (size_t j = 0; j + 1 < nodes.size(); ++j) out.push_back(s.cell_offset[cell] + nodes[j]);
    return out;
  };
  std::set<std::pair<int64_t, std::string>> used;
  for (Group& g : s.groups)
    for (int r = 0; r < g.rows(); ++r)
      if (g.node[r] >= 0)
        for (int k : g.slot_species) used.insert({g.node[r], s.species[k].name});
  for (int64_t p = 0; p < d->n_points; ++p)  // NEURON need_memb: the point's whole section
    if (d->pt_cell[p] >= 0 && s.point_group[p] >= 0)
      for (int k : s.groups[s.point_group[p]].slot_species)
        for (int64_t n : section_segments(d->pt_cell[p], d->pt_section[p])) used.insert({n, s.species[k].name});
  for (int64_t q = 0; q < d->n_ion_settings; ++q)
    for (int64_t n : section_segments(d->ionset_cell[q], d->ionset_section[q]))
      used.insert({n, s.species[d->ionset_species[q]].name});
  s.nion = (int64_t)used.size();
  s.ion_off.assign(s.total + 1, 0);
  for (int c = 0; c < 5; ++c) s.ion[c].assign(s.nion, 0.0);
  s.ion_style.assign(s.nion, 0);
  std::map<std::pair<int64_t, int>, int32_t> row_of;  // (node, species) -> ion row
  {
    int64_t q = 0;
    for (const auto& [node, name] : used) {
      const int k = species_index(name);
      const Sim::Species& sp_k = s.species[k];
      s.ion_sp.push_back(k);
      s.ion_off[node + 1] += 1;
      row_of[{node, k}] = (int32_t)q;
      s.ion[0][q] = sp_k.nernst ? nernst(sp_k.ci, sp_k.co, sp_k.z, s.celsius) : sp_k.e;
      s.ion[1][q] = sp_k.ci;
      s.ion[2][q] = sp_k.co;
      s.ion_style[q] = sp_k.nernst ? style_bits(1, 2, true, true, false) : 0;
      ++q;
    }
  }
  for (int64_t n = 0; n < s.total; ++n) s.ion_off[n + 1] += s.ion_off[n];
  for (Group& g : s.groups) {
    const size_t ns = g.slot_species.size();
    g.ion_row.assign((size_t)g.rows() * ns, -1);
    for (int r = 0; r < g.rows(); ++r)
      for (size_t k = 0; k < ns; ++k)
        if (g.node[r] >= 0) g.ion_row[(size_t)r * ns + k] = row_of.at({g.node[r], g.slot_species[k]});
  }
  // USEION values given as mechanism parameters go to the shared ion row (collect_ion_parameters!)
  for (const IonParam& ip : ion_params) {
    Group& g = s.groups[ip.gi];
    const std::vector<IonDecl>& decl = ion_decls(g.mech);
    for (size_t k = 0; k < decl.size(); ++k) {
      const int32_t q = g.ion_row[(size_t)ip.row * decl.size() + k];
      if (q < 0) continue;
      for (int c = 0; c < 4; ++c)
        if (decl[k].read[c] == ip.field || decl[k].write[c] == ip.field) s.ion[c][q] = ip.value;
    }
  }
  for (int64_t q = 0; q < d->n_ion_settings; ++q) {  // section ion values (cell_ion_values)
    const int what = d->ionset_what[q];
    if (what > 2) continue;
    const int column = what == 0 ? 1 : what == 1 ? 2 : 0;
    for (int64_t n : section_segments(d->ionset_cell[q], d->ionset_section[q]))
      s.ion[column][row_of.at({n, d->ionset_species[q]})] = d->ionset_value[q];
  }
  // styles: promotion by every instance's USEION, Nernst species, section overrides
  for (Group& g : s.groups) {
    const std::vector<IonDecl>& decl = ion_decls(g.mech);
    for (size_t k = 0; k < decl.size(); ++k) {
      const IonDecl& u = decl[k];
      const int c_style = (u.write[1] >= 0 || u.write[2] >= 0) ? 3 : (u.read[1] >= 0 || u.read[2] >= 0) ? 1 : 0;
      const int e_style = u.write[0] >= 0 ? 3 : u.read[0] >= 0 ? 1 : 0;
      for (int r = 0; r < g.rows(); ++r) {
        const int32_t q = g.ion_row[(size_t)r * decl.size() + k];
        if (q >= 0) s.ion_style[q] = promote(s.ion_style[q], c_style, e_style);
      }
    }
  }
  for (int64_t q = 0; q < s.nion; ++q)
    if (s.species[s.ion_sp[q]].nernst && ((s.ion_style[q] >> 3) & 3) != 3) s.ion_style[q] |= 0x60;
  {
    std::map<std::tuple<int, int, int>, std::pair<int, double>> last;
    for (int64_t q = 0; q < d->n_ion_settings; ++q)
      if (d->ionset_what[q] >= 3)
        last[{d->ionset_cell[q], d->ionset_section[q], d->ionset_species[q]}] = {d->ionset_what[q], d->ionset_value[q]};
    for (const auto& [key, o] : last)
      for (int64_t n : section_segments(std::get<0>(key), std::get<1>(key))) {
        const int32_t q = row_of.at({n, std::get<2>(key)});
        if (o.first == 4) {
          s.ion_style[q] = (uint8_t)o.second;
        } else if (o.second > 0) {  // policies fixed, initial, advance, dynamic
          const uint8_t flags[5] = {0, 0x00, 0x20, 0x40, 0x60};
          s.ion_style[q] = (uint8_t)((s.ion_style[q] & 0x1f) | flags[(int)o.second]);
        }
      }
  }
  for (int c = 0; c < 5; ++c) s.ion_packed[c] = s.ion[c];
  // 4. sources, detectors, fanout
  s.n_sources = d->n_sources;
  s.source_sid.assign(d->src_sid, d->src_sid + d->n_sources);
  s.source_target.assign(d->n_sources, 0);
  s.target_sources.assign(d->n_points, {});
  for (int64_t k = 0; k < d->n_sources; ++k) {
    if (d->src_kind[k] == 0) {
      s.det_node.push_back((int32_t)cell_node(d->src_cell[k], d->src_section[k], d->src_x[k], false));
      s.det_source.push_back((int32_t)k);
      s.det_threshold.push_back(d->src_threshold[k]);
    } else if (d->src_kind[k] == 2) {  // external input: driven by inject_spike! only
      s.source_external.push_back((int32_t)k);
    } else {
      s.source_target[k] = d->src_point[k] + 1;
      s.target_sources[d->src_point[k]].push_back((int32_t)k);
    }
  }
  s.det_above.assign(s.det_node.size(), 0);
  // PatternStim spikes: by time (stable over point creation), then the point's
  // sources in order (pack_micro, model.jl:604-613; reset_events!, utils/events.jl:92-110)
  std::stable_sort(pattern.begin(), pattern.end(),
                   [](const std::pair<double, int32_t>& a, const std::pair<double, int32_t>& b) {
                     return a.first < b.first;
                   });
  for (const auto& [t, p] : pattern)
    for (int64_t k = 0; k < d->n_sources; ++k)
      if (d->src_kind[k] == 1 && d->src_point[k] == p) {
        s.pattern_time.push_back(t);
        s.pattern_source.push_back((int32_t)k);
      }
  s.fanout.assign(d->n_sources + 1, 0);
  for (int64_t c = 0; c < d->n_connections; ++c) s.fanout[d->con_source[c] + 1] += 1;
  for (int64_t k = 0; k < d->n_sources; ++k) s.fanout[k + 1] += s.fanout[k];
  std::vector<int64_t> cursor(s.fanout.begin(), s.fanout.end() - 1);
  s.edge_target.resize(d->n_connections);
  s.edge_delay.resize(d->n_connections);
  s.edge_weight.resize(d->n_connections);
  s.connection_edge.resize(d->n_connections);
  for (int64_t c = 0; c < d->n_connections; ++c) {
    const int64_t e = cursor[d->con_source[c]]++;
    s.connection_edge[c] = e;
    s.edge_target[e] = d->con_target[c] + 1;
    s.edge_delay[e] = d->con_delay[c];
    s.edge_weight[e] = d->con_weight[c];
  }
  // 5. ranks (initialize.jl:629-649) and feedback receivers
  s.target_rank.assign(d->n_points, 0);
  int64_t rank = 0;
  s.feedback_group.assign(s.groups.size(), 0);
  for (size_t gi = 0; gi < s.groups.size(); ++gi) {
    Group& g = s.groups[gi];
    if (g.mech != MIND_MECH_NETSTIM) continue;
    s.feedback_group[gi] = 1;
    uint64_t max_id = 0;
    for (int r = 0; r < g.rows(); ++r) {
      if (g.target[r] > 0) s.target_rank[g.target[r] - 1] = rank + (int64_t)g.logical[r] + 1;
      max_id = std::max(max_id, g.logical[r]);
    }
    rank += (int64_t)max_id + 1;
  }
  s.rank_total = rank;
  s.source_rank.resize(d->n_sources);
  for (int64_t k = 0; k < d->n_sources; ++k)
    s.source_rank[k] = s.source_target[k] == 0 ? rank + k + 1 : s.target_rank[s.source_target[k] - 1];
  s.sequences.assign(d->n_points, 0);
  s.v = s.v_init;
  s.rhs.assign(s.total, 0.0);
  s.diag.assign(s.total, 0.0);
  return sp;
}

int oracle_initialize(void* p) {  // initialize! (initialize.jl:270-429)
  Sim& s = *static_cast<Sim*>(p);
  s.step = 0;
  s.v = s.v_init;
  s.error = 0;
  s.heap.clear();
  s.outbox.clear();
  s.next_order = 0;
  for (size_t k = 0; k < s.pattern_time.size(); ++k)  // reset_events!: pattern spikes first
    enqueue(s, Event{s.pattern_time[k], 0.0, 0, 0, -(s.pattern_source[k] + 1)});
  std::fill(s.sequences.begin(), s.sequences.end(), 0);
  std::fill(s.det_above.begin(), s.det_above.end(), 0);
  s.spike_t.clear();
  s.spike_src.clear();
  // Spikes emitted while initializing stay in the reference's macro spike
  // buffer and are recorded by the first run! (step.jl:47-49).
  s.recording_spikes = true;
  for (Group& g : s.groups) {
    g.data = g.initial;
    if (g.mech == MIND_MECH_NETSTIM) {
      // initialize_random_kernel! with digest of the NetStim source (cuda/micro.jl:79-91)
      const uint64_t digest = 0xa186833b569ff3b2ull;
      for (int r = 0; r < g.rows(); ++r) {
        g.rng[r].keys[0] = (uint32_t)g.logical[r];
        g.rng[r].keys[1] = g.logical[r] >> 32;
        g.rng[r].keys[2] = (uint64_t)(uint32_t)digest ^ 0u;
        rng_setseq(g.rng[r], s.seed, 0.0);
      }
    }
  }
  // ion rows: packed values, concentrations reset (cinit), initial reversals
  // (initialize_ions_kernel!, prepare_currents_kernel!(Initial))
  for (int c = 0; c < 5; ++c) s.ion[c] = s.ion_packed[c];
  for (int64_t q = 0; q < s.nion; ++q) {
    if (s.ion_style[q] & 0x04) {
      s.ion[1][q] = s.species[s.ion_sp[q]].ci;
      s.ion[2][q] = s.species[s.ion_sp[q]].co;
    }
    ion_prepare(s, q, 0x20);
  }
  initial_blocks(s);
  flush(s);
  deliver(s, 0.0);
  currents(s, 0.0, 0, (int64_t)s.cell_offset.size() - 1);
  flush(s);
  s.recording_spikes = false;
  s.initialized = s.error == 0;
  if (s.error) return fail(s.error, "initialization failed");
  return 0;
}

// Field read in canonical row order.
int oracle_row_count(void* p, int mech, int64_t* rows) {
  Sim& s = *static_cast<Sim*>(p);
  const int gi = mech >= 0 && mech < MIND_MECH_SLOTS ? s.group_of_mech[mech] : -1;
  *rows = gi < 0 ? 0 : s.groups[gi].rows();
  return 0;
}

int oracle_read_field(void* p, int mech, int field, double* out, int64_t n) {
  Sim& s = *static_cast<Sim*>(p);
  if (mech < 0 || mech >= MIND_MECH_SLOTS) return fail(MIND_EARG, "unknown mechanism");
  const int gi = s.group_of_mech[mech];
  if (gi < 0) return fail(MIND_EARG, "mechanism not in model");
  Group& g = s.groups[gi];
  if (n != g.rows()) return fail(MIND_EARG, "size mismatch");
  for (int r = 0; r < g.rows(); ++r) out[r] = g.at(field, r);
  return 0;
}

int oracle_node_count(void* p, int64_t* nodes, int64_t* compartments) {
  Sim& s = *static_cast<Sim*>(p);
  *nodes = s.total;
  *compartments = s.compartments;
  return 0;
}

int oracle_read_voltage(void* p, double* out, int64_t n) {
  Sim& s = *static_cast<Sim*>(p);
  if (n != s.total) return fail(MIND_EARG, "size mismatch");
  memcpy(out, s.v.data(), sizeof(double) * n);
  return 0;
}

// Packed geometry for tests: parent, area, cap, axial, axial_diag (canonical order).
int oracle_geometry(void* p, int32_t* parent, double* area, double* cap, double* axial,
                    double* axial_diag) {
  Sim& s = *static_cast<Sim*>(p);
  for (int64_t i = 0; i < s.total; ++i) {
    parent[i] = s.parent[i];
    area[i] = s.area[i];
    cap[i] = s.cap[i];
    axial[i] = s.axial[i];
    axial_diag[i] = s.axial_diag[i];
  }
  return 0;
}

int oracle_location_node(void* p, int32_t cell, int32_t section, double x, int32_t range,
                         int64_t* node) {
  Sim& s = *static_cast<Sim*>(p);
  const Template& t = s.templates[s.cell_morph[cell]];
  *node = s.cell_offset[cell] + (range ? range_node(t, section, x) : location_node(t, section, x));
  return 0;
}

int oracle_point_row(void* p, int32_t point, int32_t* row) {
  Sim& s = *static_cast<Sim*>(p);
  *row = s.point_row[point];
  return 0;
}

void oracle_log_deliveries(void* p, int32_t on) { static_cast<Sim*>(p)->log_deliveries = on != 0; }

// Applied events in delivery order: time, heap ticket, edge (1-based CSR
// index, 0 for self-events) and target point.
int64_t oracle_delivery_log(void* p, double* t, int64_t* ticket, int64_t* edge, int32_t* target,
                            int64_t cap) {
  Sim& s = *static_cast<Sim*>(p);
  const int64_t n = (int64_t)s.log_time.size();
  for (int64_t i = 0; i < n && i < cap; ++i) {
    t[i] = s.log_time[i];
    ticket[i] = s.log_order[i];
    edge[i] = s.log_edge[i];
    target[i] = s.log_target[i];
  }
  return n;
}

// run! (step.jl:4-74), micro only. Records resolve like voltage_ref / mechanism_ref.
int oracle_run(void* p, double stop, const mind_record* recs, int64_t nrec, int32_t record_spikes,
               mind_result* out) {
  Sim& s = *static_cast<Sim*>(p);
  if (!s.initialized) return fail(MIND_EARG, "simulation is not initialized");
  const double now = (double)s.step * s.dt;
  const double duration = stop - now;
  if (!mind_isfinite(duration)) return fail(MIND_EARG, "stop time must be finite");
  const int64_t count = (int64_t)nearbyint(duration / s.dt);  // round-half-even
  const double big = std::max(std::max(fabs(stop), fabs(now)), s.dt);
  const double tol = 64.0 * (nextafter(big, INFINITY) - big);       // 64 * eps(big)
  if (!(count >= 0 && fabs(duration - (double)count * s.dt) <= tol))
    return fail(MIND_EARG, "stop time must be an integer number of steps after the current time");
  // resolve records to (array, index)
  struct Ref { int kind; int64_t node; int gi; int row; int field; };
  std::vector<Ref> refs;
  for (int64_t k = 0; k < nrec; ++k) {
    const mind_record& r = recs[k];
    Ref ref{r.kind, 0, 0, 0, r.field};
    if (r.kind == 0) {
      int64_t node;
      oracle_location_node(p, r.cell, r.section, r.x, 0, &node);
      ref.node = node;
    } else if (r.kind == 1) {
      ref.gi = s.point_group[r.point];
      ref.row = s.point_row[r.point];
    } else if (r.kind == 3) {  // ion_ref (references.jl:156-178)
      int64_t node;
      oracle_location_node(p, r.cell, r.section, r.x, 1, &node);
      ref.row = -1;
      for (int64_t q = s.ion_off[node]; q < s.ion_off[node + 1]; ++q)
        if (s.ion_sp[q] == r.mech) ref.row = (int)q;
      if (ref.row < 0) return fail(MIND_EARG, "ion is not present at this location");
    } else {
      int64_t node;
      oracle_location_node(p, r.cell, r.section, r.x, 1, &node);
      ref.gi = s.group_of_mech[r.mech];
      if (ref.gi < 0) return fail(MIND_EARG, "mechanism not in model");
      Group& g = s.groups[ref.gi];
      auto it = std::lower_bound(g.node.begin(), g.node.end(), (int32_t)node);
      if (it == g.node.end() || *it != node) return fail(MIND_EARG, "no density mechanism at location");
      ref.row = (int)(it - g.node.begin());
    }
    refs.push_back(ref);
  }
  const int64_t samples = count + 1;
  std::vector<double> traces((size_t)nrec * samples);
  auto sample = [&](int64_t j) {
    for (int64_t k = 0; k < nrec; ++k) {
      const Ref& r = refs[k];
      traces[(size_t)k * samples + j] = r.kind == 0   ? s.v[r.node]
                                        : r.kind == 3 ? s.ion[r.field][r.row]
                                                      : s.groups[r.gi].at(r.field, r.row);
    }
  };
  s.recording_spikes = record_spikes != 0;
  if (s.step != 0 || !s.recording_spikes) {
    s.spike_t.clear();
    s.spike_src.clear();
  }
  const int64_t start_step = s.step;
  sample(0);
  const auto t0 = std::chrono::steady_clock::now();
  for (int64_t j = 1; j <= count; ++j) {
    advance_micro(s);
    sample(j);
    deliver(s, (double)s.step * s.dt);
    if (s.error) break;
  }
  flush(s);
  const auto t1 = std::chrono::steady_clock::now();
  s.recording_spikes = false;
  if (s.error) {
    s.initialized = false;
    return fail(s.error, s.error == MIND_ESINGULAR ? "singular or non-finite cable system"
                                                    : "invalid self-event delay or target");
  }
  for (double x : s.v)
    if (!mind_isfinite(x)) {
      s.initialized = false;
      return fail(MIND_ENONFINITE, "non-finite numerical state");
    }
  // result
  out->n_samples = samples;
  out->time = (double*)malloc(sizeof(double) * samples);
  for (int64_t j = 0; j < samples; ++j) out->time[j] = (double)(start_step + j) * s.dt;
  out->n_records = nrec;
  out->traces = (double*)malloc(sizeof(double) * traces.size() + 8);
  memcpy(out->traces, traces.data(), sizeof(double) * traces.size());
  const int64_t ns = (int64_t)s.spike_t.size();
  std::vector<int64_t> order(ns);
  for (int64_t i = 0; i < ns; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    const double ta = s.spike_t[a], tb = s.spike_t[b];
    const int64_t ia = s.source_sid[s.spike_src[a]], ib = s.source_sid[s.spike_src[b]];
    return ta < tb || (ta == tb && ia < ib);
  });
  out->n_spikes = ns;
  out->spike_times = (double*)malloc(sizeof(double) * ns + 8);
  out->spike_sids = (int64_t*)malloc(sizeof(int64_t) * ns + 8);
  for (int64_t i = 0; i < ns; ++i) {
    out->spike_times[i] = s.spike_t[order[i]];
    out->spike_sids[i] = s.source_sid[s.spike_src[order[i]]];
  }
  out->run_seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

void oracle_result_free(mind_result* r) {
  free(r->time);
  free(r->traces);
  free(r->spike_times);
  free(r->spike_sids);
  memset(r, 0, sizeof(*r));
}

// Test hook: the oracle's Nernst potential (its own restatement).
double oracle_nernst(double ci, double co, double z, double celsius) { return nernst(ci, co, z, celsius); }

// Test hooks: Philox block, uniform stream, Hines solve on a given tree.
void oracle_philox(const uint32_t ctr_in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
  philox4x32_10(c, key[0], key[1]);
  memcpy(out, c, sizeof(c));
}

void oracle_uniform_stream(uint64_t seed, uint64_t logical_id, uint64_t digest, int64_t n,
                           double* out) {
  Rng r;
  r.keys[0] = (uint32_t)logical_id;
  r.keys[1] = logical_id >> 32;
  r.keys[2] = (uint32_t)digest;
  rng_setseq(r, seed, 0.0);
  for (int64_t i = 0; i < n; ++i) out[i] = rng_uniform(r, seed);
}

// Solve the tree system with diagonal d, off-diagonal -g (g[i] couples i and
// parent[i]) and right-hand side b, in the oracle's elimination order.
int oracle_hines(int64_t n, const int32_t* parent, const double* g, double* d, double* b) {
  for (int64_t i = n - 1; i >= 0; --i) {
    const int32_t p = parent[i];
    if (d[i] == 0.0 || !mind_isfinite(d[i])) return MIND_ESINGULAR;
    if (p >= 0) {
      const double factor = g[i] / d[i];
      d[p] -= factor * g[i];
      b[p] += factor * b[i];
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    double value = b[i];
    if (parent[i] >= 0) value += g[i] * b[parent[i]];
    b[i] = value / d[i];
  }
  return 0;
}

double oracle_exp(double x) { return mind_exp(x); }
double oracle_log(double x) { return mind_log(x); }

}  // extern "C"

// ================================================================ macro ======
// Micro-macro coupling, restating advance_macro! (src/runtime/step.jl:156-204)
// with the builtin MIND-role mechanisms of this project:
//   REGION      Wilson-Cowan E/I populations, RK4 over one macro step;
//   MACRO2MACRO coupling_i = sum_j W_ij E_j(t - D_ij) over the history ring
//               (prepare_coupling_kernel! / macro_edge_kernel!, cuda/macro.jl:47-123);
//   MACRO2MICRO inhomogeneous Poisson events at rate_scale * E, rekeyed per
//               macro step (input_trigger_kernel!, cuda/macro.jl:272-301);
//   MICRO2MACRO spike count of the ROI's sources per interval, published as a
//               rate (Hz) when a bin closes.
namespace {

struct Macro {
  int n = 0;
  double dt_macro = 0.1;
  int64_t ratio = 4;  // micro steps per macro step
  std::vector<double> W, wc;           // W[target * n + source], wc[roi * WC_NPARAMS + k]
  std::vector<int64_t> delay_steps;    // [target * n + source]
  int64_t cap = 1;
  std::vector<double> hist;            // [slot * n + roi] published E
  std::vector<double> E, I, coupling, rate, acc, meanv;
  std::vector<int64_t> bin_steps;
  std::vector<int64_t> in_off;         // bindings of each ROI
  std::vector<int32_t> in_roi;         // per binding
  std::vector<uint64_t> stream_id;     // per binding
  int64_t n_points = 0, first_virtual = 0;
  std::vector<int64_t> out_off;
  std::vector<int32_t> out_src, out_node;
  std::vector<int32_t> src_roi;        // per source: ROI of the MICRO2MACRO that reads it, or -1
  int64_t step = 0;
  Rng rng;
};

// S(x) = 1/(1 + exp(-a (x - theta))) - 1/(1 + exp(a theta))
double wc_sigmoid(double x, double a, double theta) {
  return 1.0 / (1.0 + mind_exp(-a * (x - theta))) - 1.0 / (1.0 + mind_exp(a * theta));
}

void wc_deriv(const double* p, double e, double i, double u, double& de, double& di) {
  de = (-e + wc_sigmoid(p[WC_C_EE] * e - p[WC_C_EI] * i + p[WC_P] + u, p[WC_A_E], p[WC_THETA_E])) /
       p[WC_TAU_E];
  di = (-i + wc_sigmoid(p[WC_C_IE] * e - p[WC_C_II] * i + p[WC_Q], p[WC_A_I], p[WC_THETA_I])) /
       p[WC_TAU_I];
}

void wc_rk4(const double* p, double& e, double& i, double u, double h) {
  double k1e, k1i, k2e, k2i, k3e, k3i, k4e, k4i;
  wc_deriv(p, e, i, u, k1e, k1i);
  wc_deriv(p, e + 0.5 * h * k1e, i + 0.5 * h * k1i, u, k2e, k2i);
  wc_deriv(p, e + 0.5 * h * k2e, i + 0.5 * h * k2i, u, k3e, k3i);
  wc_deriv(p, e + h * k3e, i + h * k3i, u, k4e, k4i);
  e = e + h / 6.0 * (k1e + 2.0 * k2e + 2.0 * k3e + k4e);
  i = i + h / 6.0 * (k1i + 2.0 * k2i + 2.0 * k3i + k4i);
}

// PoissonDrive NET_RECEIVE at the start of an exchange interval:
//   if (rate > 0) { tn = start + negexp * 1000/rate
//                   while (tn < stop) { net_event(tn); tn = tn + negexp * 1000/rate } }
void trigger_inputs(Sim& s, Macro& M) {
  const double start = (double)M.step * M.dt_macro, stop = (double)(M.step + 1) * M.dt_macro;
  const int64_t slot = M.step % M.cap;
  for (int64_t b = 0; b < (int64_t)M.in_roi.size(); ++b) {
    const int roi = M.in_roi[b];
    Rng r;
    r.keys[0] = (uint32_t)(s.seed ^ M.stream_id[b]);
    r.keys[1] = (uint32_t)(uint64_t)b;
    r.keys[2] = (uint32_t)((s.seed >> 32) ^ (M.stream_id[b] >> 32) ^ (uint64_t)roi ^ 0ull);
    rng_setseq(r, 0, (double)M.step);  // MACRO2MICRO streams use a zero Philox key
    const double rate = M.wc[roi * WC_NPARAMS + WC_RATE_SCALE] * M.hist[slot * M.n + roi];
    if (!(rate > 0.0)) continue;
    const double scale = 1000.0 / rate;
    double tn = start + -mind_log(rng_uniform(r, 0)) * 1.0 * scale;
    while (tn < stop) {
      net_event(s, (int)(M.n_points + b), tn, 0);
      tn = tn + -mind_log(rng_uniform(r, 0)) * 1.0 * scale;
    }
  }
  flush(s);
}

void macro_record(Macro& M, std::vector<double>& tr, int64_t samples, int64_t j) {
  for (int r = 0; r < M.n; ++r) {
    tr[(0 * (size_t)M.n + r) * samples + j] = M.E[r];
    tr[(1 * (size_t)M.n + r) * samples + j] = M.I[r];
    tr[(2 * (size_t)M.n + r) * samples + j] = M.rate[r];
    tr[(3 * (size_t)M.n + r) * samples + j] = M.coupling[r];
    tr[(4 * (size_t)M.n + r) * samples + j] = M.meanv[r];
  }
}

void advance_macro(Sim& s, Macro& M) {
  // 1. delayed coupling from the history ring
  for (int i = 0; i < M.n; ++i) {
    double c = 0.0;
    for (int j = 0; j < M.n; ++j) {
      const double w = M.W[(size_t)i * M.n + j];
      if (w == 0.0) continue;
      const int64_t slot = ((M.step - M.delay_steps[(size_t)i * M.n + j]) % M.cap + M.cap) % M.cap;
      c = c + w * M.hist[slot * M.n + j];
    }
    M.coupling[i] = c;
  }
  // 2. MACRO2MICRO events for [start, stop)
  s.window_src.clear();
  s.window_t.clear();
  s.collect_window = true;
  trigger_inputs(s, M);
  // 3. micro steps; micro samples are taken by the caller's recording
  for (int64_t k = 0; k < M.ratio; ++k) {
    advance_micro(s);
    deliver(s, (double)s.step * s.dt);
  }
  s.collect_window = false;
  // 4. MICRO2MACRO: spikes of each ROI's sources in [start, stop)
  const double start = (double)M.step * M.dt_macro, stop = (double)(M.step + 1) * M.dt_macro;
  std::vector<double> count(M.n, 0.0);
  for (size_t q = 0; q < s.window_src.size(); ++q) {
    const int r = M.src_roi[s.window_src[q]];
    if (r >= 0 && s.window_t[q] >= start && s.window_t[q] < stop) count[r] += 1.0;
  }
  for (int r = 0; r < M.n; ++r) {
    M.acc[r] = M.acc[r] + count[r];
    if ((M.step + 1) % M.bin_steps[r] == 0) {
      const double nsrc = (double)(M.out_off[r + 1] - M.out_off[r]);
      const double bin = (double)M.bin_steps[r] * M.dt_macro;
      M.rate[r] = nsrc > 0.0 ? M.acc[r] / (nsrc * bin) * 1000.0 : 0.0;
      M.acc[r] = 0.0;
    }
    double sum = 0.0;
    for (int64_t q = M.out_off[r]; q < M.out_off[r + 1]; ++q) sum += s.v[M.out_node[q]];
    M.meanv[r] = M.out_off[r + 1] > M.out_off[r] ? sum / (double)(M.out_off[r + 1] - M.out_off[r]) : 0.0;
  }
  // 5. REGION advance over [start, stop)
  for (int r = 0; r < M.n; ++r) {
    const double* p = &M.wc[(size_t)r * WC_NPARAMS];
    const double u = M.coupling[r] + p[WC_K_MICRO] * (M.rate[r] / p[WC_RATE_SCALE]);
    wc_rk4(p, M.E[r], M.I[r], u, M.dt_macro);
  }
  // 6. publish the boundary into the history ring
  M.step += 1;
  for (int r = 0; r < M.n; ++r) M.hist[(M.step % M.cap) * M.n + r] = M.E[r];
}

}  // namespace

extern "C" {

void* oracle_build_macro(const mind_model_desc* d, const mind_macro_desc* m, double dt, uint64_t seed) {
  Sim* sp = static_cast<Sim*>(oracle_build(d, dt, seed));
  Sim& s = *sp;
  Macro* M = new Macro();
  s.macro = M;
  M->n = m->n_roi;
  M->dt_macro = m->dt_macro;
  M->ratio = (int64_t)nearbyint(m->dt_macro / dt);
  M->W.assign(m->weights, m->weights + (size_t)M->n * M->n);
  M->wc.assign(m->wc_params, m->wc_params + (size_t)M->n * WC_NPARAMS);
  M->delay_steps.resize((size_t)M->n * M->n);
  int64_t maxd = 0;
  for (size_t q = 0; q < M->delay_steps.size(); ++q) {
    M->delay_steps[q] = (int64_t)nearbyint(m->delays[q] / m->dt_macro);  // pack_macro:225
    if (M->W[q] != 0.0) maxd = std::max(maxd, M->delay_steps[q]);
  }
  M->cap = maxd + 1;
  M->bin_steps.resize(M->n);
  for (int r = 0; r < M->n; ++r) M->bin_steps[r] = std::max<int64_t>(1, (int64_t)nearbyint(m->bin_ms[r] / m->dt_macro));
  // virtual sources of the MACRO2MICRO bindings (model.jl:622-628), after the real ones
  const int64_t nb = m->input_offsets[M->n];
  M->n_points = d->n_points;
  M->first_virtual = d->n_sources;
  M->in_off.assign(m->input_offsets, m->input_offsets + M->n + 1);
  M->in_roi.resize(nb);
  M->stream_id.resize(nb);
  for (int r = 0; r < M->n; ++r)
    for (int64_t b = m->input_offsets[r]; b < m->input_offsets[r + 1]; ++b) M->in_roi[b] = r;
  // one random identity per (mechanism, ROI, parameters): all bindings of an ROI share it
  std::vector<int64_t> roi_stream(M->n, -1);
  int64_t next_id = 0;
  for (int64_t b = 0; b < nb; ++b) {
    if (roi_stream[M->in_roi[b]] < 0) roi_stream[M->in_roi[b]] = next_id++;
    M->stream_id[b] = (uint64_t)roi_stream[M->in_roi[b]];
  }
  for (int64_t b = 0; b < nb; ++b) {
    s.source_sid.push_back(-(b + 1));
    s.source_target.push_back((int32_t)(d->n_points + b + 1));
    s.target_sources.push_back({(int32_t)(d->n_sources + b)});
    s.sequences.push_back(0);
    s.target_rank.push_back(s.rank_total + b + 1);  // input rows rank after NetStims
    s.fanout.push_back(s.fanout.back() + 1);
    s.edge_target.push_back(m->input_point[b] + 1);
    s.edge_delay.push_back(m->input_delay[b]);
    s.edge_weight.push_back(m->input_weight[b]);
  }
  const int64_t rank_all = s.rank_total + nb;
  s.n_sources = d->n_sources + nb;
  s.source_rank.resize(s.n_sources);
  for (int64_t k = 0; k < s.n_sources; ++k)
    s.source_rank[k] = s.source_target[k] == 0 ? rank_all + k + 1 : s.target_rank[s.source_target[k] - 1];
  M->out_off.assign(m->output_offsets, m->output_offsets + M->n + 1);
  M->out_src.assign(m->output_source, m->output_source + m->output_offsets[M->n]);
  M->src_roi.assign(s.n_sources, -1);
  for (int r = 0; r < M->n; ++r)
    for (int64_t q = M->out_off[r]; q < M->out_off[r + 1]; ++q) {
      M->src_roi[M->out_src[q]] = r;
      int64_t node = -1;
      for (size_t k = 0; k < s.det_source.size(); ++k)
        if (s.det_source[k] == M->out_src[q]) node = s.det_node[k];
      M->out_node.push_back((int32_t)node);
    }
  return sp;
}

int oracle_initialize_macro(void* p) {
  Sim& s = *static_cast<Sim*>(p);
  Macro& M = *static_cast<Macro*>(s.macro);
  int rc = oracle_initialize(p);
  if (rc) return rc;
  std::fill(s.sequences.begin(), s.sequences.end(), 0);
  M.step = 0;
  M.E.resize(M.n); M.I.resize(M.n);
  M.coupling.assign(M.n, 0.0); M.rate.assign(M.n, 0.0); M.acc.assign(M.n, 0.0);
  M.meanv.assign(M.n, 0.0);
  M.hist.assign((size_t)M.cap * M.n, 0.0);
  for (int r = 0; r < M.n; ++r) {
    M.E[r] = M.wc[(size_t)r * WC_NPARAMS + WC_E0];
    M.I[r] = M.wc[(size_t)r * WC_NPARAMS + WC_I0];
    for (int64_t k = 0; k < M.cap; ++k) M.hist[k * M.n + r] = M.E[r];  // initialize_history_kernel!
    double sum = 0.0;
    for (int64_t q = M.out_off[r]; q < M.out_off[r + 1]; ++q) sum += s.v[M.out_node[q]];
    M.meanv[r] = M.out_off[r + 1] > M.out_off[r] ? sum / (double)(M.out_off[r + 1] - M.out_off[r]) : 0.0;
  }
  return 0;
}

int oracle_run_macro(void* p, double stop, int32_t record_spikes, mind_macro_result* out) {
  Sim& s = *static_cast<Sim*>(p);
  Macro& M = *static_cast<Macro*>(s.macro);
  if (!s.initialized) return fail(MIND_EARG, "simulation is not initialized");
  const double now = (double)M.step * M.dt_macro;
  const double duration = stop - now;
  const int64_t count = (int64_t)nearbyint(duration / M.dt_macro);
  const double big = std::max(std::max(fabs(stop), fabs(now)), M.dt_macro);
  const double tol = 64.0 * (nextafter(big, INFINITY) - big);
  if (!(count >= 0 && fabs(duration - (double)count * M.dt_macro) <= tol))
    return fail(MIND_EARG, "stop time must be an integer number of steps after the current time");
  const int64_t samples = count + 1;
  std::vector<double> tr((size_t)5 * M.n * samples);
  s.recording_spikes = record_spikes != 0;
  if (M.step != 0 || !s.recording_spikes) {
    s.spike_t.clear();
    s.spike_src.clear();
  }
  const int64_t start_step = M.step;
  macro_record(M, tr, samples, 0);
  const auto t0 = std::chrono::steady_clock::now();
  for (int64_t j = 1; j <= count; ++j) {
    advance_macro(s, M);
    macro_record(M, tr, samples, j);
    if (s.error) break;
  }
  flush(s);
  const auto t1 = std::chrono::steady_clock::now();
  s.recording_spikes = false;
  if (s.error) {
    s.initialized = false;
    return fail(s.error, "macro run failed");
  }
  out->n_samples = samples;
  out->n_roi = M.n;
  out->time = (double*)malloc(sizeof(double) * samples);
  for (int64_t j = 0; j < samples; ++j) out->time[j] = (double)(start_step + j) * M.dt_macro;
  out->traces = (double*)malloc(sizeof(double) * tr.size());
  memcpy(out->traces, tr.data(), sizeof(double) * tr.size());
  const int64_t ns = (int64_t)s.spike_t.size();
  std::vector<int64_t> order(ns);
  for (int64_t i = 0; i < ns; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    const double ta = s.spike_t[a], tb = s.spike_t[b];
    const int64_t ia = s.source_sid[s.spike_src[a]], ib = s.source_sid[s.spike_src[b]];
    return ta < tb || (ta == tb && ia < ib);
  });
  out->n_spikes = ns;
  out->spike_times = (double*)malloc(sizeof(double) * ns + 8);
  out->spike_sids = (int64_t*)malloc(sizeof(int64_t) * ns + 8);
  for (int64_t i = 0; i < ns; ++i) {
    out->spike_times[i] = s.spike_t[order[i]];
    out->spike_sids[i] = s.source_sid[s.spike_src[order[i]]];
  }
  out->run_seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

void oracle_macro_result_free(mind_macro_result* r) {
  free(r->time);
  free(r->traces);
  free(r->spike_times);
  free(r->spike_sids);
  memset(r, 0, sizeof(*r));
}

void oracle_free_macro(void* p) {
  Sim* s = static_cast<Sim*>(p);
  delete static_cast<Macro*>(s->macro);
  s->macro = nullptr;
  delete s;
}

// Test hook: one RK4 step of the Wilson-Cowan REGION.
void oracle_wc_rk4(const double* params, double* e, double* i, double u, double h) {
  wc_rk4(params, *e, *i, u, h);
}

}  // extern "C"

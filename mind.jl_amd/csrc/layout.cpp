// Packing of mind_model_desc into the device layout (see layout.h).
//
// Geometry follows the reference's compartmentalisation: a root node of area
// 100 um^2 with no capacitance, one node per segment holding the area of its
// two half-segments, and a zero-capacitance endpoint per section
// (src/model/micro/morphology.jl:406-484). Half-segment area, axial resistance
// and mean diameter are integrated piecewise-linearly over the pt3d samples
// with a cursor carried across half-segments
// (src/model/micro/utils/morphology.jl:28-63). Compiled with -ffp-contract=off:
// these numbers feed a bit-exact comparison.
#include "layout.h"
#include "mechreg.h"

#include <math.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>

namespace mind {

int field_count(int mech) {
  static const int n[MIND_MECH_COUNT] = {HH_NFIELDS,     PAS_NFIELDS,     EXPSYN_NFIELDS,
                                         EXP2SYN_NFIELDS, ICLAMP_NFIELDS,  NETSTIM_NFIELDS,
                                         ALPHASYNAPSE_NFIELDS, PATTERNSTIM_NFIELDS, SECLAMP_NFIELDS,
                                      VCLAMP_NFIELDS};
  if (mech >= 0 && mech < MIND_MECH_COUNT) return n[mech];
  const UserSlot* u = user_slot(mech);
  return u ? (int)u->m.columns.size() : 0;
}

// Point processes and artificial cells (created by add_pointprocess!).
bool mech_is_point(int mech) {
  if (mech >= 0 && mech < MIND_MECH_COUNT) return mech >= MIND_MECH_EXPSYN;
  return mech_known(mech) && mech_kind(mech) != MIND_KIND_DENSITY;
}

void field_defaults(int mech, double* f) {
  for (int i = 0; i < field_count(mech); ++i) f[i] = 0.0;
  if (const UserSlot* u = user_slot(mech)) {  // PARAMETER defaults and ion values of a compiled mechanism
    for (size_t i = 0; i < u->m.defaults.size(); ++i) f[i] = u->m.defaults[i];
    return;
  }
  switch (mech) {
    case MIND_MECH_HH:  // hh.mod PARAMETER block; ena/ek from the default ions
      f[HH_GNABAR] = 0.12; f[HH_GKBAR] = 0.036; f[HH_GL] = 0.0003; f[HH_EL] = -54.3;
      f[HH_ENA] = 50.0; f[HH_EK] = -77.0;
      break;
    case MIND_MECH_PAS: f[PAS_G] = 0.001; f[PAS_E] = -70.0; break;
    case MIND_MECH_EXPSYN: f[EXPSYN_TAU] = 0.1; break;
    case MIND_MECH_EXP2SYN: f[EXP2SYN_TAU1] = 0.1; f[EXP2SYN_TAU2] = 10.0; break;
    case MIND_MECH_NETSTIM:
      f[NETSTIM_INTERVAL] = 10.0; f[NETSTIM_NUMBER] = 10.0; f[NETSTIM_START] = 50.0;
      break;
    case MIND_MECH_ALPHASYNAPSE: f[ALPHASYNAPSE_TAU] = 0.1; break;  // AlphaSynapse.mod PARAMETER
    case MIND_MECH_SECLAMP: f[SECLAMP_RS] = 1.0; break;              // SEClamp.mod PARAMETER
    case MIND_MECH_VCLAMP:                                             // VClamp.mod PARAMETER
      f[VCLAMP_GAIN] = 1e5; f[VCLAMP_RSTIM] = 1.0; f[VCLAMP_TAU1] = 0.001;
      break;
    default: break;
  }
}

struct IonParam {  // a mechanism parameter (group, row, field, value), maybe an ion variable
  int group;
  int64_t row;
  int field;
  double value;
};
std::string pack_ions(const mind_model_desc& d, const std::vector<IonParam>& params, Layout& L);

namespace {

struct Sample { double x, y, z, d; };

// Walks the pt3d polyline of one section, integrating area, axial resistance
// and diameter between successive boundaries.
class GeometryCursor {
 public:
  GeometryCursor(const std::vector<Sample>& s, double Ra) : s_(s), Ra_(Ra), arc_(s.size(), 0.0) {
    for (size_t i = 1; i < s.size(); ++i) {
      const double dx = s[i].x - s[i - 1].x, dy = s[i].y - s[i - 1].y, dz = s[i].z - s[i - 1].z;
      arc_[i] = arc_[i - 1] + sqrt(dx * dx + dy * dy + dz * dz);
    }
    last_arc_ = arc_[0];
    last_d_ = s[0].d;
  }
  double total() const { return arc_.back(); }

  // Integrates from the previous boundary to `boundary`; returns area (um^2),
  // resistance (Mohm) and the diameter integral (um^2).
  void advance(double boundary, double& area, double& resistance, double& diam) {
    double a = 0.0, r = 0.0, di = 0.0;
    const size_t last = s_.size() - 1;
    for (;;) {
      const size_t nx = span_ + 1;
      const bool stop = arc_[nx] > boundary || nx == last;
      double arc, d;
      if (stop) {
        const double width = arc_[nx] - arc_[span_];
        const double frac = fabs(width) < 1.0e-10 ? 1.0 : (boundary - arc_[span_]) / width;
        arc = boundary;
        d = (1.0 - frac) * s_[span_].d + frac * s_[nx].d;
      } else {
        arc = arc_[nx];
        d = s_[nx].d;
      }
      const double step = arc - last_arc_;
      di += (last_d_ + d) * step;
      const double len = step > 1.0e-15 ? step : 1.0e-15;
      const double cond = last_d_ * d / len;
      r += 1.0 / (cond == 0.0 ? 1.0e-15 : cond);
      const double dd = d - last_d_;
      a += (last_d_ + d) * sqrt(len * len + 0.25 * (dd * dd));
      last_arc_ = arc;
      last_d_ = d;
      if (stop) break;
      span_ = nx;
    }
    area = 0.5 * M_PI * a;
    resistance = 0.04 * Ra_ / M_PI * r;
    diam = 0.5 * di;
  }

 private:
  const std::vector<Sample>& s_;
  double Ra_;
  std::vector<double> arc_;
  size_t span_ = 0;
  double last_arc_, last_d_;
};

int32_t template_location(const Template& t, int sec, double x) {
  if (x == 0.0) return t.sec_start[sec];
  const std::vector<int32_t>& nodes = t.sec_nodes[sec];
  const int64_t segs = (int64_t)nodes.size() - 1;
  if (x == 1.0) return nodes[segs];
  return nodes[std::min((int64_t)floor(x * segs), segs - 1)];
}

int32_t template_range(const Template& t, int sec, double x) {
  const std::vector<int32_t>& nodes = t.sec_nodes[sec];
  const int64_t segs = (int64_t)nodes.size() - 1;
  return nodes[std::min((int64_t)floor(x * segs), segs - 1)];
}

std::string build_template(const mind_model_desc& d, int m, Template& t) {
  const int64_t s0 = d.morph_offsets[m], s1 = d.morph_offsets[m + 1];
  const int nsec = (int)(s1 - s0);
  t.sec_nodes.assign(nsec, {});
  t.sec_start.assign(nsec, -1);
  // Parent-first section order, otherwise input order (utils/morphology.jl:77-104).
  std::vector<int> order;
  order.reserve(nsec);
  std::vector<uint8_t> state(nsec, 0);
  std::vector<int> chain;
  for (int start = 0; start < nsec; ++start) {
    int s = start;
    while (state[s] != 2) {
      if (state[s] == 1) return "morphology contains a cycle";
      state[s] = 1;
      chain.push_back(s);
      const int p = d.sec_parent[s0 + s];
      if (p < 0) break;
      if (p >= nsec) return "unknown parent section";
      s = p;
    }
    for (auto it = chain.rbegin(); it != chain.rend(); ++it) {
      state[*it] = 2;
      order.push_back(*it);
    }
    chain.clear();
  }
  auto node = [&](int32_t parent, double area, double g, double c) {
    t.parent.push_back(parent);
    t.area.push_back(area);
    t.axial.push_back(g);
    t.cap.push_back(c);
    return (int32_t)t.parent.size() - 1;
  };
  for (int s : order) {
    const int64_t i = s0 + s;
    const int nseg = d.sec_nseg[i];
    if (nseg <= 0 || !(d.sec_Ra[i] > 0) || !(d.sec_cm[i] >= 0))
      return "invalid segmentation or biophysics";
    std::vector<Sample> pts;
    if (d.sec_point_offsets && d.sec_point_offsets[i + 1] - d.sec_point_offsets[i] >= 2) {
      for (int64_t p = d.sec_point_offsets[i]; p < d.sec_point_offsets[i + 1]; ++p)
        pts.push_back({d.points[4 * p], d.points[4 * p + 1], d.points[4 * p + 2], d.points[4 * p + 3]});
    } else {
      pts = {{0.0, 0.0, 0.0, d.sec_diameter[i]}, {d.sec_length[i], 0.0, 0.0, d.sec_diameter[i]}};
    }
    GeometryCursor geo(pts, d.sec_Ra[i]);
    if (!(geo.total() > 0.0)) return "section has zero arc length";
    const int p = d.sec_parent[i];
    int32_t attach;
    if (p < 0) {
      attach = node(-1, 100.0, 0.0, 0.0);
    } else {
      if (t.sec_nodes[p].empty()) return "unknown parent section";
      const double px = d.sec_parent_x[i];
      if (!(px >= 0.0 && px <= 1.0)) return "parent_x must be in [0, 1]";
      attach = template_location(t, p, px);
    }
    const bool flipped = d.sec_child_x[i] == 1.0;
    std::vector<int32_t>& nodes = t.sec_nodes[s];
    nodes.assign(nseg + 1, -1);
    const double h = geo.total() / nseg;
    int32_t prev = attach;
    double carry = 0.0;  // right half-segment resistance of the previous segment
    for (int k = 1; k <= nseg; ++k) {
      double aL, rL, dL, aR, rR, dR;
      geo.advance((k - 0.5) * h, aL, rL, dL);
      geo.advance(k * h, aR, rR, dR);
      const double area = aL + aR;
      const int32_t n = node(prev, area, 1.0 / (carry + rL), d.sec_cm[i] * area * 1.0e-5);
      nodes[flipped ? nseg - k : k - 1] = n;
      prev = n;
      carry = rR;
    }
    const int32_t end = node(prev, 100.0, 1.0 / carry, 0.0);
    t.sec_start[s] = flipped ? end : attach;
    nodes[nseg] = flipped ? attach : end;
    t.compartments += nseg;
  }
  return "";
}

}  // namespace

// ---------------------------------------------------------------- ions ----
// IonStyle bits (ion_style_bits, src/model/micro/utils/biophysics.jl:3-4): c_style
// (bits 0-1), cinit 0x04, e_style (bits 3-4), einit 0x20, eadvance 0x40.
uint8_t ion_style_bits(int c, int e, bool einit, bool eadvance, bool cinit) {
  return (uint8_t)(c + 4 * (cinit ? 1 : 0) + 8 * e + 32 * (einit ? 1 : 0) + 64 * (eadvance ? 1 : 0));
}
// promote_ion_style (utils/biophysics.jl:9-22)
uint8_t promote_ion_style(uint8_t style, int concentration, int reversal) {
  concentration = std::max((int)(style & 0x03), concentration);
  reversal = std::max((int)((style >> 3) & 0x03), reversal);
  if (concentration > 0 && reversal < 2) reversal = 2;
  return ion_style_bits(concentration, reversal, concentration > 0 && reversal == 2,
                        concentration == 3 && reversal == 2, concentration == 3);
}

int64_t Layout::ion_row_at(int64_t node, int32_t sp) const {
  if (node < 0 || node >= nodes) return -1;
  for (int64_t q = node_ion_off[node]; q < node_ion_off[node + 1]; ++q)
    if (ion_species[q] == sp) return q;
  return -1;
}

// Ion species, rows, styles and initial values, and each group's transfers
// (src/model/micro/model.jl:672-855, src/model/utils/build.jl:111-135).
std::string pack_ions(const mind_model_desc& d, const std::vector<IonParam>& params, Layout& L) {
  std::vector<Species>& sp = L.species;
  if (d.n_species == 0) {  // default_ions (utils/biophysics.jl:31-42)
    sp = {{"na", 1.0, 10.0, 140.0, 50.0, false},
          {"k", 1.0, 54.4, 2.5, -77.0, false},
          {"ca", 2.0, 5.0e-5, 2.0, 132.4579341637009, false}};
  } else {
    if (!d.species_name || !d.species_valence || !d.species_internal || !d.species_external ||
        !d.species_reversal || !d.species_nernst)
      return "ion species arrays missing";
    for (int32_t k = 0; k < d.n_species; ++k) {
      Species s;
      s.name = d.species_name[k] ? d.species_name[k] : "";
      s.valence = d.species_valence[k];
      s.internal = d.species_internal[k];
      s.external = d.species_external[k];
      s.reversal = d.species_reversal[k];
      s.nernst = d.species_nernst[k] != 0;
      if (s.name.empty()) return "ion species without a name";
      if (!isfinite(s.valence)) return "ion valence must be finite";
      if (!(isfinite(s.internal) && isfinite(s.external) && isfinite(s.reversal)))
        return "non-finite ion parameter";
      if (s.internal < 0.0 || s.external < 0.0) return "ion concentrations must be nonnegative";
      for (const Species& o : sp)
        if (o.name == s.name) return "duplicate ion species " + s.name;
      sp.push_back(s);
    }
  }
  auto species_of = [&](const std::string& name) {
    for (size_t k = 0; k < sp.size(); ++k)
      if (sp[k].name == name) return (int32_t)k;
    return (int32_t)-1;
  };
  // species of every USEION, checked against its VALENCE; a species only a
  // mechanism declares starts at 1 mM on both sides (model.jl:690-703)
  std::vector<std::vector<int32_t>> slot_species(L.groups.size());
  for (size_t g = 0; g < L.groups.size(); ++g) {
    const std::vector<nmodl::IonUse>& uses = mech_ions(L.groups[g].mech);
    int ncur = 0;
    for (const nmodl::IonUse& u : uses) {
      int32_t k = species_of(u.name);
      if (k < 0) {
        if (!u.has_valence) return "USEION " + u.name + " needs a VALENCE declaration";
        Species s;
        s.name = u.name;
        s.valence = u.valence;
        sp.push_back(s);
        k = (int32_t)sp.size() - 1;
      } else if (u.has_valence && sp[k].valence != u.valence) {
        return "inconsistent valence for ion " + u.name;
      }
      slot_species[g].push_back(k);
      ncur += u.write[MIND_ION_I] >= 0;
    }
    if (ncur > kMaxIonCurrents) return "a mechanism writes more than 8 ion currents";
  }
  std::vector<int32_t> rank(sp.size());  // species by name
  {
    std::vector<int32_t> order(sp.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return sp[a].name < sp[b].name; });
    for (size_t k = 0; k < order.size(); ++k) rank[order[k]] = (int32_t)k;
  }
  // the segments (not the endpoint) of a cell's section
  auto segments = [&](int c, int sec, std::vector<int64_t>& out) -> bool {
    if (c < 0 || c >= L.ncells) return false;
    const Template& t = L.templates[L.cell_template[c]];
    if (sec < 0 || sec >= (int)t.sec_nodes.size()) return false;
    out.clear();
    for (size_t j = 0; j + 1 < t.sec_nodes[sec].size(); ++j) out.push_back(L.cell_off[c] + t.sec_nodes[sec][j]);
    return true;
  };
  // 1. occupied (node, species) pairs
  std::vector<std::pair<int64_t, int32_t>> occupied;  // (node, rank)
  std::vector<int64_t> seg;
  for (size_t g = 0; g < L.groups.size(); ++g) {
    const MechGroup& G = L.groups[g];
    for (int32_t k : slot_species[g])
      for (int64_t r = 0; r < G.rows; ++r)
        if (G.node[r] >= 0) occupied.push_back({G.node[r], rank[k]});
  }
  // a point process's ions cover its whole section (NEURON need_memb)
  for (int64_t p = 0; p < d.n_points; ++p) {
    if (d.pt_cell[p] < 0 || L.point_group[p] < 0) continue;
    for (int32_t k : slot_species[L.point_group[p]]) {
      if (!segments(d.pt_cell[p], d.pt_section[p], seg)) return "unknown section";
      for (int64_t n : seg) occupied.push_back({n, rank[k]});
    }
  }
  if (d.n_ion_settings > 0 &&
      (!d.ionset_cell || !d.ionset_section || !d.ionset_species || !d.ionset_what || !d.ionset_value))
    return "ion setting arrays missing";
  for (int64_t q = 0; q < d.n_ion_settings; ++q) {
    const int32_t k = d.ionset_species[q];
    if (k < 0 || k >= (int32_t)sp.size()) return "ion setting of an unknown species";
    if (!segments(d.ionset_cell[q], d.ionset_section[q], seg)) return "ion setting on an unknown section";
    for (int64_t n : seg) occupied.push_back({n, rank[k]});
  }
  std::sort(occupied.begin(), occupied.end());
  occupied.erase(std::unique(occupied.begin(), occupied.end()), occupied.end());
  std::vector<int32_t> by_rank(sp.size());
  for (size_t k = 0; k < sp.size(); ++k) by_rank[rank[k]] = (int32_t)k;
  L.nion = (int64_t)occupied.size();
  const int64_t n_ion = L.nion;
  L.ion_species.resize(n_ion);
  L.ion_style.assign(n_ion, 0);
  L.ion_data.assign(5 * n_ion, 0.0);
  L.node_ion_off.assign(L.nodes + 1, 0);
  for (int64_t q = 0; q < n_ion; ++q) {
    L.ion_species[q] = by_rank[occupied[q].second];
    L.node_ion_off[occupied[q].first + 1] += 1;
  }
  for (int64_t n = 0; n < L.nodes; ++n) L.node_ion_off[n + 1] += L.node_ion_off[n];
  // 2. initial values: the species' defaults (a Nernst species' reversal from its
  // concentrations), then ion parameters of mechanisms, then section values
  for (int64_t q = 0; q < n_ion; ++q) {
    const Species& s = sp[L.ion_species[q]];
    L.ion_data[MIND_ION_E * n_ion + q] =
        s.nernst ? mind_nernst(s.internal, s.external, s.valence, L.celsius) : s.reversal;
    L.ion_data[MIND_ION_CI * n_ion + q] = s.internal;
    L.ion_data[MIND_ION_CO * n_ion + q] = s.external;
    L.ion_style[q] = s.nernst ? ion_style_bits(1, 2, true, true, false) : 0;
  }
  for (const IonParam& ip : params) {
    const MechGroup& G = L.groups[ip.group];
    if (G.node[ip.row] < 0) continue;
    const std::vector<nmodl::IonUse>& uses = mech_ions(G.mech);
    for (size_t s = 0; s < uses.size(); ++s)
      for (int k = 0; k < 4; ++k) {
        if (uses[s].read[k] != ip.field && uses[s].write[k] != ip.field) continue;
        if ((k == MIND_ION_CI || k == MIND_ION_CO) && ip.value < 0.0)
          return "negative ion concentration of " + uses[s].name;
        L.ion_data[k * n_ion + L.ion_row_at(G.node[ip.row], slot_species[ip.group][s])] = ip.value;
      }
  }
  for (int64_t q = 0; q < d.n_ion_settings; ++q) {
    const int what = d.ionset_what[q];
    if (what < 0 || what > 4) return "unknown ion setting";
    if (what > 2) continue;
    const double v = d.ionset_value[q];
    if (!isfinite(v) || (what < 2 && v < 0.0)) return what < 2 ? "negative ion concentration" : "non-finite ion value";
    const int column = what == 0 ? MIND_ION_CI : what == 1 ? MIND_ION_CO : MIND_ION_E;
    segments(d.ionset_cell[q], d.ionset_section[q], seg);
    for (int64_t n : seg) L.ion_data[column * n_ion + L.ion_row_at(n, d.ionset_species[q])] = v;
  }
  // 3. transfers and styles; one writer per concentration and node
  std::vector<std::tuple<int64_t, int32_t, int>> owners;
  for (size_t g = 0; g < L.groups.size(); ++g) {
    MechGroup& G = L.groups[g];
    const std::vector<nmodl::IonUse>& uses = mech_ions(G.mech);
    G.nslot = (int32_t)uses.size();
    G.ion_row.assign((size_t)G.rows * G.nslot, -1);
    G.ion_bind.clear();
    if (!G.nslot) continue;
    for (int64_t r = 0; r < G.rows; ++r)
      for (int s = 0; s < G.nslot; ++s)
        G.ion_row[(size_t)r * G.nslot + s] = (int32_t)L.ion_row_at(G.node[r], slot_species[g][s]);
    int ncur = 0;
    std::vector<IonBind> reads, writes, recv;
    const std::vector<int>& rf = mech_receive_fields(G.mech);
    for (int s = 0; s < G.nslot; ++s) {
      const nmodl::IonUse& u = uses[s];
      for (int k = 0; k < 4; ++k) {
        if (u.read[k] >= 0) reads.push_back({u.read[k], k, s, 0});
        if (u.read[k] >= 0 && (k == MIND_ION_CI || k == MIND_ION_CO) &&
            std::find(rf.begin(), rf.end(), u.read[k]) != rf.end())
          recv.push_back({u.read[k], k, s, mech_state_column(G.mech, u.read[k]) ? 1 : 0});
      }
      for (int k = 0; k < 4; ++k)
        if (u.write[k] >= 0) {
          writes.push_back({u.write[k], k, s, k == MIND_ION_I ? ncur : -1});
          if (k != MIND_ION_I) L.ion_ordered = true;
        }
      ncur += u.write[MIND_ION_I] >= 0;
      const bool wconc = u.write[MIND_ION_CI] >= 0 || u.write[MIND_ION_CO] >= 0;
      const int cstyle = wconc ? 3 : (u.read[MIND_ION_CI] >= 0 || u.read[MIND_ION_CO] >= 0) ? 1 : 0;
      const int estyle = u.write[MIND_ION_E] >= 0 ? 3 : u.read[MIND_ION_E] >= 0 ? 1 : 0;
      for (int64_t r = 0; r < G.rows; ++r) {
        const int32_t q = G.ion_row[(size_t)r * G.nslot + s];
        if (q < 0) continue;
        L.ion_style[q] = promote_ion_style(L.ion_style[q], cstyle, estyle);
        for (int k : {MIND_ION_CI, MIND_ION_CO})
          if (u.write[k] >= 0) owners.push_back({G.node[r], slot_species[g][s], k});
      }
    }
    G.nread = (int32_t)reads.size();
    G.nwrite = (int32_t)writes.size();
    G.nrecv = (int32_t)recv.size();
    G.ion_bind = reads;
    G.ion_bind.insert(G.ion_bind.end(), writes.begin(), writes.end());
    G.ion_bind.insert(G.ion_bind.end(), recv.begin(), recv.end());
  }
  std::sort(owners.begin(), owners.end());
  for (size_t k = 1; k < owners.size(); ++k)
    if (owners[k] == owners[k - 1])
      return "multiple writers of " + sp[std::get<1>(owners[k])].name +
             (std::get<2>(owners[k]) == MIND_ION_CI ? "i" : "o") + " at node " + std::to_string(std::get<0>(owners[k]));
  // a Nernst species updates its reversal at initialization and every step unless a
  // mechanism owns the reversal (model.jl:841-847)
  for (int64_t q = 0; q < n_ion; ++q)
    if (sp[L.ion_species[q]].nernst && ((L.ion_style[q] >> 3) & 0x03) != 0x03) L.ion_style[q] |= 0x60;
  // section overrides, the last one per (cell, section, species) (ion_style_override,
  // utils/biophysics.jl:24-29)
  std::map<std::tuple<int, int, int>, std::pair<int, double>> over;
  for (int64_t q = 0; q < d.n_ion_settings; ++q)
    if (d.ionset_what[q] >= 3)
      over[{d.ionset_cell[q], d.ionset_section[q], d.ionset_species[q]}] = {d.ionset_what[q], d.ionset_value[q]};
  for (const auto& [key, o] : over) {
    segments(std::get<0>(key), std::get<1>(key), seg);
    for (int64_t n : seg) {
      const int64_t q = L.ion_row_at(n, std::get<2>(key));
      if (o.first == 4) {
        const int b = (int)o.second;
        if (!(o.second >= 0.0 && o.second < 128.0 && o.second == (double)b)) return "invalid ion style";
        L.ion_style[q] = (uint8_t)b;
      } else {
        const int policy = (int)o.second;
        if (!(policy >= 0 && policy <= 4 && o.second == (double)policy)) return "unknown ion reversal policy";
        static const uint8_t flags[5] = {0, 0x00, 0x20, 0x40, 0x60};
        if (policy > 0) L.ion_style[q] = (uint8_t)((L.ion_style[q] & 0x1f) | flags[policy]);
      }
    }
  }
  return "";
}

int64_t Layout::location_node(int32_t cell, int32_t section, double x) const {
  return cell_off[cell] + template_location(templates[cell_template[cell]], section, x);
}

int64_t Layout::range_node(int32_t cell, int32_t section, double x) const {
  return cell_off[cell] + template_range(templates[cell_template[cell]], section, x);
}

std::string pack(const mind_model_desc& d, double dt, uint64_t seed, Layout& L) {
  L = Layout();
  L.dt = dt;
  L.seed = seed;
  L.celsius = d.celsius;
  if (!(dt > 0.0) || !isfinite(dt)) return "dt_micro must be positive";
  // 1. Morphology templates.
  L.templates.resize(d.n_morphologies);
  for (int m = 0; m < d.n_morphologies; ++m) {
    std::string err = build_template(d, m, L.templates[m]);
    if (!err.empty()) return err;
  }
  // 2. Cells as contiguous node blocks.
  L.ncells = d.n_cells;
  L.cell_template.assign(d.cell_morphology, d.cell_morphology + d.n_cells);
  L.cell_off.assign(d.n_cells + 1, 0);
  for (int c = 0; c < d.n_cells; ++c) {
    const int m = d.cell_morphology[c];
    if (m < 0 || m >= d.n_morphologies) return "cell has an unknown morphology";
    L.cell_off[c + 1] = L.cell_off[c] + (int64_t)L.templates[m].parent.size();
    L.compartments += L.templates[m].compartments;
  }
  const int64_t N = L.cell_off[d.n_cells];
  if (N >= INT32_MAX) return "packed node count exceeds Int32 range";
  L.nodes = N;
  L.parent.resize(N); L.area.resize(N); L.cap.resize(N); L.axial.resize(N); L.v_init.resize(N);
  for (int c = 0; c < d.n_cells; ++c) {
    const Template& t = L.templates[L.cell_template[c]];
    const int64_t o = L.cell_off[c];
    for (size_t i = 0; i < t.parent.size(); ++i) {
      L.parent[o + i] = t.parent[i] < 0 ? -1 : (int32_t)(o + t.parent[i]);
      L.area[o + i] = t.area[i];
      L.cap[o + i] = t.cap[i];
      L.axial[o + i] = t.axial[i];
      L.v_init[o + i] = d.cell_voltage[c];
    }
  }
  // axial diagonal: own conductance, then children in discretization order
  L.adiag.assign(N, 0.0);
  L.child_off.assign(N + 1, 0);
  for (int64_t i = 0; i < N; ++i)
    if (L.parent[i] >= 0) {
      L.adiag[i] += L.axial[i];
      L.adiag[L.parent[i]] += L.axial[i];
      L.child_off[L.parent[i] + 1] += 1;
    }
  for (int64_t i = 0; i < N; ++i) L.child_off[i + 1] += L.child_off[i];
  L.child.resize(L.child_off[N]);
  {
    std::vector<int64_t> cur(L.child_off.begin(), L.child_off.end() - 1);
    for (int64_t i = 0; i < N; ++i)
      if (L.parent[i] >= 0) L.child[cur[L.parent[i]]++] = (int32_t)i;
  }
  // 3. Mechanism groups: first use over density insertions, then point processes.
  for (int k = 0; k < MIND_MECH_SLOTS; ++k) L.group_of_mech[k] = -1;
  struct Pending { int64_t p0, p1; };
  std::vector<std::vector<Pending>> params;
  auto group_of = [&](int mech) {
    if (L.group_of_mech[mech] < 0) {
      L.group_of_mech[mech] = (int)L.groups.size();
      L.groups.emplace_back();
      L.groups.back().mech = mech;
      params.emplace_back();
    }
    return L.group_of_mech[mech];
  };
  std::vector<std::vector<uint8_t>> seen(MIND_MECH_SLOTS);
  for (int64_t k = 0; k < d.n_density; ++k) {
    const int mech = d.dens_mech[k];
    if (mech < 0 || mech >= MIND_MECH_SLOTS || !mech_known(mech) || mech_is_point(mech))
      return "cell insertion requires a distributed micro mechanism";
    const int c = d.dens_cell[k];
    if (c < 0 || c >= d.n_cells) return "density insertion on an unknown cell";
    const Template& t = L.templates[L.cell_template[c]];
    const int sec = d.dens_section[k];
    if (sec < 0 || sec >= (int)t.sec_nodes.size()) return "unknown section";
    const int g = group_of(mech);
    if (seen[mech].empty()) seen[mech].assign(N, 0);
    const std::vector<int32_t>& nodes = t.sec_nodes[sec];
    for (size_t j = 0; j + 1 < nodes.size(); ++j) {
      const int64_t n = L.cell_off[c] + nodes[j];
      if (seen[mech][n]) continue;
      seen[mech][n] = 1;
      L.groups[g].node.push_back((int32_t)n);
      L.groups[g].point.push_back(-1);
      L.groups[g].logical.push_back((uint64_t)n + 1);
      params[g].push_back({d.dens_param_offsets[k], d.dens_param_offsets[k + 1]});
    }
  }
  L.point_group.assign(d.n_points, -1);
  L.point_row.assign(d.n_points, -1);
  L.point_cell.assign(d.n_points, -1);
  std::vector<std::pair<double, int32_t>> pattern;  // (time, point)
  for (int64_t p = 0; p < d.n_points; ++p) {
    const int mech = d.pt_mech[p];
    if (mech < 0 || mech >= MIND_MECH_SLOTS || !mech_is_point(mech))
      return "point creation requires a point process";
    const int64_t s0 = d.pt_spike_offsets ? d.pt_spike_offsets[p] : 0;
    const int64_t s1 = d.pt_spike_offsets ? d.pt_spike_offsets[p + 1] : 0;
    if (s1 > s0 && mech != MIND_MECH_PATTERNSTIM) return "spike_times applies to PatternStim only";
    if (mech == MIND_MECH_PATTERNSTIM) {  // no mechanism data: its spikes enter at initialize
      if (d.pt_cell[p] >= 0) return "PatternStim is an artificial cell without a location";
      for (int64_t k = s0; k < s1; ++k) {
        const double t = d.pt_spike_times[k];
        if (!(isfinite(t) && t >= 0.0)) return "invalid spike times";
        if (k > s0 && t < d.pt_spike_times[k - 1]) return "spike times must be sorted";
        pattern.push_back({t, (int32_t)p});
      }
      continue;
    }
    const int g = group_of(mech);
    int64_t n = -1;
    const int c = d.pt_cell[p];
    if (mech == MIND_MECH_NETSTIM) {
      if (c >= 0) return "NetStim is an artificial cell without a location";
    } else {
      if (c < 0 || c >= d.n_cells) return "point process needs a cell location";
      const Template& t = L.templates[L.cell_template[c]];
      if (d.pt_section[p] < 0 || d.pt_section[p] >= (int)t.sec_nodes.size()) return "unknown section";
      if (!(d.pt_x[p] >= 0.0 && d.pt_x[p] <= 1.0)) return "section location must be in [0, 1]";
      n = L.location_node(c, d.pt_section[p], d.pt_x[p]);
      L.point_cell[p] = c;
    }
    L.groups[g].node.push_back((int32_t)n);
    L.groups[g].point.push_back((int32_t)p);
    L.groups[g].logical.push_back((uint64_t)(p + 1) + (uint64_t)N);
    L.point_group[p] = g;
    params[g].push_back({d.pt_param_offsets[p], d.pt_param_offsets[p + 1]});
  }
  // Sort rows by node (stable) and fill parameter columns.
  std::vector<IonParam> ion_params;
  for (size_t g = 0; g < L.groups.size(); ++g) {
    MechGroup& G = L.groups[g];
    const int64_t rows = (int64_t)G.node.size();
    G.rows = rows;
    std::vector<int64_t> perm(rows);
    std::iota(perm.begin(), perm.end(), 0);
    if (G.mech != MIND_MECH_NETSTIM)
      std::stable_sort(perm.begin(), perm.end(),
                       [&](int64_t a, int64_t b) { return G.node[a] < G.node[b]; });
    MechGroup H;
    H.mech = G.mech;
    H.rows = rows;
    H.node.resize(rows); H.point.resize(rows); H.logical.resize(rows);
    const int nf = field_count(G.mech);
    std::vector<double> def(std::max(nf, 1));
    field_defaults(G.mech, def.data());
    H.data.resize((size_t)nf * rows);
    for (int f = 0; f < nf; ++f) std::fill(H.data.begin() + f * rows, H.data.begin() + (f + 1) * rows, def[f]);
    for (int64_t r = 0; r < rows; ++r) {
      const int64_t o = perm[r];
      H.node[r] = G.node[o];
      H.point[r] = G.point[o];
      H.logical[r] = G.logical[o];
      for (int64_t q = params[g][o].p0; q < params[g][o].p1; ++q) {
        const int f = d.param_field[q];
        if (f < 0 || f >= nf) return "unknown mechanism field";
        H.data[(size_t)f * rows + r] = d.param_value[q];
        ion_params.push_back({(int)g, r, f, d.param_value[q]});
      }
      if (H.point[r] >= 0) L.point_row[H.point[r]] = (int32_t)r;
    }
    if (G.mech != MIND_MECH_NETSTIM) {
      H.cell_rows.assign(L.ncells + 1, 0);
      int64_t r = 0;
      for (int c = 0; c < L.ncells; ++c) {
        H.cell_rows[c] = r;
        while (r < rows && H.node[r] < L.cell_off[c + 1]) ++r;
      }
      H.cell_rows[L.ncells] = rows;
      H.run_start.assign(rows, 1);
      for (int64_t q = 1; q < rows; ++q) H.run_start[q] = H.node[q] != H.node[q - 1];
    }
    L.groups[g] = std::move(H);
  }
  {
    const std::string e = pack_ions(d, ion_params, L);
    if (!e.empty()) return e;
  }
  // 4. Sources, detectors (one per voltage source, in source order) and fanout.
  L.nsources = d.n_sources;
  L.source_sid.assign(d.src_sid, d.src_sid + d.n_sources);
  L.source_point.assign(d.n_sources, -1);
  L.source_external.assign(d.n_sources, 0);
  std::vector<int32_t> det_of_cell_count(L.ncells + 1, 0);
  for (int64_t s = 0; s < d.n_sources; ++s) {
    if (d.src_kind[s] == 0) {
      const int c = d.src_cell[s];
      if (c < 0 || c >= L.ncells) return "detector on an unknown cell";
      if (!isfinite(d.src_threshold[s])) return "invalid spike detector";
      L.det_node.push_back((int32_t)L.location_node(c, d.src_section[s], d.src_x[s]));
      L.det_source.push_back((int32_t)s);
      L.det_cell.push_back(c);
      L.det_threshold.push_back(d.src_threshold[s]);
      det_of_cell_count[c + 1] += 1;
    } else if (d.src_kind[s] == 1) {
      const int p = d.src_point[s];
      if (p < 0 || p >= d.n_points) return "event source on an unknown point process";
      L.source_point[s] = p;
    } else if (d.src_kind[s] == 2) {
      L.source_external[s] = 1;
    } else {
      return "unknown spike-source kind";
    }
  }
  L.cell_det_off.assign(L.ncells + 1, 0);
  for (int c = 0; c < L.ncells; ++c) L.cell_det_off[c + 1] = L.cell_det_off[c] + det_of_cell_count[c + 1];
  L.cell_det.resize(L.det_node.size());
  {
    std::vector<int64_t> cur(L.cell_det_off.begin(), L.cell_det_off.end() - 1);
    for (size_t k = 0; k < L.det_node.size(); ++k) L.cell_det[cur[L.det_cell[k]]++] = (int32_t)k;
  }
  L.fanout.assign(d.n_sources + 1, 0);
  for (int64_t e = 0; e < d.n_connections; ++e) {
    if (d.con_source[e] < 0 || d.con_source[e] >= d.n_sources) return "connection from an unknown source";
    if (d.con_target[e] < 0 || d.con_target[e] >= d.n_points) return "connection to an unknown point";
    const int tm = d.pt_mech[d.con_target[e]];
    if (!mech_has_receive(tm)) return "target has no NET_RECEIVE block";
    if (!(d.con_delay[e] >= 0.0) || !isfinite(d.con_delay[e]) || !isfinite(d.con_weight[e]))
      return "connection weight and delay must be finite, delay nonnegative";
    L.fanout[d.con_source[e] + 1] += 1;
  }
  for (int64_t s = 0; s < d.n_sources; ++s) L.fanout[s + 1] += L.fanout[s];
  L.edge_target.resize(d.n_connections);
  L.edge_delay.resize(d.n_connections);
  L.edge_weight.resize(d.n_connections);
  L.connection_edge.resize(d.n_connections);
  {
    std::vector<int64_t> cur(L.fanout.begin(), L.fanout.end() - 1);
    for (int64_t e = 0; e < d.n_connections; ++e) {
      const int64_t k = cur[d.con_source[e]]++;
      L.connection_edge[e] = k;
      L.edge_target[k] = d.con_target[e];
      L.edge_delay[k] = d.con_delay[e];
      L.edge_weight[k] = d.con_weight[e];
    }
  }
  L.min_delay = INFINITY;
  L.max_delay = 0.0;
  for (double x : L.edge_delay) {
    L.min_delay = std::min(L.min_delay, x);
    L.max_delay = std::max(L.max_delay, x);
  }
  L.first_virtual = L.nsources;
  L.min_micro_delay = L.min_delay;
  std::stable_sort(pattern.begin(), pattern.end(),
                   [](const std::pair<double, int32_t>& a, const std::pair<double, int32_t>& b) {
                     return a.first < b.first;
                   });
  {
    std::vector<std::vector<int32_t>> sources_of(d.n_points);
    for (int64_t s = 0; s < d.n_sources; ++s)
      if (L.source_point[s] >= 0) sources_of[L.source_point[s]].push_back((int32_t)s);
    for (const auto& [t, p] : pattern)
      for (int32_t s : sources_of[p]) {
        L.pattern_time.push_back(t);
        L.pattern_source.push_back(s);
      }
  }
  const int gn = L.group_of_mech[MIND_MECH_NETSTIM];
  if (gn >= 0) {
    const MechGroup& G = L.groups[gn];
    L.netstim_source.assign(G.rows, -1);
    for (int64_t s = 0; s < d.n_sources; ++s)
      if (L.source_point[s] >= 0 && L.point_group[L.source_point[s]] == gn)
        L.netstim_source[L.point_row[L.source_point[s]]] = (int32_t)s;
  }
  return "";
}

std::string pack_inputs(const mind_model_desc& d, const mind_macro_desc& m, Layout& L) {
  const int64_t nb = m.input_offsets[m.n_roi];
  for (int64_t b = 0; b < nb; ++b) {
    const int32_t p = m.input_point[b];
    if (p < 0 || p >= d.n_points) return "macro2micro target is not a point process";
    const int tm = d.pt_mech[p];
    if (tm == MIND_MECH_NETSTIM || !mech_has_receive(tm)) return "macro2micro target has no NET_RECEIVE";
    if (L.point_cell[p] < 0) return "macro2micro target must sit on a cell";
    if (!(m.input_delay[b] >= 0.0) || !isfinite(m.input_delay[b]) || !isfinite(m.input_weight[b]))
      return "macro2micro weight and delay must be finite, delay nonnegative";
    L.source_sid.push_back(-(b + 1));  // virtual SIDs (model.jl:622-628)
    L.source_point.push_back(-1);
    L.fanout.push_back(L.fanout.back() + 1);
    L.edge_target.push_back(p);
    L.edge_delay.push_back(m.input_delay[b]);
    L.edge_weight.push_back(m.input_weight[b]);
    L.max_delay = std::max(L.max_delay, m.input_delay[b]);
  }
  L.nsources += nb;
  for (int64_t q = 0; q < m.output_offsets[m.n_roi]; ++q) {
    const int32_t s = m.output_source[q];
    if (s < 0 || s >= L.first_virtual) return "micro2macro source is not a spike source";
    bool det = false;
    for (size_t k = 0; k < L.det_source.size(); ++k) det |= L.det_source[k] == s;
    if (!det) return "micro2macro sources must be voltage detectors";
  }
  return "";
}

}  // namespace mind

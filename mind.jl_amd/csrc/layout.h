// Host-side packing of a flattened MicroModel into the structure-of-arrays
// layout the HIP kernels read.
//
// Node order: cells are contiguous blocks in cell order; inside a cell, nodes
// follow the discretization order of the reference (parents before children,
// src/model/micro/morphology.jl:406-484). Mechanism rows of a group are sorted
// by node (stable), so every cell owns a contiguous row range per group
// (src/model/micro/model.jl:533-555); artificial cells keep creation order.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mind.h"
#include "mech.h"

namespace mind {

struct MechGroup {
  int mech = 0;
  int64_t rows = 0;
  std::vector<int32_t> node;      // packed node, -1 for artificial cells
  std::vector<int32_t> point;     // point-process index, -1 for density rows
  std::vector<uint64_t> logical;  // logical identity (Random123 key)
  std::vector<double> data;       // field-major: data[field * rows + row]
  std::vector<int64_t> cell_rows; // rows of cell c: [cell_rows[c], cell_rows[c+1])
  std::vector<uint8_t> run_start; // point groups: row starts a run of equal nodes
  // ion transfers (USEION statements): ion_row[r * nslot + s] is row r's ion row
  // for statement s (-1 for artificial cells); ion_bind holds the nread loads,
  // the nwrite stores and the nrecv NET_RECEIVE exchanges
  int32_t nslot = 0, nread = 0, nwrite = 0, nrecv = 0;
  std::vector<int32_t> ion_row;
  std::vector<IonBind> ion_bind;
};

// IonSpecies (src/model/micro/biophysics.jl:3-25).
struct Species {
  std::string name;
  double valence = 1.0, internal = 1.0, external = 1.0, reversal = 0.0;
  bool nernst = false;
};

struct Template {  // one discretized morphology
  std::vector<int32_t> parent;  // local, -1 for the root
  std::vector<double> area, cap, axial;
  std::vector<std::vector<int32_t>> sec_nodes;  // nseg + 1 nodes per section
  std::vector<int32_t> sec_start;
  int64_t compartments = 0;
};

struct Layout {
  double dt = 0.025, celsius = 6.3;
  int32_t secondorder = 0;  // 0 backward Euler; 1, 2 Crank-Nicolson (assembly at dt/2, v += 2 dv)
  // dt of the cable assembly's C/dt term (assemble_cable_kernel!'s argument, step.jl:112)
  double assembly_dt() const { return secondorder ? 0.5 * dt : dt; }
  uint64_t seed = 1;
  int32_t ncells = 0;
  int64_t nodes = 0, compartments = 0;
  std::vector<Template> templates;
  std::vector<int32_t> cell_template;
  std::vector<int64_t> cell_off;  // ncells + 1
  std::vector<int32_t> parent;    // global index, -1 for roots
  std::vector<double> area, cap, axial, adiag, v_init;
  std::vector<int64_t> child_off;  // children CSR, increasing index
  std::vector<int32_t> child;

  std::vector<MechGroup> groups;   // canonical order: first use, densities first
  int group_of_mech[MIND_MECH_SLOTS];
  std::vector<int32_t> point_group, point_row, point_cell;

  int64_t nsources = 0;
  std::vector<int64_t> source_sid;
  std::vector<int32_t> source_point;  // point index of an event source, -1 for detectors
  std::vector<uint8_t> source_external;  // 1 for external spike inputs (inject_spike!)
  // PatternStim spikes as (time, source), in the order initialize enqueues them:
  // by time (stable over point creation), then the point's sources in order
  std::vector<double> pattern_time;
  std::vector<int32_t> pattern_source;
  std::vector<int32_t> det_node, det_source, det_cell;
  std::vector<double> det_threshold;
  std::vector<int64_t> cell_det_off;  // detectors of cell c (indices into det_*)
  std::vector<int32_t> cell_det;

  std::vector<int64_t> fanout;  // nsources + 1
  std::vector<int32_t> edge_target;
  std::vector<double> edge_delay, edge_weight;
  std::vector<int64_t> connection_edge;  // edge of each connection, in creation order
  double min_delay = 0.0, max_delay = 0.0;

  // ion rows, one per (node, species) pair in use, sorted by node and then by
  // species name (src/model/micro/model.jl:710-855); columns MIND_ION_E, _CI,
  // _CO, _I, _DIDV of row q at ion_data[column * nion + q]
  std::vector<Species> species;       // the description's order
  int64_t nion = 0;
  std::vector<int32_t> ion_species;
  std::vector<uint8_t> ion_style;     // IonStyle bits (ion_style_bits)
  std::vector<double> ion_data;
  std::vector<int64_t> node_ion_off;  // nodes + 1
  bool ion_ordered = false;           // some mechanism writes a reversal or a concentration
  int64_t ion_row_at(int64_t node, int32_t species) const;

  std::vector<int32_t> netstim_source;  // per NetStim row: its source index or -1
  // MACRO2MICRO bindings are virtual sources [first_virtual, nsources), one edge each
  int64_t first_virtual = 0;
  double min_micro_delay = 0.0;  // over edges of real sources only

  int64_t location_node(int32_t cell, int32_t section, double x) const;
  int64_t range_node(int32_t cell, int32_t section, double x) const;
};

// Returns an empty string on success, else the error message (ArgumentError).
std::string pack(const mind_model_desc& d, double dt, uint64_t seed, Layout& out);
// Appends the virtual sources of MACRO2MICRO bindings (src/model/build.jl:261-295).
std::string pack_inputs(const mind_model_desc& d, const mind_macro_desc& m, Layout& L);

int field_count(int mech);
void field_defaults(int mech, double* values);
bool mech_is_point(int mech);

}  // namespace mind

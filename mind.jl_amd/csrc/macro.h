// Device and host state of the macro (ROI) layer; see macro.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.h"
#include "macro_gen.h"

namespace mind {

struct MacroDev {
  int32_t n;
  int32_t ratio;
  double dt_macro;
  const double* W;           // [target * n + source]
  const int64_t* dsteps;     // delay in macro steps
  int64_t cap;
  double* hist;              // [slot * n + roi]
  const double* wc;          // [roi * WC_NPARAMS + k]
  double *E, *I, *coupling, *rate, *acc, *meanv;
  const int64_t* bin_steps;
  const int64_t* out_off;
  const int32_t* out_node;
  const int32_t* src_roi;    // per source, -1 if no MICRO2MACRO reads it
  int64_t nb, first_virtual;
  const int32_t* in_roi;     // per binding
  const uint64_t* stream_id; // per binding
  int64_t* in_count;
  int64_t* in_base;
  double* trace;             // [5][n][samples]
  int64_t samples, trace_start;
  // windows of several macro steps (one population-rate bin): macro steps per
  // full window, the E, I and coupling of each step of a window ([3][kwin][n]),
  // and the outputs' soma-voltage snapshots ([kwin][n_out], filled by the cells)
  int32_t kwin;
  double* traj;
  const double* snap;
  int64_t n_out;
};

// The macro step's MIND-role mechanisms, lowered by nmodlc from
// mind.jl_amd/mod/{WilsonCowan,LinearDelay,PoissonDrive,PopRate}.mod
// (macro_gen.h). The kernels call them through these wrappers, each on one row
// held in registers.
static_assert(gen_macro::WilsonCowan_tau_e == WC_TAU_E && gen_macro::WilsonCowan_tau_i == WC_TAU_I &&
                  gen_macro::WilsonCowan_c_ee == WC_C_EE && gen_macro::WilsonCowan_c_ei == WC_C_EI &&
                  gen_macro::WilsonCowan_c_ie == WC_C_IE && gen_macro::WilsonCowan_c_ii == WC_C_II &&
                  gen_macro::WilsonCowan_a_e == WC_A_E && gen_macro::WilsonCowan_theta_e == WC_THETA_E &&
                  gen_macro::WilsonCowan_a_i == WC_A_I && gen_macro::WilsonCowan_theta_i == WC_THETA_I &&
                  gen_macro::WilsonCowan_P == WC_P && gen_macro::WilsonCowan_Q == WC_Q &&
                  gen_macro::WilsonCowan_k_micro == WC_K_MICRO && gen_macro::WilsonCowan_rate_scale == WC_RATE_SCALE &&
                  gen_macro::WilsonCowan_E0 == WC_E0 && gen_macro::WilsonCowan_I0 == WC_I0,
              "WilsonCowan.mod's parameters must be the WC_* columns in order");

// REGION: one macro step of length dt of the ROI whose parameters are p (WC_*
// order), from (e, i), with the delayed coupling and the micro population rate
// (Hz) as inputs.
MIND_HD void region_advance(const double* p, double& e, double& i, double coupling, double rate, double dt) {
  using namespace gen_macro;
  double f[WilsonCowan_NFIELDS];
  for (int k = 0; k < WC_NPARAMS; ++k) f[k] = p[k];
  f[WilsonCowan_E] = e;
  f[WilsonCowan_I] = i;
  f[WilsonCowan_coupling] = coupling;
  f[WilsonCowan_rate] = rate;
  f[WilsonCowan_dE] = 0.0;
  f[WilsonCowan_dI] = 0.0;
  WilsonCowan_advance(Cols{f, 1}, 0, 0.0, 0.0, dt, 0.0);
  e = f[WilsonCowan_E];
  i = f[WilsonCowan_I];
}

// MACRO2MACRO: an edge's contribution to its target's coupling, from the
// source's delayed E.
MIND_HD double edge_value(double weight, double source_e) {
  using namespace gen_macro;
  double f[LinearDelay_NFIELDS] = {0.0};
  f[LinearDelay_weight] = weight;
  f[LinearDelay_x] = source_e;
  LinearDelay_evaluate(Cols{f, 1}, 0, 0.0, 0.0);
  return f[LinearDelay_y];
}

// MACRO2MICRO: the Poisson rate (Hz) of a binding from its ROI's E and rate_scale.
MIND_HD double drive_rate(double rate_scale, double e) {
  using namespace gen_macro;
  double f[PoissonDrive_NFIELDS] = {0.0};
  f[PoissonDrive_e] = e;
  f[PoissonDrive_scale] = rate_scale;
  PoissonDrive_evaluate(Cols{f, 1}, 0, 0.0, 0.0);
  return f[PoissonDrive_rate];
}

// MICRO2MACRO at a bin end: the population rate (Hz) of `count` spikes of
// `sources` cells over `width` ms. PopRate's NET_RECEIVE adds one per spike; the
// kernels add a macro step's spikes at once, the same exact integer sum.
MIND_HD double bin_rate(double count, double sources, double width) {
  using namespace gen_macro;
  double f[PopRate_NFIELDS] = {0.0};
  f[PopRate_count] = count;
  f[PopRate_sources] = sources;
  f[PopRate_width] = width;
  PopRate_evaluate(Cols{f, 1}, 0, 0.0, 0.0);
  return f[PopRate_rate];
}

struct MacroHost {
  MacroDev m{};
  bool bins = false;  // windows of one bin (kwin macro steps) instead of one macro step
  std::vector<void*> allocations;
  std::vector<double> wc, E0, I0;
  hipGraphExec_t graph = nullptr;
};

}  // namespace mind

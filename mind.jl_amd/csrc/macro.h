// Device and host state of the macro (ROI) layer; see macro.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.h"

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

// Wilson-Cowan REGION: S(x) = 1/(1 + exp(-a (x - theta))) - 1/(1 + exp(a theta)).
__host__ __device__ inline double wc_sigmoid(double x, double a, double theta) {
  return 1.0 / (1.0 + mind_exp(-a * (x - theta))) - 1.0 / (1.0 + mind_exp(a * theta));
}

__host__ __device__ inline void wc_deriv(const double* p, double e, double i, double u, double& de, double& di) {
  de = (-e + wc_sigmoid(p[WC_C_EE] * e - p[WC_C_EI] * i + p[WC_P] + u, p[WC_A_E], p[WC_THETA_E])) /
       p[WC_TAU_E];
  di = (-i + wc_sigmoid(p[WC_C_IE] * e - p[WC_C_II] * i + p[WC_Q], p[WC_A_I], p[WC_THETA_I])) /
       p[WC_TAU_I];
}

__host__ __device__ inline void wc_rk4(const double* p, double& e, double& i, double u, double h) {
  double k1e, k1i, k2e, k2i, k3e, k3i, k4e, k4i;
  wc_deriv(p, e, i, u, k1e, k1i);
  wc_deriv(p, e + 0.5 * h * k1e, i + 0.5 * h * k1i, u, k2e, k2i);
  wc_deriv(p, e + 0.5 * h * k2e, i + 0.5 * h * k2i, u, k3e, k3i);
  wc_deriv(p, e + h * k3e, i + h * k3i, u, k4e, k4i);
  e = e + h / 6.0 * (k1e + 2.0 * k2e + 2.0 * k3e + k4e);
  i = i + h / 6.0 * (k1i + 2.0 * k2i + 2.0 * k3i + k4i);
}

struct MacroHost {
  MacroDev m{};
  bool bins = false;  // windows of one bin (kwin macro steps) instead of one macro step
  std::vector<void*> allocations;
  std::vector<double> wc, E0, I0;
  hipGraphExec_t graph = nullptr;
};

}  // namespace mind

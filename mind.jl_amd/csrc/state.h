// Host-side state of one simulation (owned by the C ABI handle).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device.h"

namespace mind {

int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) return ::mind::fail(MIND_EHIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

struct BatchPlan;
struct MacroHost;
struct State;

// Kernels that run mechanism phases. A model with runtime-compiled mechanisms
// launches them from a module that hipRTC builds for its set of mechanisms
// (rtc.hip); every other model launches the library's own.
enum KernelId {
  K_PLAIN, K_INIT_GROUP, K_CURRENTS_FIELDS, K_WINDOW, K_WINDOW_LEAN, K_WINDOW_LEAN_HBM, K_WINDOW_LARGE,
  K_SERIAL_STEP, K_COUNT
};
struct KernelSet {
  hipModule_t module = nullptr;
  hipFunction_t fn[K_COUNT] = {};
  int max_groups = 0;  // mechanism groups its batched kernels cache in LDS
};
void batched_release(State& S);
void macro_release(State& S);

struct State {
  Layout L;
  Dev d{};
  int device = 0;
  int method = MIND_METHOD_PLAIN;
  // every mechanism group has an inlined path in the batched kernels (PAS, HH with
  // at most its own ion currents, EXPSYN, EXP2SYN, ICLAMP, NETSTIM), no ion
  // orders, no secondorder-2 ion correction: the *_core window kernels run it
  bool core = false;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;          // concurrent launches of oversized cells
  hipEvent_t fork = nullptr, join = nullptr;
  std::vector<void*> allocations;
  std::vector<GroupDev> groups_host;
  std::vector<double*> group_initial;  // parameter snapshot restored by initialize
  double* ion_initial = nullptr;       // packed ion rows, restored by initialize
  double* v_init = nullptr;
  uint64_t* ns_logical = nullptr;
  int32_t wmax = 1;
  int32_t window_steps_opt = 0;  // build option window_steps (0: as long as the model allows)
  int64_t step = 0;
  bool initialized = false;
  hipGraphExec_t graph = nullptr;
  int32_t graph_steps = 0;
  // batched kernel plan (cell groups, chain units and level schedules)
  BatchPlan* batch = nullptr;
  MacroHost* macro = nullptr;
  const KernelSet* ks = nullptr;  // runtime-compiled kernels (owned by the module cache), or null
  // event-exact path (delays under 1.5 dt or connections into NetStims)
  bool serial = false;
  SerialDev q{};
  double* pattern_time = nullptr;
  int32_t* pattern_source = nullptr;
  std::vector<int32_t> cell_bucket;  // event bucket of each cell: the cell, or its group
  // injected input not yet on the device: connection events beyond the delivery
  // ring's horizon, source spikes due in a later window (inject_event!, inject_spike!)
  std::vector<QEvent> pending_events;
  std::vector<InjSpike> pending_spikes;
  QEvent* inj_ev = nullptr;  // device staging of injected events (kInjCap)
  std::vector<void*> run_allocations;
  // optional timing of the dominant cell-kernel launch (HIP events on `stream`)
  bool profiling = false;
  std::vector<hipEvent_t> events;  // start/stop pairs of the current run
  size_t events_used = 0;
  double kernel_ms = 0.0;
  int64_t kernel_launches = 0;
  int64_t kernel_steps = 0;

  ~State();
  void release() {
    macro_release(*this);
    batched_release(*this);
    for (hipEvent_t e : events) hipEventDestroy(e);
    if (graph) hipGraphExecDestroy(graph);
    for (void* p : run_allocations) hipFree(p);
    for (void* p : allocations) hipFree(p);
    if (fork) hipEventDestroy(fork);
    if (join) hipEventDestroy(join);
    if (aux) hipStreamDestroy(aux);
    if (stream) hipStreamDestroy(stream);
  }

  template <typename T>
  int alloc(T** out, size_t count) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(MIND_EOOM, "device allocation failed");
    allocations.push_back(p);
    *out = static_cast<T*>(p);
    return 0;
  }

  template <typename T>
  int upload(T** out, const std::vector<T>& host) {
    int rc = alloc(out, host.size());
    if (rc) return rc;
    if (!host.empty())
      HIP_TRY(hipMemcpy(*out, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
  }
};


// Launch of kernel `id`: `fn` (the library's) or the state's module function.
hipError_t launch_kernel(const State& S, int id, const void* fn, dim3 grid, dim3 block, size_t shmem,
                         hipStream_t stream, void** args);
template <typename... A>
hipError_t launch(const State& S, int id, const void* fn, dim3 grid, dim3 block, size_t shmem, hipStream_t stream,
                  A... a) {
  void* args[] = {(void*)&a...};
  return launch_kernel(S, id, fn, grid, block, shmem, stream, args);
}
// Builds (or takes from the process-wide cache) the module of the state's
// runtime-compiled mechanisms; leaves S.ks null when the model has none.
int rtc_kernels(State& S);

int launch_window(State& S, int32_t n);
int stage_injections(State& S, int32_t n);
int setup_state(State& S, const mind_build_options& o);
int check_device_error(State& S);
int collect_spikes(State& S, int64_t* n, double** times, int64_t** sids);
int macro_setup(State& S, const mind_macro_desc& md);
int macro_initialize(State& S);
int launch_macro_window(State& S, int32_t K);
int timer_start(State& S);
int timer_stop(State& S);
int batched_setup(State& S, int32_t cells_per_group);
int batched_launch(State& S, int32_t n);
int cells_launch(State& S, int32_t n);
int serial_setup(State& S, int64_t capacity);
int serial_initialize(State& S);
int serial_launch(State& S, int32_t n);
int serial_inject(State& S, const SEvent& e);

}  // namespace mind

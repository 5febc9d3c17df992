// Runtime compilation of the cell kernels for models with runtime-compiled
// mechanisms (mind_compile_mod).
//
// The reference generates device code per model from its compiled MOD files
// (src/compiler/codegen.jl) and lets its GPU toolchain compile it at first use.
// Here the kernels that run mechanism phases (cell_kernels.h,
// batched_kernels.h, serial_kernels.h) are compiled once more with hipRTC,
// from the same headers embedded in the library, with user_mechs.h holding the
// lowering of the model's compiled mechanisms and a dispatch over their slots.
// The module is cached per device and mechanism set for the life of the
// process; all other kernels (exchange, gather, queues) stay the library's.
// Builtin phases compile to the same IEEE operations in both (the same source
// and -ffp-contract=off), so a model gives the same results either way.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mechreg.h"
#include "state.h"

namespace mind {

namespace {

struct Embedded {
  const char* name;
  const char* text;
};

// The device headers of the library, as built (rtc_sources.inc is generated
// from them by the Makefile).
const Embedded kHeaders[] = {
#include "rtc_sources.inc"
};

const char* const kKernelNames[K_COUNT] = {
    "&mind::k_plain",       "&mind::k_init_group",      "&mind::k_currents_fields", "&mind::k_window",
    "&mind::k_window_lean", "&mind::k_window_lean_hbm", "&mind::k_window_large",    "&mind::k_serial_step"};

std::mutex g_mu;
std::map<std::string, KernelSet*> g_modules;  // key: device, program and user header text

// user_mechs.h: the lowering of each slot, then mind::user's dispatch over them.
std::string user_header(const std::vector<int>& slots) {
  std::string h = "// Runtime-compiled mechanisms of this module.\n#pragma once\n";
  for (int id : slots) h += user_slot(id)->lowering;
  auto cases = [&](const char* phase, const char* args, bool only_state) {
    std::string c;
    for (int id : slots) {
      if (only_state && !user_slot(id)->m.has_state) continue;
      c += "    case " + std::to_string(id) + ": u" + std::to_string(id) + phase + args + "; break;\n";
    }
    return c;
  };
  h += "namespace mind {\nnamespace user {\n";
  h += "MIND_HD void current(int mech, const Cols& c, int64_t r, double v, double t, double dt, double& value, "
       "double& shifted, bool store, double celsius, double* ish) {\n  switch (mech) {\n" +
       cases("_current", "(c, r, v, t, dt, value, shifted, store, celsius, ish)", false) +
       "    default: value = shifted = 0.0;\n  }\n}\n";
  h += "MIND_HD void advance(int mech, const Cols& c, int64_t r, double v, double t, double dt, double celsius) {\n"
       "  switch (mech) {\n" +
       cases("_advance", "(c, r, v, t, dt, celsius)", true) + "    default: break;\n  }\n}\n";
  h += "MIND_HD void initial(int mech, const Cols& c, int64_t r, double v, double celsius) {\n  switch (mech) {\n" +
       cases("_initial", "(c, r, v, celsius, nullptr, 0)", false) + "    default: break;\n  }\n}\n";
  h += "MIND_HD void receive(int mech, const Cols& c, int64_t r, double t, double weight) {\n  switch (mech) {\n" +
       cases("_receive", "(c, r, t, weight, 0.0, nullptr, 0)", false) + "    default: break;\n  }\n}\n";
  h += "MIND_HD bool has_state(int mech) {\n  switch (mech) {\n";
  for (int id : slots)
    if (user_slot(id)->m.has_state) h += "    case " + std::to_string(id) + ":\n";
  h += "      return true;\n    default: return false;\n  }\n}\n";
  h += "}  // namespace user\n}  // namespace mind\n";
  return h;
}

// Compiles the cell kernels of a mechanism set for `arch`; returns 0 and the code
// object in *code, else the MIND_E* code with the message set.
int compile_module(const std::vector<int>& slots, int max_groups, const std::string& arch, std::vector<char>* code,
                   std::string* lowered) {
  std::string program = "#define MIND_USER_MECHANISMS 1\n#define MIND_MAX_GROUPS " + std::to_string(max_groups) + "\n";
#ifdef MIND_GENERATED_MECHANISMS
  program += "#define MIND_GENERATED_MECHANISMS 1\n";
#endif
  program += "#include \"cell_kernels.h\"\n#include \"serial_kernels.h\"\n#include \"batched_kernels.h\"\n";
  const std::string uh = user_header(slots);
  std::vector<const char*> names, texts;
  for (const Embedded& e : kHeaders) {
    names.push_back(e.name);
    texts.push_back(e.text);
  }
  names.push_back("user_mechs.h");
  texts.push_back(uh.c_str());
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, program.c_str(), "mind_cells.hip", (int)names.size(), texts.data(),
                          names.data()) != HIPRTC_SUCCESS)
    return fail(MIND_EHIP, "hiprtcCreateProgram failed");
  for (const char* k : kKernelNames) hiprtcAddNameExpression(prog, k);
  const std::string target = "--offload-arch=" + arch;
  const char* opts[] = {target.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(MIND_EHIP, "hipRTC compilation of the cell kernels failed: " + log.substr(0, 4000));
  }
  size_t size = 0;
  hiprtcGetCodeSize(prog, &size);
  code->assign(size, 0);
  hiprtcGetCode(prog, code->data());
  for (int k = 0; k < K_COUNT; ++k) {
    const char* name = nullptr;
    if (hiprtcGetLoweredName(prog, kKernelNames[k], &name) != HIPRTC_SUCCESS || !name) {
      hiprtcDestroyProgram(&prog);
      return fail(MIND_EHIP, std::string("hipRTC has no kernel ") + kKernelNames[k]);
    }
    lowered[k] = name;
  }
  hiprtcDestroyProgram(&prog);
  return 0;
}

// Compiled mechanisms of a model, in id order.
int model_slots(const Layout& L, std::vector<int>* slots) {
  for (const MechGroup& G : L.groups)
    if (G.mech >= MIND_MECH_COUNT) slots->push_back(G.mech);
  std::sort(slots->begin(), slots->end());
  for (int id : *slots)
    if (!user_slot(id)) return fail(MIND_EARG, "unknown mechanism " + std::to_string(id));
  return 0;
}

}  // namespace

int rtc_kernels(State& S) {
  S.ks = nullptr;
  std::vector<int> slots;
  if (int rc = model_slots(S.L, &slots)) return rc;
  if (slots.empty()) return 0;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, S.device));
  const int max_groups = std::max<int>(MIND_MECH_COUNT, (int)S.L.groups.size());
  const std::string key = std::to_string(S.device) + "\n" + prop.gcnArchName + "\n" + std::to_string(max_groups) +
                          "\n" + user_header(slots);
  std::lock_guard<std::mutex> lock(g_mu);
  auto hit = g_modules.find(key);
  if (hit != g_modules.end()) {
    S.ks = hit->second;
    return 0;
  }
  std::vector<char> code;
  std::string lowered[K_COUNT];
  if (int rc = compile_module(slots, max_groups, prop.gcnArchName, &code, lowered)) return rc;
  auto* ks = new KernelSet();
  ks->max_groups = max_groups;
  hipError_t e = hipModuleLoadData(&ks->module, code.data());
  for (int k = 0; e == hipSuccess && k < K_COUNT; ++k) e = hipModuleGetFunction(&ks->fn[k], ks->module, lowered[k].c_str());
  if (e != hipSuccess) {
    delete ks;
    return fail(MIND_EHIP, std::string("loading the runtime-compiled cell kernels: ") + hipGetErrorString(e));
  }
  g_modules[key] = ks;
  S.ks = ks;
  return 0;
}

}  // namespace mind

extern "C" int mind_rtc_check(const int32_t* mechs, int32_t n, const char* arch) {
  using namespace mind;
  if ((n > 0 && !mechs) || !arch) return fail(MIND_EARG, "mind_rtc_check: null argument");
  std::vector<int> slots;
  for (int32_t k = 0; k < n; ++k) {
    if (!user_slot(mechs[k])) return fail(MIND_EARG, "mechanism " + std::to_string(mechs[k]) + " is not compiled");
    if (std::find(slots.begin(), slots.end(), mechs[k]) == slots.end()) slots.push_back(mechs[k]);
  }
  if (slots.empty()) return fail(MIND_EARG, "mind_rtc_check: no compiled mechanisms");
  std::sort(slots.begin(), slots.end());
  std::vector<char> code;
  std::string lowered[K_COUNT];
  return compile_module(slots, MIND_MECH_COUNT + (int)slots.size(), arch, &code, lowered);
}

namespace mind {

hipError_t launch_kernel(const State& S, int id, const void* fn, dim3 grid, dim3 block, size_t shmem,
                         hipStream_t stream, void** args) {
  if (S.ks)
    return hipModuleLaunchKernel(S.ks->fn[id], grid.x, grid.y, grid.z, block.x, block.y, block.z,
                                 (unsigned)shmem, stream, args, nullptr);
  return hipLaunchKernel(fn, grid, block, args, shmem, stream);
}

}  // namespace mind

// NMODL -> HIP C lowering, host side (no device code).
//
// Replaces the reference's compiler front end and code generator
// (src/compiler/frontend.jl, src/compiler/codegen.jl) for the micro mechanisms.
// The same lowering serves two callers:
//  * the library build: the builtin MOD files become mech_gen.h, compiled
//    into libmind (nmodlc, the command-line driver);
//  * runtime compilation: mind_compile_mod lowers a caller's MOD text into a
//    new mechanism slot whose phases are compiled for the device with hipRTC
//    when a model uses it, and by g++ for host checks.
// Each mechanism becomes MIND_HD functions over its field-major columns:
//   <P>_current / _current_load / _current_eval   nrn_cur: BREAKPOINT without
//       its SOLVE statements, evaluated at v + 0.001, then at v
//       (src/runtime/cuda/micro.jl:447-472); the current sum takes NONSPECIFIC
//       currents first, then ion WRITE currents in USEION order, minus
//       ELECTRODE currents (src/runtime/initialize.jl:496-502)
//       The `ish` argument, when not null, receives each written ion current
//       k (USEION order) at v + 0.001 in ish[k] and at v in ish[kIshSlots + k]
//       (the di/dv probe and the summed value, without reading the stored field
//       back); a mechanism writes at most kIshSlots ion currents.
//   <P>_advance   nrn_state: the SOLVE statements hoisted out of BREAKPOINT
//       (codegen.jl:372-378): cnexp relaxations x + (1 - exp(dt*(-1/tau))) *
//       (xinf - x) (codegen.jl:569-641), Euler, or a PROCEDURE
//   <P>_initial, <P>_receive   INITIAL and NET_RECEIVE; net_send / net_event
//       are captured in the returned NetStimOut
// where P is the mechanism name for the builtins and u<slot> for runtime
// slots. Fields are PARAMETER, then STATE, then ASSIGNED names in declaration
// order without the execution inputs v, t, dt and celsius; an array field
// takes one column per element. Expressions are emitted fully parenthesised
// in the MOD's left-to-right association, with exp/log/pow/fabs mapped to
// libmind's own FP64 functions, so that the IEEE operation sequence is the
// MOD's on every compiler that builds it with -ffp-contract=off.
#pragma once

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mind {
namespace nmodl {

// Written ion currents of one mechanism (the halves of the ish array).
constexpr int kIshSlots = 8;

// A lowering failure: "<file>:<line>: <message>".
struct ModError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Source digest: the first 16 hex digits of the SHA-256 of the text, as a
// number (src/compiler/ir.jl:316-324).
uint64_t digest(const std::string& text);

// The builtin set (filename, text) lowered to the mech_gen.h header of libmind.
std::string builtin_header(const std::vector<std::pair<std::string, std::string>>& sources);

// One USEION statement: per ion column (0 reversal, 1 internal and 2 external
// concentration, 3 current) the field column it reads into or writes from, -1
// if none. Reads include written concentrations (NMODL loads them).
struct IonUse {
  std::string name;
  bool has_valence = false;
  double valence = 0.0;
  int read[4] = {-1, -1, -1, -1};
  int write[4] = {-1, -1, -1, -1};
};

// A mechanism lowered for a runtime slot.
struct UserMechanism {
  std::string name, filename;
  int kind = 0;  // 0 density (SUFFIX), 1 point process
  uint64_t digest = 0;
  std::vector<std::string> columns;  // field columns; element k of array a is "a[k]" (1-based)
  std::vector<double> defaults;      // per column: PARAMETER defaults, ion values for ion reads
  bool has_net_receive = false;
  bool has_state = false;            // BREAKPOINT has SOLVE statements
  std::string code;                  // C++ phases named <prefix>_*, for namespace mind::user
  std::vector<IonUse> ions;          // USEION statements in declaration order
  std::vector<uint8_t> state_column; // per column: 1 for a STATE variable
  std::vector<uint8_t> global_column; // per column: 1 for a GLOBAL (shared, never assigned)
  std::vector<int> receive_fields;   // columns NET_RECEIVE reads or writes
};

// Lowers one MOD text for slot functions named `prefix` (a C identifier).
// Refuses what the device path cannot run for a runtime mechanism: artificial
// cells, RANDOM, net_send / net_event, celsius in NET_RECEIVE.
UserMechanism lower_user(const std::string& filename, const std::string& text, const std::string& prefix);

// The MIND-role mechanisms of the macro step (a MIND block with ROLE REGION,
// MACRO2MACRO, MACRO2MICRO or MICRO2MACRO) lowered to the macro_gen.h header of
// libmind, namespace mind::gen_macro: per mechanism its column enum, defaults,
// routines, <name>_evaluate (BREAKPOINT without SOLVE, once), _advance (the
// SOLVE statements), _initial and _receive.
std::string macro_header(const std::vector<std::pair<std::string, std::string>>& sources);

// Definitions the lowered phases of runtime slots use (GenCtx, gen_ctx,
// gen_index), for namespace mind::user.
std::string user_prelude();

}  // namespace nmodl
}  // namespace mind

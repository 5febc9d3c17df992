// Mechanism slots (host side): builtin metadata and the registry of mechanisms
// compiled at run time by mind_compile_mod.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "nmodl.h"

namespace mind {

// A runtime slot: its id, the MOD text it came from and its lowering, whose
// functions are named u<id>_*.
struct UserSlot {
  int id = 0;
  std::string source;
  std::string lowering;  // self-contained text for namespace mind::user
  nmodl::UserMechanism m;
};

// The slot of a compiled mechanism, or nullptr (builtins, unused ids).
const UserSlot* user_slot(int id);
// Whether `id` names a builtin with mechanism data or a compiled mechanism.
bool mech_known(int id);
// MIND_KIND_* of a known mechanism.
int mech_kind(int id);
bool mech_has_receive(int id);
// USEION statements of a known mechanism (hh: na and k; the other builtins none).
const std::vector<nmodl::IonUse>& mech_ions(int id);
// Whether column `col` of a known mechanism is a STATE variable, and the columns
// its NET_RECEIVE block touches.
bool mech_state_column(int id, int col);
const std::vector<int>& mech_receive_fields(int id);

}  // namespace mind

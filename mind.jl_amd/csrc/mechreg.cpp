// Mechanism slots: builtin metadata and the registry of runtime-compiled
// mechanisms behind mind_compile_mod (compile_mod / load_mod / builtin,
// src/compiler/frontend.jl:58-110). Slots live until the process ends.
#include "mechreg.h"

#include <string.h>

#include <algorithm>
#include <deque>
#include <mutex>

#include "../../include/mind.h"
#include "layout.h"
#include "mech.h"

namespace mind {

int fail(int code, const std::string& msg);

namespace {

struct Builtin {
  const char* name;
  int kind;
  bool receive;
  std::vector<const char*> fields;
};

// Column names of the builtins in include/mind.h order.
const std::vector<Builtin>& builtins() {
  static const std::vector<Builtin> b = {
      {"hh", MIND_KIND_DENSITY, false, {"gnabar", "gkbar", "gl", "el", "m", "h", "n", "ena", "ek", "gna", "gk",
                                        "ina", "ik", "il", "minf", "hinf", "ninf", "mtau", "htau", "ntau"}},
      {"pas", MIND_KIND_DENSITY, false, {"g", "e", "i"}},
      {"ExpSyn", MIND_KIND_POINT, true, {"tau", "e", "g", "i"}},
      {"Exp2Syn", MIND_KIND_POINT, true, {"tau1", "tau2", "e", "A", "B", "i", "g", "factor"}},
      {"IClamp", MIND_KIND_POINT, false, {"del", "dur", "amp", "i"}},
      {"NetStim", MIND_KIND_ARTIFICIAL, true, {"interval", "number", "start", "noise", "event", "on", "ispike"}},
      {"AlphaSynapse", MIND_KIND_POINT, false, {"onset", "tau", "gmax", "e", "i", "g"}},
      {"PatternStim", MIND_KIND_ARTIFICIAL, false, {}},
      {"SEClamp", MIND_KIND_POINT, false,
       {"rs", "dur1", "amp1", "dur2", "amp2", "dur3", "amp3", "i", "vc", "tc2", "tc3", "on"}},
      {"VClamp", MIND_KIND_POINT, false,
       {"dur[1]", "dur[2]", "dur[3]", "amp[1]", "amp[2]", "amp[3]", "gain", "rstim", "tau1", "tau2", "e0", "vo0",
        "vi0", "fac", "e", "vo", "vi", "i", "stim", "tc"}},
  };
  return b;
}

std::mutex g_mu;
std::deque<UserSlot> g_slots;  // stable addresses; slot id = MIND_MECH_COUNT + index

// A slot's lowering as one self-contained text for namespace mind::user.
std::string slot_text(const UserSlot& s) {
  return "// ---- runtime slot " + std::to_string(s.id) + " ----\nnamespace mind {\nnamespace user {\n"
         "#ifndef MIND_USER_PRELUDE\n#define MIND_USER_PRELUDE\n" +
         nmodl::user_prelude() + "#endif\n" + s.m.code + "\n}  // namespace user\n}  // namespace mind\n";
}

}  // namespace

const UserSlot* user_slot(int id) {
  std::lock_guard<std::mutex> lock(g_mu);
  const int k = id - MIND_MECH_COUNT;
  if (k < 0 || k >= (int)g_slots.size()) return nullptr;
  return &g_slots[k];
}

bool mech_known(int id) { return (id >= 0 && id < MIND_MECH_COUNT) || user_slot(id) != nullptr; }

int mech_kind(int id) {
  if (id >= 0 && id < MIND_MECH_COUNT) return builtins()[id].kind;
  const UserSlot* s = user_slot(id);
  return s ? s->m.kind : MIND_KIND_DENSITY;
}

bool mech_has_receive(int id) {
  if (id >= 0 && id < MIND_MECH_COUNT) return builtins()[id].receive;
  const UserSlot* s = user_slot(id);
  return s && s->m.has_net_receive;
}

const std::vector<nmodl::IonUse>& mech_ions(int id) {
  static const std::vector<nmodl::IonUse> none;
  // hh.mod: USEION na READ ena WRITE ina, USEION k READ ek WRITE ik
  static const std::vector<nmodl::IonUse> hh = [] {
    nmodl::IonUse na, k;
    na.name = "na";
    na.read[0] = HH_ENA;
    na.write[3] = HH_INA;
    k.name = "k";
    k.read[0] = HH_EK;
    k.write[3] = HH_IK;
    return std::vector<nmodl::IonUse>{na, k};
  }();
  if (id == MIND_MECH_HH) return hh;
  const UserSlot* s = user_slot(id);
  return s ? s->m.ions : none;
}

bool mech_state_column(int id, int col) {
  const UserSlot* s = user_slot(id);
  return s && col >= 0 && col < (int)s->m.state_column.size() && s->m.state_column[col];
}

const std::vector<int>& mech_receive_fields(int id) {
  static const std::vector<int> none;
  const UserSlot* s = user_slot(id);
  return s ? s->m.receive_fields : none;
}

}  // namespace mind

using namespace mind;

extern "C" int mind_compile_mod(const char* source, size_t length, const char* filename, int32_t* mech) {
  if (!source || !mech) return fail(MIND_EARG, "mind_compile_mod: null argument");
  const std::string text(source, length);
  const std::string file = filename ? filename : "<mod>";
  std::lock_guard<std::mutex> lock(g_mu);
  for (const UserSlot& s : g_slots)
    if (s.source == text && s.m.filename == file) {
      *mech = s.id;
      return 0;
    }
  if (MIND_MECH_COUNT + (int)g_slots.size() >= MIND_MECH_SLOTS)
    return fail(MIND_EQUEUE, "every runtime mechanism slot is taken");
  UserSlot s;
  s.id = MIND_MECH_COUNT + (int)g_slots.size();
  s.source = text;
  try {
    s.m = nmodl::lower_user(file, text, "u" + std::to_string(s.id));
  } catch (const nmodl::ModError& e) {
    return fail(MIND_EARG, e.what());
  }
  s.lowering = slot_text(s);
  g_slots.push_back(std::move(s));
  *mech = g_slots.back().id;
  return 0;
}

extern "C" int mind_mech_info(int32_t mech, mind_mech_desc* out) {
  if (!out) return fail(MIND_EARG, "mind_mech_info: null argument");
  if (mech >= 0 && mech < MIND_MECH_COUNT) {
    const Builtin& b = builtins()[mech];
    *out = mind_mech_desc{b.name, "", b.kind, (int32_t)b.fields.size(), b.receive ? 1 : 0,
                          mech_has_state(mech) ? 1 : 0, 0, 0};
    return 0;
  }
  const UserSlot* s = user_slot(mech);
  if (!s) return fail(MIND_EARG, "unknown mechanism " + std::to_string(mech));
  *out = mind_mech_desc{s->m.name.c_str(), s->m.filename.c_str(), s->m.kind, (int32_t)s->m.columns.size(),
                        s->m.has_net_receive ? 1 : 0, s->m.has_state ? 1 : 0, 1, s->m.digest};
  return 0;
}

extern "C" int mind_mech_fields(int32_t mech, const char** names, double* defaults, int32_t capacity,
                                int32_t* count) {
  if (!count) return fail(MIND_EARG, "mind_mech_fields: null count");
  if (mech >= 0 && mech < MIND_MECH_COUNT) {
    const Builtin& b = builtins()[mech];
    const int n = (int)b.fields.size();
    double def[32] = {0};
    if (n) field_defaults(mech, def);
    for (int k = 0; k < n && k < capacity; ++k) {
      if (names) names[k] = b.fields[k];
      if (defaults) defaults[k] = def[k];
    }
    *count = n;
    return 0;
  }
  const UserSlot* s = user_slot(mech);
  if (!s) return fail(MIND_EARG, "unknown mechanism " + std::to_string(mech));
  const int n = (int)s->m.columns.size();
  for (int k = 0; k < n && k < capacity; ++k) {
    if (names) names[k] = s->m.columns[k].c_str();
    if (defaults) defaults[k] = s->m.defaults[k];
  }
  *count = n;
  return 0;
}

extern "C" int mind_mech_lookup(const char* name, int32_t* mech) {
  if (!name || !mech) return fail(MIND_EARG, "mind_mech_lookup: null argument");
  {
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto it = g_slots.rbegin(); it != g_slots.rend(); ++it)
      if (it->m.name == name) {
        *mech = it->id;
        return 0;
      }
  }
  for (int k = 0; k < MIND_MECH_COUNT; ++k)
    if (strcmp(builtins()[k].name, name) == 0) {
      *mech = k;
      return 0;
    }
  return fail(MIND_EARG, std::string("unknown mechanism ") + name);
}

extern "C" int mind_mech_lowering(int32_t mech, const char** text) {
  if (!text) return fail(MIND_EARG, "mind_mech_lowering: null argument");
  const UserSlot* s = user_slot(mech);
  if (!s) return fail(MIND_EARG, "mechanism " + std::to_string(mech) + " is not a compiled mechanism");
  *text = s->lowering.c_str();
  return 0;
}

extern "C" int mind_mech_ions(int32_t mech, int32_t index, const char** name, int32_t* has_valence, double* valence,
                              int32_t read[4], int32_t write[4], int32_t* count) {
  if (!count) return fail(MIND_EARG, "mind_mech_ions: null count");
  if (!mech_known(mech)) return fail(MIND_EARG, "unknown mechanism " + std::to_string(mech));
  const std::vector<nmodl::IonUse>& ions = mech_ions(mech);
  *count = (int32_t)ions.size();
  if (index < 0 || index >= (int32_t)ions.size()) return 0;
  const nmodl::IonUse& u = ions[index];
  if (name) *name = u.name.c_str();
  if (has_valence) *has_valence = u.has_valence ? 1 : 0;
  if (valence) *valence = u.valence;
  for (int k = 0; k < 4; ++k) {
    if (read) read[k] = u.read[k];
    if (write) write[k] = u.write[k];
  }
  return 0;
}

extern "C" int mind_mech_globals(int32_t mech, uint8_t* is_global, int32_t capacity, int32_t* count) {
  if (!count) return fail(MIND_EARG, "mind_mech_globals: null count");
  if (!mech_known(mech)) return fail(MIND_EARG, "unknown mechanism " + std::to_string(mech));
  const int n = field_count(mech);
  const UserSlot* s = user_slot(mech);  // the builtins declare no GLOBAL
  for (int k = 0; k < n && k < capacity; ++k)
    if (is_global) is_global[k] = s && k < (int)s->m.global_column.size() && s->m.global_column[k] ? 1 : 0;
  *count = n;
  return 0;
}

extern "C" int mind_mech_columns(int32_t mech, uint8_t* is_state, uint8_t* in_receive, int32_t capacity,
                                 int32_t* count) {
  if (!count) return fail(MIND_EARG, "mind_mech_columns: null count");
  if (!mech_known(mech)) return fail(MIND_EARG, "unknown mechanism " + std::to_string(mech));
  const int n = field_count(mech);
  const std::vector<int>& rf = mech_receive_fields(mech);
  for (int k = 0; k < n && k < capacity; ++k) {
    if (is_state) is_state[k] = mech_state_column(mech, k) ? 1 : 0;
    if (in_receive) in_receive[k] = std::find(rf.begin(), rf.end(), k) != rf.end() ? 1 : 0;
  }
  *count = n;
  return 0;
}

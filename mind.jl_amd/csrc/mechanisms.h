// Mechanism phase dispatch used by the kernels: the NMODL lowering generated from
// MOD sources (mech_gen.h, produced by mind_amd.nmodl) when the library is built
// with MIND_GENERATED_MECHANISMS, else the hand lowering of mech.h. Both follow
// the reference's lowering rules and perform the same IEEE operations.
#pragma once

#include "mech.h"

#ifdef MIND_GENERATED_MECHANISMS
#include "mech_gen.h"
#endif

// Runtime-compiled mechanisms (slots MIND_MECH_COUNT and up): a module built
// with hipRTC for a model that uses them defines MIND_USER_MECHANISMS and
// provides user_mechs.h, whose mind::user::{current, advance, initial, receive,
// has_state} dispatch over its slots.
#ifdef MIND_USER_MECHANISMS
#include "user_mechs.h"
#define MIND_USER_CASE(call) \
  if (mech >= MIND_MECH_COUNT) { call; return; }
#define MIND_USER_VALUE(call) \
  if (mech >= MIND_MECH_COUNT) return call;
#else
#define MIND_USER_CASE(call)
#define MIND_USER_VALUE(call)
#endif

namespace mind {

#ifdef MIND_GENERATED_MECHANISMS
static_assert(gen::kComplete,
              "the generated mechanisms must cover hh, pas, ExpSyn, Exp2Syn, IClamp, NetStim, AlphaSynapse, SEClamp "
              "and VClamp");
constexpr const char* kMechanismSource = gen::kSource;
constexpr uint64_t kNetStimDigest = gen::kNetStimDigest;
MIND_HD void phase_current(int mech, const Cols& c, int64_t r, double v, double t, double dt, double& value,
                           double& shifted, bool store, double celsius, double* ish) {
  MIND_USER_CASE(user::current(mech, c, r, v, t, dt, value, shifted, store, celsius, ish))
  gen::current(mech, c, r, v, t, dt, value, shifted, store, celsius, ish);
}
// nrn_cur of mechanism slot MECH in two parts: the column reads of a row (load),
// then the evaluation and stores (eval).
template <int MECH>
using CurOps = gen::CurOf<MECH>;
MIND_HD void phase_advance(int mech, const Cols& c, int64_t r, double v, double t, double dt, double celsius) {
  MIND_USER_CASE(user::advance(mech, c, r, v, t, dt, celsius))
  gen::advance(mech, c, r, v, t, dt, celsius);
}
MIND_HD void phase_initial(int mech, const Cols& c, int64_t r, double v, double celsius) {
  MIND_USER_CASE(user::initial(mech, c, r, v, celsius))
  gen::initial(mech, c, r, v, celsius);
}
// NET_RECEIVE of a connection event delivered at time t (the builtin receivers
// other than NetStim do not read t).
MIND_HD void phase_receive(int mech, const Cols& c, int64_t r, double t, double weight) {
  MIND_USER_CASE(user::receive(mech, c, r, t, weight))
  (void)t;
  gen::receive(mech, c, r, weight);
}
MIND_HD NetStimOut phase_netstim_initial(const Cols& c, int64_t r, Stream& s, uint64_t seed) {
  return gen::NetStim_initial(c, r, 0.0, 0.0, &s, seed);
}
MIND_HD NetStimOut phase_netstim_receive(const Cols& c, int64_t r, Stream& s, uint64_t seed, double t, double w,
                                         double flag) {
  return gen::NetStim_receive(c, r, t, w, flag, &s, seed);
}
#else
constexpr const char* kMechanismSource = "hand lowering (mech.h)";
// Source digest of the NetStim mechanism: the first 16 hex digits of the
// SHA-256 of NetStim.mod (src/compiler/ir.jl:316-324).
constexpr uint64_t kNetStimDigest = 0xa186833b569ff3b2ull;
MIND_HD void phase_current(int mech, const Cols& c, int64_t r, double v, double t, double dt, double& value,
                           double& shifted, bool store, double celsius, double* ish) {
  MIND_USER_CASE(user::current(mech, c, r, v, t, dt, value, shifted, store, celsius, ish))
  (void)celsius;  // no builtin reads celsius in BREAKPOINT
  mech_current(mech, c, r, v, t, dt, value, shifted, store, ish);
}
// nrn_cur of mechanism slot MECH in two parts; the hand lowering reads its
// columns in the evaluation.
template <int MECH>
struct CurOps {
  struct In {};
  static MIND_HD In load(const Cols&, int64_t) { return In{}; }
  static MIND_HD void eval(const Cols& c, int64_t r, const In&, double v, double t, double dt, double& value,
                           double& shifted, bool store, double celsius) {
    (void)celsius;
    mech_current(MECH, c, r, v, t, dt, value, shifted, store, nullptr);
  }
  static MIND_HD void eval_ish(const Cols& c, int64_t r, const In&, double v, double t, double dt, double& value,
                               double& shifted, bool store, double celsius, double* ish) {
    (void)celsius;
    mech_current(MECH, c, r, v, t, dt, value, shifted, store, ish);
  }
  // the evaluation reads the column itself, which the caller has stored
  static MIND_HD void set(In&, int, double) {}
};
MIND_HD void phase_advance(int mech, const Cols& c, int64_t r, double v, double t, double dt, double celsius) {
  MIND_USER_CASE(user::advance(mech, c, r, v, t, dt, celsius))
  mech_advance(mech, c, r, v, t, dt, celsius);
}
MIND_HD void phase_initial(int mech, const Cols& c, int64_t r, double v, double celsius) {
  MIND_USER_CASE(user::initial(mech, c, r, v, celsius))
  mech_initial(mech, c, r, v, celsius);
}
// NET_RECEIVE of a connection event delivered at time t (the builtin receivers
// other than NetStim do not read t).
MIND_HD void phase_receive(int mech, const Cols& c, int64_t r, double t, double weight) {
  MIND_USER_CASE(user::receive(mech, c, r, t, weight))
  (void)t;
  synapse_receive(mech, c, r, weight);
}
MIND_HD NetStimOut phase_netstim_initial(const Cols& c, int64_t r, Stream& s, uint64_t seed) {
  return netstim_initial(c, r, s, seed);
}
MIND_HD NetStimOut phase_netstim_receive(const Cols& c, int64_t r, Stream& s, uint64_t seed, double t, double w,
                                         double flag) {
  return netstim_receive(c, r, s, seed, t, w, flag);
}
#endif

// Whether a mechanism has a state update (nrn_state).
MIND_HD bool phase_has_state(int mech) {
  MIND_USER_VALUE(user::has_state(mech))
  return mech_has_state(mech);
}

}  // namespace mind

// Mechanism phase dispatch used by the kernels: the NMODL lowering generated from
// MOD sources (mech_gen.h, produced by mind_amd.nmodl) when the library is built
// with MIND_GENERATED_MECHANISMS, else the hand lowering of mech.h. Both follow
// the reference's lowering rules and perform the same IEEE operations.
#pragma once

#include "mech.h"

#ifdef MIND_GENERATED_MECHANISMS
#include "mech_gen.h"
#endif

namespace mind {

#ifdef MIND_GENERATED_MECHANISMS
static_assert(gen::kComplete,
              "the generated mechanisms must cover hh, pas, ExpSyn, Exp2Syn, IClamp, NetStim, AlphaSynapse, SEClamp "
              "and VClamp");
constexpr const char* kMechanismSource = gen::kSource;
constexpr uint64_t kNetStimDigest = gen::kNetStimDigest;
MIND_HD void phase_current(int mech, const Cols& c, int64_t r, double v, double t, double dt, double& value,
                           double& shifted, bool store) {
  gen::current(mech, c, r, v, t, dt, value, shifted, store);
}
// nrn_cur of mechanism slot MECH in two parts: the column reads of a row (load),
// then the evaluation and stores (eval).
template <int MECH>
using CurOps = gen::CurOf<MECH>;
MIND_HD void phase_advance(int mech, const Cols& c, int64_t r, double v, double t, double dt, double celsius) {
  gen::advance(mech, c, r, v, t, dt, celsius);
}
MIND_HD void phase_initial(int mech, const Cols& c, int64_t r, double v, double celsius) {
  gen::initial(mech, c, r, v, celsius);
}
MIND_HD void phase_receive(int mech, const Cols& c, int64_t r, double weight) { gen::receive(mech, c, r, weight); }
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
                           double& shifted, bool store) {
  mech_current(mech, c, r, v, t, dt, value, shifted, store);
}
// nrn_cur of mechanism slot MECH in two parts; the hand lowering reads its
// columns in the evaluation.
template <int MECH>
struct CurOps {
  struct In {};
  static MIND_HD In load(const Cols&, int64_t) { return In{}; }
  static MIND_HD void eval(const Cols& c, int64_t r, const In&, double v, double t, double dt, double& value,
                           double& shifted, bool store) {
    mech_current(MECH, c, r, v, t, dt, value, shifted, store);
  }
};
MIND_HD void phase_advance(int mech, const Cols& c, int64_t r, double v, double t, double dt, double celsius) {
  mech_advance(mech, c, r, v, t, dt, celsius);
}
MIND_HD void phase_initial(int mech, const Cols& c, int64_t r, double v, double celsius) {
  mech_initial(mech, c, r, v, celsius);
}
MIND_HD void phase_receive(int mech, const Cols& c, int64_t r, double weight) { synapse_receive(mech, c, r, weight); }
MIND_HD NetStimOut phase_netstim_initial(const Cols& c, int64_t r, Stream& s, uint64_t seed) {
  return netstim_initial(c, r, s, seed);
}
MIND_HD NetStimOut phase_netstim_receive(const Cols& c, int64_t r, Stream& s, uint64_t seed, double t, double w,
                                         double flag) {
  return netstim_receive(c, r, s, seed, t, w, flag);
}
#endif

}  // namespace mind

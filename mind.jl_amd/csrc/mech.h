// Device (and host) lowering of the seven builtin mechanisms of the simulation
// step: hh, pas, ExpSyn, Exp2Syn, IClamp, NetStim, AlphaSynapse, SEClamp and VClamp.
//
// Each function is the nrn_cur / nrn_state / NET_RECEIVE / INITIAL phase of one
// MOD file as the reference lowers it (src/compiler/codegen.jl:362-485):
// BREAKPOINT without its SOLVE for the current phase, the hoisted SOLVE with
// cnexp relaxation x += (1 - exp(dt*(-1/tau))) * (xinf - x) for the state
// phase (codegen.jl:625-627), NET_RECEIVE with the connection weight and flag.
// Expressions keep the MOD's left-to-right evaluation order so that the IEEE
// results match the CPU restatement bit for bit (-ffp-contract=off).
#pragma once

#ifdef __HIPCC_RTC__
#include "mind.h"
#else
#include "../../include/mind.h"
#endif
#include "mind_math.h"

namespace mind {

// hh.mod FUNCTION vtrap (hh.mod:116-122)
MIND_HD double hh_vtrap(double x, double y) {
  if (mind_fabs(x / y) < 1e-6) return y * (1.0 - x / y / 2.0);
  return x / (mind_exp(x / y) - 1.0);
}

struct HHRates { double minf, hinf, ninf, mtau, htau, ntau; };

// hh.mod PROCEDURE rates (hh.mod:90-114)
MIND_HD HHRates hh_rates(double v, double celsius) {
  HHRates r;
  const double q10 = mind_pow(3.0, (celsius - 6.3) / 10.0);
  double alpha = 0.1 * hh_vtrap(-(v + 40.0), 10.0);
  double beta = 4.0 * mind_exp(-(v + 65.0) / 18.0);
  double sum = alpha + beta;
  r.mtau = 1.0 / (q10 * sum);
  r.minf = alpha / sum;
  alpha = 0.07 * mind_exp(-(v + 65.0) / 20.0);
  beta = 1.0 / (mind_exp(-(v + 35.0) / 10.0) + 1.0);
  sum = alpha + beta;
  r.htau = 1.0 / (q10 * sum);
  r.hinf = alpha / sum;
  alpha = 0.01 * hh_vtrap(-(v + 55.0), 10.0);
  beta = 0.125 * mind_exp(-(v + 65.0) / 80.0);
  sum = alpha + beta;
  r.ntau = 1.0 / (q10 * sum);
  r.ninf = alpha / sum;
  return r;
}

MIND_HD double cnexp_relax(double x, double xinf, double tau, double dt) {
  return x + (1.0 - mind_exp(dt * (-1.0 / tau))) * (xinf - x);
}

// Column accessor over field-major mechanism data.
struct Cols {
  double* data;
  int64_t rows;
  MIND_HD double& operator()(int field, int64_t row) const { return data[field * rows + row]; }
};

// One ion transfer of a mechanism group: mechanism field `field` and ion column
// `column` (MIND_ION_*) of the row's USEION statement `slot`. aux: for a current
// write, its index among the written ion currents (the di/dv probe); for a
// NET_RECEIVE exchange, 1 if the field is a STATE (published back).
struct IonBind {
  int32_t field, column, slot, aux;
};
// Written ion currents one mechanism may have (USEION statements writing i).
constexpr int kMaxIonCurrents = 8;

// VClamp.mod clamp(): the three amplifier equations solved for e, vo and vi from
// the previous step's e0, vo0, vi0 (VClamp.mod PROCEDURE clamp).
MIND_HD void vclamp_amplifier(const Cols& c, int64_t r, double v, double dt, double stim, double e0, double vo0,
                              double& e, double& vo, double& vi) {
  const double t1 = c(VCLAMP_TAU1, r) / dt;
  const double t2 = c(VCLAMP_TAU2, r) / dt;
  const double fac = c(VCLAMP_FAC, r), gain = c(VCLAMP_GAIN, r);
  e = (((((1.0 + t1) * (((1.0 - fac) * v) + (((fac * t2) * vo0) / (1.0 + t2)))) - stim) -
        (t1 * (c(VCLAMP_VI0, r) - e0))) /
       ((2.0 + t1) + ((((1.0 + t1) * fac) * gain) / (1.0 + t2))));
  vo = (((t2 * vo0) - (gain * e)) / (1.0 + t2));
  vi = ((v + (fac * vo)) - (fac * v));
}

// Current phase. Returns the membrane current sum I(v) in the units of the
// mechanism (mA/cm2 for density, nA for point processes) and its probe at
// v + 0.001; stores the assigned fields of the evaluation at v.
// Current fields sum NONSPECIFIC currents first, then ion writes (hh: il, ina,
// ik), and subtract ELECTRODE currents (src/runtime/initialize.jl:496-502).
// ish (if not null) receives the written ion currents in USEION order (hh: ina,
// ik) at v + 0.001 in ish[k], for the ion rows' di/dv, and at v in
// ish[kMaxIonCurrents + k].
MIND_HD void mech_current(int mech, const Cols& c, int64_t r, double v, double t, double dt,
                          double& value, double& shifted, bool store, double* ish) {
  (void)dt;
  for (int probe = 0; probe < 2; ++probe) {
    const double vv = probe == 0 ? v + 0.001 : v;
    double sum = 0.0;
    switch (mech) {
      case MIND_MECH_HH: {
        const double m = c(HH_M, r), h = c(HH_H, r), n = c(HH_N, r);
        const double gna = c(HH_GNABAR, r) * m * m * m * h;
        const double ina = gna * (vv - c(HH_ENA, r));
        const double gk = c(HH_GKBAR, r) * n * n * n * n;
        const double ik = gk * (vv - c(HH_EK, r));
        const double il = c(HH_GL, r) * (vv - c(HH_EL, r));
        sum += il;
        sum += ina;
        sum += ik;
        if (ish) {
          ish[probe == 0 ? 0 : kMaxIonCurrents] = ina;
          ish[probe == 0 ? 1 : kMaxIonCurrents + 1] = ik;
        }
        if (probe == 1 && store) {
          c(HH_GNA, r) = gna; c(HH_INA, r) = ina; c(HH_GK, r) = gk; c(HH_IK, r) = ik;
          c(HH_IL, r) = il;
        }
        break;
      }
      case MIND_MECH_PAS: {
        const double i = c(PAS_G, r) * (vv - c(PAS_E, r));
        sum += i;
        if (probe == 1 && store) c(PAS_I, r) = i;
        break;
      }
      case MIND_MECH_EXPSYN: {
        const double i = c(EXPSYN_G, r) * (vv - c(EXPSYN_E, r));
        sum += i;
        if (probe == 1 && store) c(EXPSYN_I, r) = i;
        break;
      }
      case MIND_MECH_EXP2SYN: {
        const double g = c(EXP2SYN_B, r) - c(EXP2SYN_A, r);
        const double i = g * (vv - c(EXP2SYN_E, r));
        sum += i;
        if (probe == 1 && store) { c(EXP2SYN_G, r) = g; c(EXP2SYN_I, r) = i; }
        break;
      }
      case MIND_MECH_ICLAMP: {
        const double del = c(ICLAMP_DEL, r), dur = c(ICLAMP_DUR, r);
        const double i = (t < del + dur && t >= del) ? c(ICLAMP_AMP, r) : 0.0;
        sum -= i;
        if (probe == 1 && store) c(ICLAMP_I, r) = i;
        break;
      }
      case MIND_MECH_ALPHASYNAPSE: {  // AlphaSynapse.mod BREAKPOINT; at_time only marks a discontinuity
        const double x = (t - c(ALPHASYNAPSE_ONSET, r)) / c(ALPHASYNAPSE_TAU, r);
        const double a = (x < 0.0 || x > 10.0) ? 0.0 : x * mind_exp(1.0 - x);
        const double g = c(ALPHASYNAPSE_GMAX, r) * a;
        const double i = g * (vv - c(ALPHASYNAPSE_E, r));
        sum += i;
        if (probe == 1 && store) { c(ALPHASYNAPSE_G, r) = g; c(ALPHASYNAPSE_I, r) = i; }
        break;
      }
      case MIND_MECH_SECLAMP: {  // SEClamp.mod BREAKPOINT without its SOLVE: vstim(), which calls icur()
        double on = 1.0, vc = 0.0;
        if (t < c(SECLAMP_DUR1, r)) vc = c(SECLAMP_AMP1, r);
        else if (t < c(SECLAMP_TC2, r)) vc = c(SECLAMP_AMP2, r);
        else if (t < c(SECLAMP_TC3, r)) vc = c(SECLAMP_AMP3, r);
        else on = 0.0;
        const double i = on != 0.0 ? (vc - vv) / c(SECLAMP_RS, r) : 0.0;
        sum -= i;
        if (probe == 1 && store) { c(SECLAMP_ON, r) = on; c(SECLAMP_VC, r) = vc; c(SECLAMP_I, r) = i; }
        break;
      }
      case MIND_MECH_VCLAMP: {  // VClamp.mod BREAKPOINT without its SOLVE: vstim(), then i = icur()
        double tc = 0.0, stim = c(VCLAMP_STIM, r), j = 0.0, selected = 0.0;
        while (j < 3.0 && !selected) {
          stim = c(VCLAMP_AMP + (int)j, r);
          tc = tc + c(VCLAMP_DUR + (int)j, r);
          if (t < tc) {
            tc = tc + 100.0;
            selected = 1.0;
          }
          j = j + 1.0;
        }
        double e0 = c(VCLAMP_E0, r), vo0 = c(VCLAMP_VO0, r);
        double e = c(VCLAMP_E, r), vo = c(VCLAMP_VO, r), vi = c(VCLAMP_VI, r), i;
        if (t > tc) {
          e0 = 0.0;
          vo0 = 0.0;
          i = 0.0;
        } else {
          vclamp_amplifier(c, r, vv, dt, stim, e0, vo0, e, vo, vi);
          i = (vo - vv) / c(VCLAMP_RSTIM, r);
        }
        sum -= i;
        if (probe == 1 && store) {
          c(VCLAMP_TC, r) = tc; c(VCLAMP_STIM, r) = stim; c(VCLAMP_I, r) = i; c(VCLAMP_E0, r) = e0;
          c(VCLAMP_VO0, r) = vo0; c(VCLAMP_E, r) = e; c(VCLAMP_VO, r) = vo; c(VCLAMP_VI, r) = vi;
        }
        break;
      }
      default: break;
    }
    if (probe == 0) shifted = sum; else value = sum;
  }
}

// Mechanisms whose BREAKPOINT has a SOLVE, i.e. a state phase (nrn_state).
MIND_HD bool mech_has_state(int mech) {
  return mech == MIND_MECH_HH || mech == MIND_MECH_EXP2SYN || mech == MIND_MECH_EXPSYN ||
         mech == MIND_MECH_SECLAMP || mech == MIND_MECH_VCLAMP;
}


// State phase at t + dt: hoisted SOLVE of BREAKPOINT (pas and IClamp have none;
// SEClamp's is `SOLVE icur METHOD after_cvode`, the clamp current at the new v).
MIND_HD void mech_advance(int mech, const Cols& c, int64_t r, double v, double t, double dt,
                          double celsius) {
  switch (mech) {
    case MIND_MECH_HH: {
      const HHRates q = hh_rates(v, celsius);
      c(HH_MINF, r) = q.minf; c(HH_HINF, r) = q.hinf; c(HH_NINF, r) = q.ninf;
      c(HH_MTAU, r) = q.mtau; c(HH_HTAU, r) = q.htau; c(HH_NTAU, r) = q.ntau;
      c(HH_M, r) = cnexp_relax(c(HH_M, r), q.minf, q.mtau, dt);
      c(HH_H, r) = cnexp_relax(c(HH_H, r), q.hinf, q.htau, dt);
      c(HH_N, r) = cnexp_relax(c(HH_N, r), q.ninf, q.ntau, dt);
      break;
    }
    case MIND_MECH_EXP2SYN:
      c(EXP2SYN_A, r) = cnexp_relax(c(EXP2SYN_A, r), 0.0, c(EXP2SYN_TAU1, r), dt);
      c(EXP2SYN_B, r) = cnexp_relax(c(EXP2SYN_B, r), 0.0, c(EXP2SYN_TAU2, r), dt);
      break;
    case MIND_MECH_EXPSYN:
      c(EXPSYN_G, r) = cnexp_relax(c(EXPSYN_G, r), 0.0, c(EXPSYN_TAU, r), dt);
      break;
    case MIND_MECH_SECLAMP:
      c(SECLAMP_I, r) = c(SECLAMP_ON, r) != 0.0 ? (c(SECLAMP_VC, r) - v) / c(SECLAMP_RS, r) : 0.0;
      break;
    case MIND_MECH_VCLAMP: {  // SOLVE update METHOD after_cvode: i = icur() at the new v, then e0, vo0, vi0
      double i;
      if (t > c(VCLAMP_TC, r)) {
        c(VCLAMP_E0, r) = 0.0;
        c(VCLAMP_VO0, r) = 0.0;
        i = 0.0;
      } else {
        double e, vo, vi;
        vclamp_amplifier(c, r, v, dt, c(VCLAMP_STIM, r), c(VCLAMP_E0, r), c(VCLAMP_VO0, r), e, vo, vi);
        c(VCLAMP_E, r) = e; c(VCLAMP_VO, r) = vo; c(VCLAMP_VI, r) = vi;
        i = (c(VCLAMP_VO, r) - v) / c(VCLAMP_RSTIM, r);
      }
      c(VCLAMP_I, r) = i;
      c(VCLAMP_E0, r) = c(VCLAMP_E, r);
      c(VCLAMP_VO0, r) = c(VCLAMP_VO, r);
      c(VCLAMP_VI0, r) = c(VCLAMP_VI, r);
      break;
    }
    default: break;
  }
}

// INITIAL blocks of the membrane mechanisms (NetStim is handled by the queue).
MIND_HD void mech_initial(int mech, const Cols& c, int64_t r, double v, double celsius) {
  switch (mech) {
    case MIND_MECH_HH: {
      const HHRates q = hh_rates(v, celsius);
      c(HH_MINF, r) = q.minf; c(HH_HINF, r) = q.hinf; c(HH_NINF, r) = q.ninf;
      c(HH_MTAU, r) = q.mtau; c(HH_HTAU, r) = q.htau; c(HH_NTAU, r) = q.ntau;
      c(HH_M, r) = q.minf; c(HH_H, r) = q.hinf; c(HH_N, r) = q.ninf;
      break;
    }
    case MIND_MECH_EXP2SYN: {  // Exp2Syn.mod:56-69
      double tau1 = c(EXP2SYN_TAU1, r);
      const double tau2 = c(EXP2SYN_TAU2, r);
      if (tau1 / tau2 > 0.9999) tau1 = 0.9999 * tau2;
      if (tau1 / tau2 < 1e-9) tau1 = tau2 * 1e-9;
      c(EXP2SYN_TAU1, r) = tau1;
      c(EXP2SYN_A, r) = 0.0;
      c(EXP2SYN_B, r) = 0.0;
      const double tp = (tau1 * tau2) / (tau2 - tau1) * mind_log(tau2 / tau1);
      double factor = -mind_exp(-tp / tau1) + mind_exp(-tp / tau2);
      c(EXP2SYN_FACTOR, r) = 1.0 / factor;
      break;
    }
    case MIND_MECH_EXPSYN: c(EXPSYN_G, r) = 0.0; break;
    case MIND_MECH_ICLAMP: c(ICLAMP_I, r) = 0.0; break;
    case MIND_MECH_VCLAMP:  // VClamp.mod INITIAL
      c(VCLAMP_E0, r) = 0.0;
      c(VCLAMP_VO, r) = v;
      c(VCLAMP_VO0, r) = v;
      c(VCLAMP_VI, r) = v;
      c(VCLAMP_VI0, r) = v;
      c(VCLAMP_E, r) = 0.0;
      break;
    case MIND_MECH_SECLAMP:  // SEClamp.mod INITIAL
      c(SECLAMP_TC2, r) = c(SECLAMP_DUR1, r) + c(SECLAMP_DUR2, r);
      c(SECLAMP_TC3, r) = c(SECLAMP_TC2, r) + c(SECLAMP_DUR3, r);
      c(SECLAMP_ON, r) = 0.0;
      break;
    default: break;
  }
}

// NET_RECEIVE of the synapses (Exp2Syn.mod:82-85, ExpSyn.mod NET_RECEIVE).
MIND_HD void synapse_receive(int mech, const Cols& c, int64_t r, double weight) {
  if (mech == MIND_MECH_EXP2SYN) {
    const double f = c(EXP2SYN_FACTOR, r);
    c(EXP2SYN_A, r) = c(EXP2SYN_A, r) + weight * f;
    c(EXP2SYN_B, r) = c(EXP2SYN_B, r) + weight * f;
  } else if (mech == MIND_MECH_EXPSYN) {
    c(EXPSYN_G, r) = c(EXPSYN_G, r) + weight;
  }
}

// ---- Random123 Philox4x32-10 streams (src/runtime/cuda/micro.jl:3-91) -------
MIND_HD void philox_block(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d, uint64_t seed) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int i = 0; i < 10; ++i) {
    const uint64_t p0 = (uint64_t)0xd2511f53u * a;
    const uint64_t p1 = (uint64_t)0xcd9e8d57u * c;
    const uint32_t x0 = (uint32_t)(p1 >> 32) ^ b ^ k0;
    const uint32_t x2 = (uint32_t)(p0 >> 32) ^ d ^ k1;
    b = (uint32_t)p1;
    d = (uint32_t)p0;
    a = x0;
    c = x2;
    k0 += 0x9e3779b9u;
    k1 += 0xbb67ae85u;
  }
}

struct Stream {  // one row's stream: counter, 3 keys, 4 buffered words
  uint64_t counter;
  uint64_t key[3];
  uint32_t word[4];
};

MIND_HD void stream_refresh(Stream& s, uint64_t seed) {
  uint32_t a = (uint32_t)(s.counter >> 2), b = (uint32_t)s.key[2], c = (uint32_t)s.key[0],
           d = (uint32_t)s.key[1];
  philox_block(a, b, c, d, seed);
  s.word[0] = a; s.word[1] = b; s.word[2] = c; s.word[3] = d;
}

MIND_HD double stream_uniform(Stream& s, uint64_t seed) {
  const uint64_t counter = s.counter;
  const uint32_t w = s.word[counter & 3u];
  const uint64_t next = (counter + 1) & 0x3ffffffffull;
  s.counter = next;
  if ((counter & 3u) == 3u) {
    if (next == 0) {
      const int order[3] = {2, 0, 1};
      for (int i = 0; i < 3; ++i) {
        const uint32_t x = (uint32_t)(s.key[order[i]] + 1);
        s.key[order[i]] = x;
        if (x != 0) break;
      }
    }
    stream_refresh(s, seed);
  }
  return ((double)w + 1.0) * (1.0 / 4294967297.0);
}

MIND_HD void stream_setseq(Stream& s, uint64_t seed, double seq) {
  const double bounded = (seq >= 0.0 && seq <= 1099511627775.0) ? seq : 0.0;
  s.counter = ((uint64_t)bounded) & 0x3ffffffffull;
  stream_refresh(s, seed);
}

// ---- NetStim (NetStim.mod) ---------------------------------------------------
// Outcome of one NET_RECEIVE: an optional net_event and an optional net_send.
struct NetStimOut {
  int spiked;
  double spike_time;
  int send;
  double send_delay, send_flag;  // net_send(delay, flag) at the event time
};

MIND_HD double netstim_invl(const Cols& c, int64_t r, Stream& s, uint64_t seed, double mean) {
  if (mean <= 0.0) mean = 0.01;
  const double noise = c(NETSTIM_NOISE, r);
  if (noise == 0.0) return mean;
  const double erand = -mind_log(stream_uniform(s, seed)) * 1.0;
  return (1.0 - noise) * mean + noise * mean * erand;
}

MIND_HD void netstim_init_sequence(const Cols& c, int64_t r) {
  if (c(NETSTIM_NUMBER, r) > 0.0) {
    c(NETSTIM_ON, r) = 1.0;
    c(NETSTIM_EVENT, r) = 0.0;
    c(NETSTIM_ISPIKE, r) = 0.0;
  }
}

MIND_HD void netstim_next_invl(const Cols& c, int64_t r, Stream& s, uint64_t seed) {
  if (c(NETSTIM_NUMBER, r) > 0.0)
    c(NETSTIM_EVENT, r) = netstim_invl(c, r, s, seed, c(NETSTIM_INTERVAL, r));
  if (c(NETSTIM_ISPIKE, r) >= c(NETSTIM_NUMBER, r)) c(NETSTIM_ON, r) = 0.0;
}

// INITIAL (NetStim.mod:23-45) after random_setseq(ranvar, 0).
MIND_HD NetStimOut netstim_initial(const Cols& c, int64_t r, Stream& s, uint64_t seed) {
  NetStimOut o = {0, 0.0, 0, 0.0, 0.0};
  stream_setseq(s, seed, 0.0);
  c(NETSTIM_ON, r) = 0.0;
  c(NETSTIM_ISPIKE, r) = 0.0;
  if (c(NETSTIM_NOISE, r) < 0.0) c(NETSTIM_NOISE, r) = 0.0;
  if (c(NETSTIM_NOISE, r) > 1.0) c(NETSTIM_NOISE, r) = 1.0;
  if (c(NETSTIM_START, r) >= 0.0 && c(NETSTIM_NUMBER, r) > 0.0) {
    c(NETSTIM_ON, r) = 1.0;
    const double interval = c(NETSTIM_INTERVAL, r);
    double event = c(NETSTIM_START, r) + netstim_invl(c, r, s, seed, interval) -
                   interval * (1.0 - c(NETSTIM_NOISE, r));
    if (event < 0.0) event = 0.0;
    c(NETSTIM_EVENT, r) = event;
    o.send = 1;
    o.send_delay = event;
    o.send_flag = 3.0;
  }
  return o;
}

// NET_RECEIVE (NetStim.mod:77-105) at event time t with weight w and flag.
MIND_HD NetStimOut netstim_receive(const Cols& c, int64_t r, Stream& s, uint64_t seed, double t,
                                   double w, double flag) {
  NetStimOut o = {0, 0.0, 0, 0.0, 0.0};
  if (flag == 0.0) {
    if (w > 0.0 && c(NETSTIM_ON, r) == 0.0) {
      netstim_init_sequence(c, r);
      netstim_next_invl(c, r, s, seed);
      c(NETSTIM_EVENT, r) =
          c(NETSTIM_EVENT, r) - c(NETSTIM_INTERVAL, r) * (1.0 - c(NETSTIM_NOISE, r));
      o.send = 1; o.send_delay = c(NETSTIM_EVENT, r); o.send_flag = 1.0;
    } else if (w < 0.0) {
      c(NETSTIM_ON, r) = 0.0;
    }
  }
  if (flag == 3.0) {
    if (c(NETSTIM_ON, r) == 1.0) {
      netstim_init_sequence(c, r);
      o.send = 1; o.send_delay = 0.0; o.send_flag = 1.0;
    }
  }
  if (flag == 1.0 && c(NETSTIM_ON, r) == 1.0) {
    c(NETSTIM_ISPIKE, r) = c(NETSTIM_ISPIKE, r) + 1.0;
    o.spiked = 1;
    o.spike_time = t;
    netstim_next_invl(c, r, s, seed);
    if (c(NETSTIM_ON, r) == 1.0) {
      o.send = 1; o.send_delay = c(NETSTIM_EVENT, r); o.send_flag = 1.0;
    }
  }
  return o;
}

}  // namespace mind

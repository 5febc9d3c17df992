: Calcium-activated potassium current, instantaneous in cai; written for the
: ion-dynamics tests of this repository.
NEURON {
    SUFFIX kca
    USEION k READ ek WRITE ik
    USEION ca READ cai
    RANGE gbar, kd
}
PARAMETER { gbar = 0.01 (S/cm2) kd = 1e-3 (mM) }
ASSIGNED { v (mV) ek (mV) ik (mA/cm2) cai (mM) }
BREAKPOINT { ik = gbar*cai/(cai + kd)*(v - ek) }

: High-voltage-activated calcium current through the GHK flux equation,
: written for the ion-dynamics tests of this repository.
NEURON {
    SUFFIX cal
    USEION ca READ cai, cao WRITE ica
    RANGE pbar
}
UNITS { (mA) = (milliamp) (mV) = (millivolt) (mM) = (milli/liter) }
PARAMETER {
    pbar = 2e-4 (cm/s)
}
STATE { m }
ASSIGNED { v (mV) ica (mA/cm2) cai (mM) cao (mM) minf mtau (ms) }
BREAKPOINT {
    SOLVE gate METHOD cnexp
    ica = pbar*m*m*ghk(v, cai, cao, 2)
}
INITIAL { rates(v)  m = minf }
DERIVATIVE gate {
    rates(v)
    m' = (minf - m)/mtau
}
PROCEDURE rates(vm (mV)) {
    minf = 1/(1 + exp(-(vm + 25)/5))
    mtau = 1.5
}

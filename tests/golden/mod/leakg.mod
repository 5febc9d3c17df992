: A passive leak whose conductance and reversal are GLOBALs: one value shared by
: every instance, read (never assigned) by BREAKPOINT. With gl = g and el = e it
: restates pas.mod's current.
NEURON {
    SUFFIX leakg
    NONSPECIFIC_CURRENT i
    GLOBAL gl, el
    RANGE i
}
UNITS {
    (mA) = (milliamp)
    (mV) = (millivolt)
}
PARAMETER {
    gl = 0.0001 (S/cm2)
    el = -65 (mV)
}
ASSIGNED {
    v (mV)
    i (mA/cm2)
}
BREAKPOINT {
    i = gl * (v - el)
}

: Persistent sodium current with one activation gate integrated by forward
: Euler, a two-element array parameter for its half-activation voltage and
: slope, written for the runtime MOD tests.
NEURON {
    SUFFIX nap
    USEION na READ ena WRITE ina
    RANGE gbar, shape
}

PARAMETER {
    gbar = 0.0002 (S/cm2)
    shape[2]
    tau = 1.5 (ms)
}

STATE { p }

ASSIGNED {
    v (mV)
    ena (mV)
    ina (mA/cm2)
    pinf
}

BREAKPOINT {
    SOLVE act METHOD euler
    ina = gbar*p*(v - ena)
}

INITIAL {
    shape[0] = -48
    shape[1] = 6
    p = steady(v)
}

DERIVATIVE act {
    pinf = steady(v)
    p' = (pinf - p)/tau
}

FUNCTION steady(vm (mV)) {
    steady = 1/(1 + exp(-(vm - shape[0])/shape[1]))
}

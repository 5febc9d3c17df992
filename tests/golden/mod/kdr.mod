: Delayed-rectifier potassium channel with Hodgkin-Huxley (1952) kinetics,
: the n gate of the squid axon model, written for the runtime MOD tests.
NEURON {
    SUFFIX kdr
    USEION k READ ek WRITE ik
    RANGE gkbar, gk
}

PARAMETER {
    gkbar = 0.036 (S/cm2) <0,1e9>
}

STATE { n }

ASSIGNED {
    v (mV)
    celsius (degC)
    ek (mV)
    gk (S/cm2)
    ik (mA/cm2)
    ninf
    ntau (ms)
}

BREAKPOINT {
    SOLVE gate METHOD cnexp
    gk = gkbar*n*n*n*n
    ik = gk*(v - ek)
}

INITIAL {
    kinetics(v)
    n = ninf
}

DERIVATIVE gate {
    kinetics(v)
    n' = (ninf - n)/ntau
}

PROCEDURE kinetics(v (mV)) {
    LOCAL a, b, total, phi
    phi = 3^((celsius - 6.3)/10)
    a = .01*expm1ratio(-(v + 55), 10)
    b = .125*exp(-(v + 65)/80)
    total = a + b
    ntau = 1/(phi*total)
    ninf = a/total
}

: x / (exp(x/y) - 1), continued by its limit near x = 0
FUNCTION expm1ratio(x, y) {
    if (fabs(x/y) < 1e-6) {
        expm1ratio = y*(1 - x/y/2)
    } else {
        expm1ratio = x/(exp(x/y) - 1)
    }
}

: Conductance-based synapse with a single exponential decay: each event adds
: its weight to g, which decays with time constant tau; written for the
: runtime MOD tests.
NEURON {
    POINT_PROCESS ExpDecaySyn
    RANGE tau, e, i
    NONSPECIFIC_CURRENT i
}

PARAMETER {
    tau = 0.1 (ms) <1e-9,1e9>
    e = 0 (mV)
}

ASSIGNED {
    v (mV)
    i (nA)
}

STATE { g (uS) }

INITIAL { g = 0 }

BREAKPOINT {
    SOLVE decay METHOD cnexp
    i = g*(v - e)
}

DERIVATIVE decay { g' = -g/tau }

NET_RECEIVE(weight (uS)) { g = g + weight }

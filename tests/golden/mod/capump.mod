: A point source of calcium current (nA) switched on by events, with a
: NET_RECEIVE that reads the local calcium; written for the ion-dynamics tests
: of this repository.
NEURON {
    POINT_PROCESS CaPump
    USEION ca READ cai WRITE ica
    RANGE amp, on, seen
}
PARAMETER { amp = -0.002 (nA) }
ASSIGNED { ica (nA) cai (mM) seen (mM) }
STATE { on }
INITIAL { on = 0  seen = 0 }
BREAKPOINT { ica = amp*on }
NET_RECEIVE(w) {
    on = 1 - on
    seen = cai
}

: Population rate of an ROI's micro cells: every spike of its sources adds one
: to count; at the end of each bin the rate is count per source per bin width
: (ms), in Hz, and the count restarts.
NEURON {
    SUFFIX PopRate
    RANGE bin, rate
}

MIND {
    ROLE MICRO2MACRO
    WRITE_TARGET rate
}

PARAMETER {
    bin = 1 (ms)
}

ASSIGNED {
    count
    sources
    width (ms)
    rate (Hz)
}

INITIAL {
    count = 0
    rate = 0
}

NET_RECEIVE(weight) {
    count = count + 1
}

BREAKPOINT {
    if (sources > 0) {
        rate = count/(sources*width)*1000
    } else {
        rate = 0
    }
    count = 0
}

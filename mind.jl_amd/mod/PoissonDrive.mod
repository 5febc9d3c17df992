: Poisson drive of a point process from its ROI: events at rate rate_scale x E
: (Hz), E taken at the start of each macro step. The macro step draws the
: events of each step from the binding's random stream, rekeyed per step.
NEURON {
    SUFFIX PoissonDrive
    RANGE rate
}

MIND {
    ROLE MACRO2MICRO
    READ_SOURCE E AS e, rate_scale AS scale
}

ASSIGNED {
    e
    scale (Hz)
    rate (Hz)
}

BREAKPOINT {
    rate = scale*e
}

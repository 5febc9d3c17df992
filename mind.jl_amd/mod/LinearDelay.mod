: Linear delayed coupling between ROIs: each connectome edge adds
: weight x E(t - delay) of its source ROI to its target's coupling input. The
: delay is the connectome's, in whole macro steps; the macro step sums a
: target's incoming edges in source order.
NEURON {
    SUFFIX LinearDelay
    RANGE weight
}

MIND {
    ROLE MACRO2MACRO
    READ_SOURCE E AS x
    WRITE_TARGET coupling AS y
}

ASSIGNED {
    weight
    x
    y
}

BREAKPOINT {
    y = weight*x
}

: Submembrane calcium: influx from the calcium current into a shell of the
: given depth, first-order decay toward a resting level. Written for the
: ion-dynamics tests of this repository.
NEURON {
    SUFFIX cadecay
    USEION ca READ ica WRITE cai
    RANGE depth, taur, cainf, influx
}
UNITS { (mA) = (milliamp) (mM) = (milli/liter) (um) = (micron) }
PARAMETER {
    depth = 0.1 (um)
    taur = 80 (ms)
    cainf = 5e-5 (mM)
}
STATE { cai (mM) }
ASSIGNED { ica (mA/cm2) influx (mM/ms) }
INITIAL { cai = cainf }
BREAKPOINT { SOLVE shell METHOD euler }
DERIVATIVE shell {
    influx = -10000*ica/(2*96485.3321*depth)
    if (influx < 0) { influx = 0 }
    cai' = influx + (cainf - cai)/taur
}

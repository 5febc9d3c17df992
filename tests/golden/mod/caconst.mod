: A constant calcium current density; written for the ion-dynamics tests of
: this repository.
NEURON {
    SUFFIX caconst
    USEION ca WRITE ica
    RANGE amp
}
PARAMETER { amp = -1e-3 (mA/cm2) }
ASSIGNED { ica (mA/cm2) }
BREAKPOINT { ica = amp }

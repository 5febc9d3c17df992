: Wilson-Cowan neural mass of one ROI (Wilson & Cowan 1972): the mean activities
: E and I of coupled excitatory and inhibitory populations,
:   tau_e dE/dt = -E + S_e(c_ee E - c_ei I + P + u)
:   tau_i dI/dt = -I + S_i(c_ie E - c_ii I + Q)
: with S(x) = 1/(1 + exp(-a (x - theta))) - 1/(1 + exp(a theta)). The input u is
: the delayed coupling from other ROIs plus k_micro times the ROI's micro
: population rate relative to rate_scale (the rate that E = 1 drives). One macro
: step of length dt advances E and I by the classical fourth-order Runge-Kutta
: method.
NEURON {
    SUFFIX WilsonCowan
    RANGE tau_e, tau_i, c_ee, c_ei, c_ie, c_ii, a_e, theta_e, a_i, theta_i, P, Q
    RANGE k_micro, rate_scale, E0, I0, coupling, rate
}

MIND {
    ROLE REGION
    EXPOSURE E, I, rate_scale
    READ_TARGET coupling, rate
}

PARAMETER {
    tau_e = 10 (ms)
    tau_i = 10 (ms)
    c_ee = 16
    c_ei = 12
    c_ie = 15
    c_ii = 3
    a_e = 1.3
    theta_e = 4
    a_i = 2
    theta_i = 3.7
    P = 1.25
    Q = 0
    k_micro = 0
    rate_scale = 100 (Hz)
    E0 = 0.1
    I0 = 0.05
}

STATE { E I }

ASSIGNED {
    coupling
    rate (Hz)
    dE (/ms)
    dI (/ms)
}

INITIAL {
    E = E0
    I = I0
}

BREAKPOINT {
    SOLVE rk4step METHOD after_cvode
}

PROCEDURE rk4step() {
    LOCAL u, k1e, k1i, k2e, k2i, k3e, k3i, k4e, k4i
    u = coupling + k_micro*(rate/rate_scale)
    slopes(E, I, u)
    k1e = dE
    k1i = dI
    slopes(E + 0.5*dt*k1e, I + 0.5*dt*k1i, u)
    k2e = dE
    k2i = dI
    slopes(E + 0.5*dt*k2e, I + 0.5*dt*k2i, u)
    k3e = dE
    k3i = dI
    slopes(E + dt*k3e, I + dt*k3i, u)
    k4e = dE
    k4i = dI
    E = E + dt/6*(k1e + 2*k2e + 2*k3e + k4e)
    I = I + dt/6*(k1i + 2*k2i + 2*k3i + k4i)
}

: time derivatives of E and I at (e, i) under input u
PROCEDURE slopes(e, i, u) {
    dE = (-e + response(c_ee*e - c_ei*i + P + u, a_e, theta_e))/tau_e
    dI = (-i + response(c_ie*e - c_ii*i + Q, a_i, theta_i))/tau_i
}

: population response, shifted so that response(0) = 0
FUNCTION response(x, a, theta) {
    response = 1/(1 + exp(-a*(x - theta))) - 1/(1 + exp(a*theta))
}

"""Physical constants and safety floors of the nondimensional system.

Scaling (as the reference, ``src/jaxrans/core/constants.py:1-30``): density by
rho_inf, velocity by a_inf, length by the chord.  The freestream is then
rho = 1, |V| = M, p = 1/gamma, and the dynamic viscosity mu_inf = M / Re.
"""

from __future__ import annotations

GAMMA_AIR = 1.4
PRANDTL_LAMINAR = 0.72
PRANDTL_TURBULENT = 0.9

# Detection thresholds of a non-physical state (src/jaxrans/core/constants.py:49-52).
MIN_DENSITY = 1.0e-10
MIN_PRESSURE = 1.0e-10

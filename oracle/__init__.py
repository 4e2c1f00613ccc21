"""CPU restatement of the reference's Newton-Krylov hot path, in FP64 NumPy.

THIS PACKAGE IS TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it, and only as the checker
or as the timed CPU baseline -- never as the thing measured on the GPU and
never from the product package ``jaxrans_amd``.

Modules and the reference files they follow (paths relative to the
reference repository):

* ``dual``     -- forward-mode dual numbers with JAX's JVP tie rules
  (``src/jaxrans/solvers/implicit.py:105-108`` uses ``jax.linearize``);
* ``physics``  -- limiters, HLLC/Roe, boundary states, viscous stress, SA
  (``numerics/limiters.py``, ``numerics/riemann.py``, ``physics/euler.py``,
  ``numerics/boundary.py``, ``physics/navier_stokes.py``,
  ``physics/spalart_allmaras.py``);
* ``residual`` -- ghost fill, MUSCL, face gradients, SA row, residual, norm,
  wave speeds (``numerics/residual.py``, ``reconstruction.py``,
  ``gradients.py``, ``turbulence.py``);
* ``solver``   -- line preconditioner (Thomas and cyclic reduction), GMRES,
  Newton step, forces (``core/linalg.py``, ``solvers/preconditioners.py``,
  ``solvers/krylov.py``, ``solvers/implicit.py``, ``post/forces.py``).
* ``cpu``      -- ctypes binding of ``cpu/newton2d.cpp``, a C++/OpenMP
  restatement of the 2D residual, tangent, wave speeds, block-Thomas line
  preconditioner, GMRES and Newton step above (the benchmark's CPU
  baseline; pinned against this package by ``tests/test_oracle_cpu.py``).

Pinning (each checked in ``tests/test_oracle_*.py``):

* fluxes: consistency, conservation, linearity in S, rotation, stationary
  contact, Roe jump property;
* limiters: psi(a, a) = a, symmetry, TVD bounds, zero at extrema;
* residual: freestream preservation off the walls, O-grid up/down mirror
  symmetry, CL = 0 at zero incidence on a symmetric grid;
* tangent: central differences of the residual, linearity, first-order
  Euler homogeneity J(U) U = R(U) on interior faces;
* SA pointwise: an independent scalar transcription of the TMR formulas at
  seven branch points (the reference's own sample points), c_w1 and the
  freestream eddy viscosity;
* line solve: Thomas and cyclic reduction against a dense solve;
* GMRES: full-subspace exactness and the Givens solve against
  ``numpy.linalg.lstsq``.

Parity unpinned: JAX's tie conventions at exactly non-smooth points of the
tangent (``max``/``min`` ties, ``abs`` at 0), which no reference test or
golden vector fixes; the oracle and the device follow JAX's rules.
"""

"""MI355X-native implementation of jax-rans' pseudo-transient Newton-Krylov hot path.

Public API, as the reference's ``src/jaxrans/__init__.py:22-26``:
``load_case(directory, overrides)``, ``run(case)`` and ``sensitivity(result)``;
``run_polar`` batches freestreams on one grid.  All numerics run in the HIP library
``libjrhip.so`` (``jaxrans_amd.device``).
"""

from .case import Case, CaseError, load_case  # noqa: F401
from .simulation import Forces, Result, run, run_polar  # noqa: F401
from .derivatives import Sensitivity, sensitivity  # noqa: F401

__version__ = "0.1.0"

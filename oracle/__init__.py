"""CPU oracle for the PM force step: test infrastructure only.

This package is a float64 NumPy restatement of PM++'s single-device particle-mesh
operators (CIC enmesh / scatter / gather and their adjoints, the stable cell sort
contract, the spectral Poisson + gradient force, and the power-spectrum metric).
Each function cites the reference file:line it follows.

It is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it.  The product path
(``pmpp_amd``) never imports anything from here and fails loudly when its HIP
library is missing.

Pinning (what fixes each function independently of retyping its formula) is
listed in the module docstrings and exercised in ``tests/test_oracle_*.py``.
No function of this oracle is "parity unpinned" except ``lpt`` helpers marked
as such in ``oracle/ics.py``.
"""

from .cic import enmesh, scatter, gather, scatter_adj, gather_adj  # noqa: F401
from .spectral import (  # noqa: F401
    kvec, particle_nyquist_masks, neg_k, gravity, gravity_vjp, green_multiplier,
)
from .sort import cell_keys, sort_cells  # noqa: F401
from . import power_spectrum  # noqa: F401  (module: delta_to_pk, power_spectrum)

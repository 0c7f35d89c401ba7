"""MI355X (gfx950) Newton-step provider for SDPX's dense Schur route.

Host mirror of the provider a Julia extension would register behind SDPX's
``la_backend`` / ``kkt_backend`` seams; all array work runs in the HIP library
``libsdpx_mi355x.so`` (C-ABI in include/sdpx_mi355x.h).
"""

from .backend import MI355XDenseBackend, SDPProblem, ingest
from .ipm import SDPResult, solve
from .policy import SolverOptions, settings

__all__ = ["MI355XDenseBackend", "SDPProblem", "ingest", "SDPResult", "solve", "SolverOptions", "settings"]

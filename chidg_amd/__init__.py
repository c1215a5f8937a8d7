"""MI355X (gfx950) implementation of the linear solve inside ChiDG's
Newton-Krylov loop: block-sparse matvec, block ILU(0) and restarted FGMRES,
behind ChiDG's own interfaces (see include/chidg_hip.h for the C ABI)."""
from . import pattern
from .linsolve import (Context, chidg_matrix, chidg_mv, create_preconditioner, default_context, fgmres,
                       precon_identity, precon_ILU0)

__all__ = ["pattern", "Context", "chidg_matrix", "chidg_mv", "create_preconditioner", "default_context",
           "fgmres", "precon_identity", "precon_ILU0"]

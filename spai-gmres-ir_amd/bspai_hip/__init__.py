"""gfx950 SPAI-GMRES-IR: Python front end over libbspai_hip.so."""
from . import problems  # noqa: F401
from ._lib import LIB_PATH, LibraryMissing, load  # noqa: F401
from .api import (Context, Matrix, adjoint_solve, construct_spai, csr_from_entries,  # noqa: F401
                  default_context, evaluate_spai_score, left_preconditioned_gmres,
                  refine_with_preconditioner, spai_gmres_ir)

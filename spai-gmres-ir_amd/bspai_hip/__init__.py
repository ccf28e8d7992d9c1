"""gfx950 SPAI-GMRES-IR: Python front end over libbspai_hip.so."""
from . import problems  # noqa: F401
from ._lib import LIB_PATH, LibraryMissing, load  # noqa: F401
from .api import (Context, Matrix, MatrixMarketError, adjoint_solve, construct_spai,  # noqa: F401
                  csr_from_entries, default_context, evaluate_spai_score,
                  left_preconditioned_gmres, parse_matrix_market, read_matrix_market,
                  refine_with_preconditioner, spai_gmres_ir)

"""MI355X-native Toeplitz subspace normal operator and subspace backprojection
(drop-in for MRISubspaceRecon.jl's NFFTNormalOp / calculate_backprojection)."""
from . import synth  # noqa: F401

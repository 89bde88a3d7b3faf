"""altar_amd: the CATMIP inner loop of AlTar on MI355X (gfx950).

Forward models (Mogi, CDM, linear), the L2 data log-likelihood with
Cd^-1/2 whitening on the matrix cores, and the Metropolis / COV annealing
step, as hand-written HIP kernels behind a C ABI (include/altar_amd.h).
"""
from .job import Job
from .distributions import Contiguous, Gaussian, Uniform

__version__ = "0.1.0"

__all__ = ["Job", "Contiguous", "Gaussian", "Uniform"]

"""botorch_amd: BoTorch's GP-posterior / qLogNEI / exact-MLL hot path on MI355X.

Host-side mirror of BoTorch's interfaces (SingleTaskGP.posterior,
qLogNoisyExpectedImprovement.forward, optimize_acqf, fit_gpytorch_mll) over
hand-written HIP kernels for gfx950 in libbotorch_amd.so (C ABI:
include/botorch_amd.h).
"""
from .acquisition import (TAU_MAX, TAU_RELU, BatchedQLogNEI, prune_inferior_points,
                          qLogNoisyExpectedImprovement)
from .exceptions import (BotorchError, BotorchWarning, InputDataError, NanError, NotPSDError,
                         NumericalWarning, OptimizationGradientError, OptimizationWarning,
                         UnsupportedError)
from .gp import (ConstantMean, GaussianLikelihood, GPyTorchPosterior, GreaterThan,
                 LogNormalPrior, MaternKernel, Positive, RBFKernel, ScaleKernel, SingleTaskGP,
                 Standardize, psd_safe_cholesky)
from .sampling import SobolQMCNormalSampler, draw_sobol_normal_samples, draw_sobol_samples

__version__ = "0.1.0"

"""gprx_amd: gfx950 (MI355X) backend for gprx's Gaussian-process hot path.

The compute lives in libgprx_hip.so (hand-written HIP kernels behind the C
ABI of include/gprx_hip.h); this package is the host-side mirror of gprx's
model API over that ABI.  There is no CPU fallback: calls fail loudly when
the library or a GPU is missing.
"""
from .errors import (CholeskyFailed, DeviceError, EmptyInput, GprError, InsufficientData, InvalidConfig,
                     InvalidHyperparameter, InvalidInterval, InvalidNoiseVariance, LengthMismatch,
                     NonFiniteInput, NonFiniteKernelValue, OptimizationNotConverged, Unsupported)
from .gpr import (Fallback, FittedGpr, GaussianLikelihood, Gpr, JitterPolicy, Precision,
                  Prediction, VarianceKind)
from .batch import GprBatch
from .persist import FORMAT_VERSION, PersistFailed, UnsupportedPersistVersion
from .optimizer import Fixed, Lbfgs, OptResult, TrustRegion
from .sgpr import FittedSgpr, Sgpr
from .svgp import FittedSvgp, Svgp, prior_params
from .kernels import (ConstantKernel, KernelSpec, PeriodicKernel, Product, RbfArdKernel, RbfKernel,
                      Sum)

__all__ = [
    "CholeskyFailed", "DeviceError", "EmptyInput", "GprError", "InvalidHyperparameter",
    "InvalidNoiseVariance", "LengthMismatch", "NonFiniteInput", "NonFiniteKernelValue",
    "Unsupported", "Fallback", "FittedGpr", "GaussianLikelihood", "Gpr", "JitterPolicy", "Precision",
    "Prediction", "VarianceKind", "ConstantKernel", "KernelSpec", "PeriodicKernel", "Product",
    "RbfArdKernel", "RbfKernel", "Sum", "Sgpr", "FittedSgpr", "GprBatch", "Svgp", "FittedSvgp", "prior_params",
    "Fixed", "Lbfgs", "OptResult", "TrustRegion", "OptimizationNotConverged", "InvalidConfig", "InvalidInterval",
    "InsufficientData", "FORMAT_VERSION", "PersistFailed", "UnsupportedPersistVersion",
]

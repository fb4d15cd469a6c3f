"""GprError variants raised by the Python mirror (src/error.rs)."""
from __future__ import annotations

from . import _lib as L


class GprError(RuntimeError):
    status = -1

    def __init__(self, message: str, stage: int = 0, jitter: float = 0.0, matrix_size: int = 0):
        super().__init__(message)
        self.stage, self.jitter, self.matrix_size = stage, jitter, matrix_size


class CholeskyFailed(GprError):
    status = L.CHOLESKY_FAILED


class NonFiniteKernelValue(GprError):
    status = L.NON_FINITE_KERNEL_VALUE


class NonFiniteInput(GprError):
    status = L.NON_FINITE_INPUT


class InvalidHyperparameter(GprError):
    status = L.INVALID_HYPERPARAMETER


class InvalidNoiseVariance(GprError):
    status = L.INVALID_NOISE_VARIANCE


class LengthMismatch(GprError):
    status = L.LENGTH_MISMATCH


class EmptyInput(GprError):
    status = L.EMPTY_INPUT


class Unsupported(GprError):
    status = L.UNSUPPORTED


class DeviceError(GprError):
    status = L.DEVICE_ERROR


class OptimizationNotConverged(GprError):
    status = L.NOT_CONVERGED


class InvalidConfig(GprError):
    status = L.INVALID_CONFIG


class InvalidInterval(GprError):
    status = L.INVALID_INTERVAL


class InsufficientData(GprError):
    status = L.INSUFFICIENT_DATA


_BY_STATUS = {c.status: c for c in (CholeskyFailed, NonFiniteKernelValue, NonFiniteInput,
                                      InvalidHyperparameter, InvalidNoiseVariance, LengthMismatch,
                                      EmptyInput, Unsupported, DeviceError, OptimizationNotConverged,
                                      InvalidConfig, InvalidInterval, InsufficientData)}


def raise_for(rc: int, err) -> None:
    if rc == L.OK:
        return
    cls = _BY_STATUS.get(rc, GprError)
    if isinstance(err, str):
        raise cls(err)
    if err is None:
        raise cls(f"status {rc}")
    raise cls(err.message.decode(errors="replace"), err.stage, err.jitter, err.matrix_size)

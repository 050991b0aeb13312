"""Exception types with the meaning of the reference's (botorch/exceptions/errors.py
and linear_operator's NotPSDError / NanError), so callers' fallbacks keep working."""


class BotorchError(Exception):
    """Base botorch_amd exception (exceptions/errors.py:16)."""


class UnsupportedError(BotorchError):
    """Shape or option outside the implemented envelope (exceptions/errors.py:40)."""


class InputDataError(BotorchError):
    """Invalid input data."""


class OptimizationGradientError(BotorchError, RuntimeError):
    """NaN gradient during candidate generation (generation/gen.py:471-479)."""

    def __init__(self, *args, current_x=None):
        super().__init__(*args)
        self.current_x = current_x


class NotPSDError(RuntimeError):
    """Cholesky failed after all jitter tries (linear_operator NotPSDError)."""


class NanError(RuntimeError):
    """NaN in a factorisation input or in posterior samples (linear_operator NanError)."""


class BotorchWarning(Warning):
    """Base botorch_amd warning (exceptions/warnings.py:12)."""


class OptimizationWarning(BotorchWarning):
    """Optimisation-related warnings (exceptions/warnings.py:38)."""


class NumericalWarning(RuntimeWarning):
    """Jitter added to a factorisation (linear_operator NumericalWarning)."""

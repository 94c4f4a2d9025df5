"""Exception and warning types of the drop-in calculator.

Names mirror upstream (``src/dxtb/_src/typing/exceptions/scf.py:27-36`` and
tad_mctc.exceptions) so calling code can catch the same types.
"""


class DtypeError(ValueError):
    """Tensor of wrong dtype."""


class DeviceError(RuntimeError):
    """Tensor on the wrong device."""


class SCFConvergenceError(RuntimeError):
    """SCF did not converge and convergence was forced."""


class SCFConvergenceWarning(UserWarning):
    """SCF did not converge within maxiter."""


class PropertyNotImplementedError(NotImplementedError):
    """Requested property is not provided by this calculator."""


class NativeLibraryError(RuntimeError):
    """The HIP library is missing or a native call failed."""

"""MI355X-native GFN1-xTB single point (drop-in for dxtb's GFN1 path).

Public surface:

* ``GFN1Calculator`` -- upstream-compatible calculator (energy, forces,
  charges) on PyTorch-ROCm tensors, backed by hand-written HIP kernels.
* ``GFN1_XTB`` -- name of the bundled parametrisation.
* ``RaggedServer`` -- continuous-batching serving loop (``dxtb_amd.serving``).
"""

from .calculator import GFN1Calculator  # noqa: F401
from .exceptions import (  # noqa: F401
    DeviceError,
    DtypeError,
    NativeLibraryError,
    SCFConvergenceError,
    SCFConvergenceWarning,
)
from .param import GFN1_XTB  # noqa: F401

__version__ = "0.1.0"

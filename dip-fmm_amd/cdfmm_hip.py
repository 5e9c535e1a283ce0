"""Python front-end of the HipFull backend (the reference's ``cdfmm`` interface).

The extension ``_cdfmm_hip`` links the HIP runtime.  PyTorch-ROCm ships its
own copy of that runtime; two copies in one process cannot share the GPU, so
torch (when installed) is imported first and the extension then binds to the
runtime already loaded.
"""

try:  # one HIP runtime per process
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for host-memory use
    torch = None

from _cdfmm_hip import *  # noqa: F401,F403
from _cdfmm_hip import _laplace_derivatives, _m2l_class_matrix  # noqa: F401

"""tenex_amd -- MI355X (gfx950) transfer-entropy engine with the tnx API.

    import tenex_amd as tnx
    res = tnx.TransferEntropyEngine(data, names).compute(accelerator="gpu")
    grn = tnx.NetWeaver(res).infer(method="surrogate_test")
"""
__version__ = "0.1.0"

from ._log import is_verbose, set_verbose
from .kernels import TEKernel, auto_select, get_kernel, registered_kernels
from .result import TransferEntropyResult
from .transferentropy import TransferEntropyEngine

__all__ = [
    "TransferEntropyEngine", "TransferEntropyResult", "TEKernel", "auto_select", "get_kernel",
    "registered_kernels", "set_verbose", "is_verbose", "__version__",
]

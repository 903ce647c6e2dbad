"""tenex_amd -- MI355X (gfx950) transfer-entropy engine with the tnx API.

    import tenex_amd as tnx
    res = tnx.TransferEntropyEngine(data, names).compute(accelerator="gpu")
    grn = tnx.NetWeaver(res).infer(method="surrogate_test")

Public names mirror the reference package tnx (tenex/__init__.py:16-45).
"""
__version__ = "0.1.0"

from ._log import is_verbose, set_verbose
from .binarize import available_binarizers, binarize
from .inference import GRN, NetWeaver, available_methods
from .inference.surrogate_test import SurrogateTestResult
from .inference.trace import TRACEResult
from .io import ScRnaData, load_scrna
from .kernels import TEKernel, auto_select, get_kernel, registered_kernels
from .pipeline import Pipeline, PipelineResult
from .result import TransferEntropyResult
from .transferentropy import TransferEntropyEngine

__all__ = [
    "TransferEntropyEngine", "TransferEntropyResult", "Pipeline", "PipelineResult", "NetWeaver",
    "GRN", "SurrogateTestResult", "TRACEResult", "available_methods", "TEKernel", "auto_select",
    "get_kernel", "registered_kernels", "set_verbose", "is_verbose", "__version__", "binarize",
    "available_binarizers", "load_scrna", "ScRnaData",
]

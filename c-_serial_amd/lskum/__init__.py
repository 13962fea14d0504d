"""MI355X q-LSKUM iteration: host interface over liblskum.so (HIP, gfx950)."""
from ._lib import LskumError, FP64, MIXED, REORDER_NONE, REORDER_RCM, REORDER_HILBERT  # noqa: F401
from .cloud import Cloud, WORKLOADS, workload  # noqa: F401
from .batch import Batch, batch_workload  # noqa: F401

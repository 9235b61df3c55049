"""rdmce_amd: maximal clique enumeration on AMD Instinct MI355X.

Host-side mirror of the reference's interface (``Graph``, ``MceGpuSolver``,
``BkSolverWrapper``) over the C ABI of librdmce.so, whose HIP kernels do all
the enumeration.
"""
from ._lib import (MceError, ROOT_EDGE, ROOT_VERTEX, LIB_PATH, lib)  # noqa: F401
from .graph import Graph, OrderType, batch_graph_params, bin_path_for  # noqa: F401
from .solver import (Batch, BatchOutput, BkSolverWrapper, MceGpuSolver, MceOutput, Plan,  # noqa: F401
                     choose_root_mode, device_count, solve)

__all__ = ["Graph", "OrderType", "MceGpuSolver", "BkSolverWrapper", "solve", "Plan", "Batch", "BatchOutput",
           "choose_root_mode",
           "MceOutput", "MceError",
           "ROOT_EDGE", "ROOT_VERTEX", "batch_graph_params", "bin_path_for", "device_count", "lib", "LIB_PATH"]

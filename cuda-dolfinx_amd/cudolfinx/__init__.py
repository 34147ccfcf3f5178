"""cudolfinx -- the cuDOLFINx assembly API on MI355X (HIP / hipRTC, gfx950).

Drop-in names of python/cudolfinx/__init__.py: `form`, `CUDAAssembler`,
`CUDAForm`, `CUDAMatrix`, `CUDAVector`, `CUDADirichletBC`, `Coefficient`;
`CUDASystem` and `CUDABatch` fuse several forms into one launch.  The work happens
in libmifem.so through its C ABI; forms, meshes and function spaces come
from `mifem` (the DOLFINx/FFCx/Basix stand-in of this repository).
"""
from .assemble import CUDAAssembler
from .batch import CUDABatch
from .bcs import CUDADirichletBC
from .coefficient import Coefficient
from .context import CUDAContext, DeviceArray, get_cuda_context
from .form import BlockCUDAForm, CUDAForm, form
from .graph import CUDAGraph
from .la import CUDAMatrix, CUDAVector
from . import jit
from .system import CUDASystem

__all__ = ["CUDAAssembler", "CUDADirichletBC", "CUDAContext", "DeviceArray", "get_cuda_context",
           "CUDAForm", "BlockCUDAForm", "form", "CUDAMatrix", "CUDAVector", "CUDAGraph", "Coefficient", "CUDASystem",
           "CUDABatch"]

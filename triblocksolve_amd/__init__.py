"""triblocksolve_amd -- MI355X (gfx950) line-implicit Jacobi block solver.

The product is the HIP library libtriblock_hip.so (C ABI: include/triblock_hip.h)
and the C++ host API / triblock.exe built on it (include/{Platform,Mesh,Jacobian}.hxx).
This Python package is a thin ctypes layer over that C ABI for tests and the
benchmark: device memory and events come from the library itself, every
computation runs in the HIP library, and a missing library raises instead of
falling back to anything.
"""
from ._capi import (TBS_FP32, TBS_FP64, TBS_STEP_NORM, TBS_STEP_RESIDUAL, TBS_STEP_THOMAS,  # noqa: F401
                    TbsError, load, LIB_PATH)
from .solver import (Context, DeviceBuffer, StructuredMesh, HostMesh, Plan, Batch, Server, Gmres,  # noqa: F401
                     generate_host, structured_mesh, ragged_mesh, batch_shapes, read_mesh, write_mesh,
                     read_jacobian, write_jacobian, model_bytes, reference_model_bytes, variant_name,
                     factor_plans)

__all__ = ["Context", "DeviceBuffer", "StructuredMesh", "HostMesh", "Plan", "Batch", "Server", "Gmres", "TBS_FP32", "TBS_FP64",
           "TBS_STEP_THOMAS", "TBS_STEP_RESIDUAL", "TBS_STEP_NORM",
           "TbsError", "load", "generate_host", "structured_mesh", "ragged_mesh", "batch_shapes", "read_mesh", "write_mesh",
           "read_jacobian", "write_jacobian", "model_bytes", "reference_model_bytes", "variant_name",
           "factor_plans"]

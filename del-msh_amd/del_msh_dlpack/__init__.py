"""del_msh_dlpack -- MI355X (gfx950) drop-in for the LBVH / ray-cast /
differentiable-rasterizer entry points of del-msh's del_msh_dlpack package.

Module and function names, argument meaning and tensor conventions follow the
reference package (del-msh-dlpack/del_msh_dlpack/<Module>/torch.py); the
kernels behind them are libdmsh's HIP kernels, reached through the C ABI in
include/dmsh.h.  Only ROCm device tensors are accepted.
"""

from ._lib import raise_on_device_errors, device_errors  # noqa: F401

__all__ = ["raise_on_device_errors", "device_errors"]

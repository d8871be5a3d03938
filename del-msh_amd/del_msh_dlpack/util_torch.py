"""Tensor checks with the conventions of del_msh_dlpack/util_torch.py in the
reference (contiguous, exact shape / dtype / device, no requires_grad)."""

import torch


def to_dlpack_safe(t: torch.Tensor, stream_ptr: int):
    """DLPack export with the reference's stream convention (legacy default = 1)."""
    assert t.is_contiguous()
    dev = t.device.type
    if dev == "cuda":
        if stream_ptr == 0:
            stream_ptr = 1
        return t.__dlpack__(stream=stream_ptr)
    if dev == "cpu":
        return t.__dlpack__()
    raise RuntimeError(f"Unsupported device for DLPack export: {t.device}")


def assert_shape_dtype_device(
    t: torch.Tensor, shape: tuple, dtype: torch.dtype, device: torch.device
):
    assert t.shape == shape, (t.shape, shape)
    assert t.dtype == dtype, (t.dtype, dtype)
    assert t.device == device, (t.device, device)
    assert t.is_contiguous()
    assert not t.requires_grad


def as_tensor(x) -> torch.Tensor:
    """Accepts a tensor or a DLPack capsule / producer (the reference's native
    entry points take capsules)."""
    if isinstance(x, torch.Tensor):
        return x
    return torch.from_dlpack(x)

"""Vtx2Xyz helpers (reference: del_msh_dlpack/Vtx2Xyz/torch.py)."""

import torch


def transform_affine(vtx2xyz: torch.Tensor, transform: torch.Tensor) -> torch.Tensor:
    """p' = (M @ [p, 1])[:3] for every point; M is a row-major 4x4 affine map."""
    ones = torch.ones((vtx2xyz.shape[0], 1), dtype=vtx2xyz.dtype, device=vtx2xyz.device)
    return (torch.cat([vtx2xyz, ones], dim=1) @ transform.T)[:, 0:3].clone()

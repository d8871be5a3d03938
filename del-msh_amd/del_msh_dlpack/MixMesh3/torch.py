"""MixMesh3.to_polyhedral_mesh (reference: del_msh_dlpack/MixMesh3/torch.py:
38-90, native mix_mesh.rs): the four connectivity arrays as one CSR mesh,
tets first, then pyramids, prisms and hexes, each in its own order.  A
layout change, done with device tensor copies."""

import torch


def to_polyhedral_mesh(tet2vtx: torch.Tensor, pyrmd2vtx: torch.Tensor, prism2vtx: torch.Tensor,
                       hex2vtx: torch.Tensor):
    """Returns (elem2idx_offset [num_elem+1] u32, idx2vtx [num_idx] u32)."""
    device = tet2vtx.device
    parts = []
    for a, n in ((tet2vtx, 4), (pyrmd2vtx, 5), (prism2vtx, 6), (hex2vtx, 8)):
        assert a.dtype == torch.uint32 and a.dim() == 2 and a.shape[1] == n and a.device == device
        parts.append((a, n))
    counts = [a.shape[0] for a, _ in parts]
    sizes = torch.cat([torch.full((c,), n, dtype=torch.int64, device=device)
                       for (a, n), c in zip(parts, counts)])
    off = torch.zeros(sizes.shape[0] + 1, dtype=torch.int64, device=device)
    off[1:] = torch.cumsum(sizes, 0)
    idx2vtx = torch.cat([a.reshape(-1).view(torch.int32) for a, _ in parts]).view(torch.uint32)
    return off.to(torch.int32).view(torch.uint32), idx2vtx

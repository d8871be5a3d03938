"""GPU: the allocation-free build (dmsh_build_trimesh3_noalloc, scratch in a
fixed device array inside the library) gives the same trees as
dmsh_build_trimesh3 with a caller workspace, on two calls in a row (the
second with a smaller mesh, so stale scratch from the first would show), and
refuses a mesh too large for its scratch before enqueuing anything."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives

pytestmark = pytest.mark.gpu

DMSH_ERR_TOO_LARGE = 4
SCRATCH_BYTES = 64 << 20


def _build(L, stream, tri, vtx, bits, noalloc):
    n = tri.shape[0]
    nodes = torch.full((2 * n - 1, 3), 7, dtype=torch.uint32, device=tri.device)
    aabb = torch.full((2 * n - 1, 6), 7.0, dtype=torch.float32, device=tri.device)
    if noalloc:
        rc = L.dmsh_build_trimesh3_noalloc(stream, n, tri.data_ptr(), vtx.shape[0], vtx.data_ptr(),
                                           bits, nodes.data_ptr(), aabb.data_ptr())
        _lib.check(rc, "dmsh_build_trimesh3_noalloc")
    else:
        ws = _lib.workspace(L.dmsh_build_workspace_bytes(n, bits), tri.device)
        _lib.check(L.dmsh_build_trimesh3(stream, n, tri.data_ptr(), vtx.shape[0], vtx.data_ptr(),
                                         bits, nodes.data_ptr(), aabb.data_ptr(), ws.data_ptr(),
                                         ws.numel()), "dmsh_build_trimesh3")
    torch.cuda.synchronize()
    return nodes.cpu().numpy(), aabb.cpu().numpy()


def test_noalloc_build_equals_workspace_build(gpu, oracle):
    L = _lib.lib()
    stream = torch.cuda.current_stream(gpu).cuda_stream
    meshes = [
        (primitives.octahedral_sphere(8, noise_sigma=0.01, seed=1), 63),  # 524,288 triangles
        (primitives.torus_zup(1.3, 0.4, 64, 32), 30),  # 4,096 triangles, second and smaller
    ]
    for (tri, vtx), bits in meshes:
        assert L.dmsh_build_workspace_bytes(tri.shape[0], bits) <= SCRATCH_BYTES
        t = torch.from_numpy(tri).to(gpu)
        v = torch.from_numpy(vtx).to(gpu)
        got = _build(L, stream, t, v, bits, noalloc=True)
        want = _build(L, stream, t, v, bits, noalloc=False)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
        if tri.shape[0] < 10_000:
            wbn, waabb = oracle.build_trimesh3(tri, vtx, morton_bits=bits)
            assert np.array_equal(got[0], wbn)
            assert np.array_equal(got[1].view(np.uint32), waabb.view(np.uint32))
    assert _lib.device_errors() == 0


def test_noalloc_build_refuses_meshes_beyond_its_scratch(gpu):
    L = _lib.lib()
    n = 1 << 21  # 2^21 triangles need about 182 MB with 63-bit codes
    assert L.dmsh_build_workspace_bytes(n, 63) > SCRATCH_BYTES
    small = torch.zeros(16, dtype=torch.float32, device=gpu)
    # rejected on the host from the sizes alone: nothing is read or written
    rc = L.dmsh_build_trimesh3_noalloc(torch.cuda.current_stream(gpu).cuda_stream, n,
                                       small.data_ptr(), 3, small.data_ptr(), 63, small.data_ptr(),
                                       small.data_ptr())
    assert rc == DMSH_ERR_TOO_LARGE
    assert "allocation-free" in L.dmsh_status_string(rc).decode()
    assert L.dmsh_build_trimesh3_noalloc(None, 10, small.data_ptr(), 3, small.data_ptr(), 31,
                                         small.data_ptr(), small.data_ptr()) == 1
    torch.cuda.synchronize()
    assert torch.count_nonzero(small).item() == 0

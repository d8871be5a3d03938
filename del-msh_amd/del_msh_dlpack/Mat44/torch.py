"""4x4 transform helpers (reference: del_msh_dlpack/Mat44/torch.py), row-major
matrices applied as p' = M @ [p, 1]."""

import math

import torch


def _mat(rows, device=None, dtype=torch.float32):
    return torch.tensor(rows, device=device, dtype=dtype)


def from_x_rotation(angle: float, device=None, dtype=torch.float32) -> torch.Tensor:
    c, s = math.cos(angle), math.sin(angle)
    return _mat([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]], device, dtype)


def from_translation(tx: float, ty: float, tz: float, device=None, dtype=torch.float32):
    return _mat([[1, 0, 0, tx], [0, 1, 0, ty], [0, 0, 1, tz], [0, 0, 0, 1]], device, dtype)


def from_scale(sx: float, sy: float, sz: float, device=None, dtype=torch.float32):
    return _mat([[sx, 0, 0, 0], [0, sy, 0, 0], [0, 0, sz, 0], [0, 0, 0, 1]], device, dtype)


def from_uniform_scale(s, device):
    return from_scale(s, s, s, device)


def from_transform_ndc2pix(img_shape, device=None, dtype=torch.float32):
    """NDC [-1,1]^3 -> pixel coordinates (x right, y down) and depth [0, 1];
    img_shape = (width, height)."""
    w, h = float(img_shape[0]), float(img_shape[1])
    return _mat([[0.5 * w, 0, 0, 0.5 * w], [0, -0.5 * h, 0, 0.5 * h], [0, 0, 0.5, 0.5],
                 [0, 0, 0, 1]], device, dtype)


def from_transfrom_world2unit(vtx2xyz, device):
    """translate(0.5) @ scale(1 / longest extent) @ translate(-box centre): the
    bounding box of the points, centred in the unit cube (reference
    Mat44/torch.py:107-117; composed as products of 4x4 matrices there too)."""
    xyz_min = vtx2xyz.min(dim=0).values
    xyz_max = vtx2xyz.max(dim=0).values
    scale = 1.0 / (xyz_max - xyz_min).max().item()
    c = (xyz_min + xyz_max) * 0.5
    m1 = from_translation(-c[0], -c[1], -c[2], device)
    m2 = from_uniform_scale(scale, device)
    m3 = from_translation(0.5, 0.5, 0.5, device)
    return m3 @ m2 @ m1


def from_fit_vtx2xyz_into_unit_cube(vtx2xyz: torch.Tensor) -> torch.Tensor:
    """Uniform scale s = 1 / max extent and translation -center * s + 0.5, mapping
    the bounding box of the points into [0, 1]^3 (reference Mat44/torch.py:120-140)."""
    assert vtx2xyz.shape[1] == 3 and vtx2xyz.dtype == torch.float32
    lo = vtx2xyz.min(dim=0).values
    hi = vtx2xyz.max(dim=0).values
    center = (lo + hi) * 0.5
    s = 1.0 / (hi - lo).max()
    mat = torch.zeros(4, 4, dtype=torch.float32, device=vtx2xyz.device)
    mat[0, 0] = s
    mat[1, 1] = s
    mat[2, 2] = s
    mat[3, 3] = 1.0
    mat[:3, 3] = -center * s + 0.5
    return mat


def perspective(fovy_deg: float, aspect: float, near: float, far: float,
                device=None, dtype=torch.float32) -> torch.Tensor:
    """Symmetric-frustum projection for a camera looking down -z, with the
    depth convention of camera_perspective_blender: the near plane maps to
    NDC z = +1 and the far plane to -1, so pixel rays (which start at NDC
    z = +1) leave the camera and their first hit is the visible surface.
    w = -z_camera > 0 in front of the camera."""
    f = 1.0 / math.tan(math.radians(fovy_deg) * 0.5)
    return _mat([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                 [0, 0, -(far + near) / (near - far), -2.0 * far * near / (near - far)],
                 [0, 0, -1, 0]], device, dtype)


def camera_perspective_blender(aspect: float, lens: float, near: float, far: float,
                               proj_direction: bool, *, dtype=torch.float32,
                               device=None) -> torch.Tensor:
    """Blender-like projection (reference Mat44/torch.py:143-206): focal
    length `lens` in mm on an 18 mm half-sensor; with proj_direction the
    camera looks down -z and w = z_camera, the near plane maps to NDC z = +1.
    Without it, the same frustum seen down +z (the matrix for (-far, -near)
    followed by a flip of all three axes)."""
    if not proj_direction:
        t0 = camera_perspective_blender(aspect, lens, -far, -near, True, dtype=dtype,
                                        device=device)
        flip = torch.diag(torch.tensor([-1.0, -1.0, -1.0, 1.0], dtype=dtype, device=device))
        return t0 @ flip
    aspect = torch.as_tensor(aspect, dtype=dtype, device=device)
    lens = torch.as_tensor(lens, dtype=dtype, device=device)
    near = torch.as_tensor(near, dtype=dtype, device=device)
    far = torch.as_tensor(far, dtype=dtype, device=device)
    one = torch.ones((), dtype=dtype, device=device)
    zero = torch.zeros((), dtype=dtype, device=device)
    sx, sy = (one / aspect, one) if aspect < one else (one, aspect)
    half = torch.atan2(torch.tensor(18.0, dtype=dtype, device=device), lens)
    focus = torch.cos(half) / torch.sin(half)
    zz = (near + far) / (near - far)
    zw = 2.0 * far * near / (near - far)
    return torch.stack([
        torch.stack([-sx * focus, zero, zero, zero]),
        torch.stack([zero, -sy * focus, zero, zero]),
        torch.stack([zero, zero, zz, zw]),
        torch.stack([zero, zero, one, zero]),
    ])


def look_at(eye, target, up, device=None, dtype=torch.float32) -> torch.Tensor:
    """World -> camera (right-handed, camera looks down -z)."""
    import numpy as np

    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return torch.tensor(m, device=device, dtype=dtype)

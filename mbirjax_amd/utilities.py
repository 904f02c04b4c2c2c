"""Phantoms for tests and benchmarks (host NumPy).

generate_3d_shepp_logan_low_dynamic_range and gen_cube_phantom produce the same
volumes as mbirjax/utilities.py:658-695 and 1034-1054 (ellipsoid membership
test of add_ellipsoid, mbirjax/utilities.py:856-895).  gen_random_ellipsoid_phantom
is the seeded random-ellipsoid phantom used by the benchmark configurations
(the reference has no seeded generator).
"""
import numpy as np


def _add_ellipsoid(volume, x_grid, y_grid, z_locations, x0, y0, z0, a, b, c, angle=0.0, intensity=1.0):
    ca, sa = np.cos(np.deg2rad(angle)), np.sin(np.deg2rad(angle))
    xr = ca * (x_grid - x0) + sa * (y_grid - y0)
    yr = -sa * (x_grid - x0) + ca * (y_grid - y0)
    xy = xr ** 2 / a ** 2 + yr ** 2 / b ** 2
    inside = (xy[:, :, None] + ((z_locations[None, None, :] - z0) ** 2) / c ** 2) <= 1
    volume += intensity * inside.astype(volume.dtype)
    return volume


def generate_3d_shepp_logan_low_dynamic_range(phantom_shape, device=None):
    """Low-dynamic-range 3D Shepp-Logan phantom with values in [0, 1]."""
    N, M, P = phantom_shape
    vol = np.zeros(phantom_shape, dtype=np.float32)
    xg, yg = np.meshgrid(np.linspace(-1, 1, N, dtype=np.float32), np.linspace(-1, 1, M, dtype=np.float32),
                         indexing='ij')
    z = np.linspace(-1, 1, P, dtype=np.float32)
    for args, kw in [((0, 0, 0, 0.69, 0.92, 0.9), dict(intensity=1)),
                     ((0, 0.0184, 0, 0.6624, 0.874, 0.88), dict(intensity=-0.8)),
                     ((0.22, 0, 0, 0.41, 0.16, 0.21), dict(angle=108, intensity=-0.2)),
                     ((-0.22, 0, 0, 0.31, 0.11, 0.22), dict(angle=72, intensity=-0.2)),
                     ((0, 0.35, 0, 0.21, 0.25, 0.5), dict(intensity=0.1)),
                     ((0, 0.1, 0, 0.046, 0.046, 0.046), dict(intensity=0.1)),
                     ((0, -0.1, 0, 0.046, 0.046, 0.046), dict(intensity=0.1)),
                     ((-0.08, -0.605, 0, 0.046, 0.023, 0.02), dict(angle=0, intensity=0.1)),
                     ((0, -0.605, 0, 0.023, 0.023, 0.02), dict(angle=0, intensity=0.1))]:
        _add_ellipsoid(vol, xg, yg, z, *args, **kw)
    return _place(vol, device)


def gen_cube_phantom(recon_shape, device=None):
    """A sheared cube of value 1/max(size/4) (mbirjax/utilities.py:1034-1054)."""
    rows, cols, slices = recon_shape[:3]
    pr, pc = rows // 4, cols // 4
    vol = np.zeros((rows, cols, slices), dtype=np.float32)
    r0, r1 = (rows - pr) // 2, (rows + pr) // 2
    c0, c1 = (cols - pc) // 2, (cols + pc) // 2
    for k in range(slices):
        sh = int(k * pc / slices)
        vol[r0:r1, (sh + c0):(sh + c1), k] = 1.0 / max(pr, pc)
    return _place(vol, device)


def gen_random_ellipsoid_phantom(recon_shape, num_ellipsoids=10, seed=0, device=None):
    """Seeded phantom: a sum of num_ellipsoids random ellipsoids inside the
    unit ball of normalised coordinates, rescaled to [0, 1]."""
    rng = np.random.default_rng(seed)
    N, M, P = recon_shape
    vol = np.zeros(recon_shape, dtype=np.float32)
    xg, yg = np.meshgrid(np.linspace(-1, 1, N, dtype=np.float32), np.linspace(-1, 1, M, dtype=np.float32),
                         indexing='ij')
    z = np.linspace(-1, 1, P, dtype=np.float32)
    for _ in range(num_ellipsoids):
        x0, y0, z0 = rng.uniform(-0.45, 0.45, size=3)
        a, b, c = rng.uniform(0.08, 0.45, size=3)
        angle = rng.uniform(0, 180)
        intensity = rng.uniform(0.1, 1.0)
        _add_ellipsoid(vol, xg, yg, z, x0, y0, z0, a, b, c, angle=angle, intensity=intensity)
    mx = float(vol.max())
    if mx > 0:
        vol /= mx
    return _place(vol, device)


def _place(vol, device):
    if device is None:
        return vol
    import torch
    return torch.as_tensor(vol, device=device)


def gen_translation_vectors(num_x_translations, num_z_translations, x_spacing, z_spacing):
    """(views, 3) translation vectors on a centred x-z grid, x fastest, y = 0
    (mbirjax/utilities.py:1003-1031)."""
    num_views = num_x_translations * num_z_translations
    tv = np.zeros((num_views, 3))
    x_center = (num_x_translations - 1) / 2
    z_center = (num_z_translations - 1) / 2
    for row in range(num_z_translations):
        for col in range(num_x_translations):
            tv[row * num_x_translations + col] = [(col - x_center) * x_spacing, 0.0, (row - z_center) * z_spacing]
    return tv

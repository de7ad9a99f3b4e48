"""Seeded synthetic registration workloads (there is no network for real data).

Volumes are numpy arrays in C order (z, y, x), i.e. x fastest, matching the
reference's tipl::image layout.  Affine parameter vectors are the reference's
affine_param order: translocation[3], rotation[3] (radians), scaling[3],
affine (shear)[3].

  blob_pair        -- 96x112x96: 20 Gaussian blobs + 1% noise, moving = fixed
                      under a random 12-DOF affine.
  tissue_pair      -- 182x218x182: 4-class nested-ellipsoid phantom
                      (0/300/600/900, sigma 1, 2% noise); moving = monotone
                      cubic intensity remap + random affine.
  warped_pair      -- 160x192x160: the tissue phantom warped by 30 Gaussian
                      radial-basis displacements (max 6 voxels, Jacobian > 0).
Seeds follow 1000*config + index.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import ndimage


def random_affine(rng: np.random.Generator, t=8.0, r_deg=10.0, s=0.05, a=0.03) -> np.ndarray:
    """Random affine_param within the given half-ranges (voxel units, degrees)."""
    p = np.zeros(12, np.float32)
    p[0:3] = rng.uniform(-t, t, 3)
    p[3:6] = np.deg2rad(rng.uniform(-r_deg, r_deg, 3))
    p[6:9] = 1.0 + rng.uniform(-s, s, 3)
    p[9:12] = rng.uniform(-a, a, 3)
    return p


def affine_matrix_np(p, from_shape, to_shape, from_vs=(1, 1, 1), to_vs=(1, 1, 1)) -> np.ndarray:
    """transformation_matrix(affine_param, ...) in float64 (for data generation only;
    the product computes it in the reference's float arithmetic)."""
    p = np.asarray(p, np.float64)
    sx, cx = math.sin(p[3]), math.cos(p[3])
    sy, cy = math.sin(p[4]), math.cos(p[4])
    sz, cz = math.sin(p[5]), math.cos(p[5])
    m = np.array([[cy * cz, cy * sz, -sy],
                  [-cx * sz + sx * cz * sy, cx * cz + sx * sz * sy, sx * cy],
                  [sx * sz + cx * cz * sy, -sx * cz + cx * sz * sy, cx * cy]])
    m = m * p[6:9][:, None]
    m[0] += m[1] * p[9] + m[2] * p[10]
    m[1] += m[2] * p[11]
    fdim = np.array(from_shape[::-1], np.float64)
    tdim = np.array(to_shape[::-1], np.float64)
    t = fdim * -0.5 * np.asarray(from_vs, np.float64) + p[0:3]
    sh = m @ t
    m = m * np.asarray(from_vs, np.float64)[None, :]
    m = m / np.asarray(to_vs, np.float64)[:, None]
    sh = sh / np.asarray(to_vs, np.float64) + tdim * 0.5
    return m, sh


def warp_affine(vol: np.ndarray, p) -> np.ndarray:
    """moving(q) = vol(M q + s): vol resampled on its own grid (float, linear)."""
    m, sh = affine_matrix_np(p, vol.shape, vol.shape)
    zz, yy, xx = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in vol.shape], indexing="ij")
    pts = np.stack([xx, yy, zz], 0).reshape(3, -1)
    src = m @ pts + sh[:, None]
    coords = src[::-1]  # (z, y, x)
    out = ndimage.map_coordinates(vol, coords, order=1, mode="constant", cval=0.0)
    return out.reshape(vol.shape).astype(np.float32)


def blobs(shape=(96, 112, 96), n=20, seed=0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(*[np.arange(s, dtype=np.float32) for s in shape], indexing="ij")
    v = np.zeros(shape, np.float32)
    for _ in range(n):
        c = [rng.uniform(0, s) for s in shape]
        sig = rng.uniform(3, 12)
        amp = rng.uniform(100, 1000)
        r2 = (zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2
        v += amp * np.exp(-r2 / (2 * sig * sig)).astype(np.float32)
    v += rng.normal(0.0, 0.01 * float(v.max()), shape).astype(np.float32)
    return v


def blob_pair(seed=0, shape=(96, 112, 96)):
    """Config 0: (fixed, moving, truth_param); moving = fixed under truth."""
    rng = np.random.default_rng(seed)
    fixed = blobs(shape, 20, seed)
    truth = random_affine(rng)
    moving = warp_affine(fixed, truth)
    return fixed, moving, truth


def tissue(shape=(182, 218, 182), seed=1000) -> np.ndarray:
    """4-class nested ellipsoids (0/300/600/900), sigma 1 smoothing, 2% noise."""
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(*[np.linspace(-1, 1, s, dtype=np.float32) for s in shape], indexing="ij")
    v = np.zeros(shape, np.float32)
    radii = [(0.85, 0.80, 0.75), (0.62, 0.58, 0.55), (0.36, 0.30, 0.33)]
    for cls, (rz, ry, rx) in enumerate(radii, start=1):
        jit = rng.uniform(-0.04, 0.04, 3)
        inside = ((zz - jit[0]) / rz) ** 2 + ((yy - jit[1]) / ry) ** 2 + ((xx - jit[2]) / rx) ** 2 <= 1
        v[inside] = 300.0 * cls
    v = ndimage.gaussian_filter(v, 1.0)
    v += rng.normal(0.0, 0.02 * 900.0, shape).astype(np.float32)
    return v.astype(np.float32)


def monotone_remap(v: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """A random monotone cubic intensity remap of [min, max] onto itself."""
    lo, hi = float(v.min()), float(v.max())
    u = (v - lo) / (hi - lo)
    a, b = rng.uniform(0.2, 1.0), rng.uniform(0.0, 0.8)
    c = rng.uniform(0.0, 0.5)
    f = a * u + b * u ** 2 + c * u ** 3  # derivative > 0 on [0, 1]
    f = f / (a + b + c)
    return (lo + f * (hi - lo)).astype(np.float32)


def tissue_pair(seed=1000, shape=(182, 218, 182)):
    """Config 1: (fixed, moving, truth_param)."""
    rng = np.random.default_rng(seed)
    fixed = tissue(shape, seed)
    truth = random_affine(rng)
    moving = warp_affine(monotone_remap(fixed, rng), truth)
    return fixed, moving, truth


def rbf_field(shape, n=30, max_disp=6.0, seed=2000, step=1) -> np.ndarray:
    """Smooth displacement (z, y, x, 3) in voxels, components (dx, dy, dz),
    sum of n Gaussian radial-basis displacements (sigma 8-20) scaled so that
    the largest displacement is max_disp; raises if the Jacobian is not > 0.
    step > 1 evaluates the field on every step-th voxel and interpolates it
    linearly (a cheaper field for large volumes; sigma >= 8 keeps it smooth)."""
    rng = np.random.default_rng(seed)
    grids = [np.arange(0, s, step, dtype=np.float32) for s in shape]
    zz, yy, xx = np.meshgrid(*grids, indexing="ij")
    cshape = zz.shape
    d = np.zeros(cshape + (3,), np.float32)
    for _ in range(n):
        c = [rng.uniform(0, s) for s in shape]
        sig = rng.uniform(8, 20)
        vec = rng.normal(0, 1, 3).astype(np.float32)
        w = np.exp(-((zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2) / (2 * sig * sig))
        d += w[..., None] * vec[None, None, None, :]
    mag = np.sqrt((d ** 2).sum(-1)).max()
    d *= max_disp / mag
    # Jacobian determinant of x -> x + d(x), FP64, in voxels of the full grid
    g = [np.gradient(d[..., k].astype(np.float64), float(step)) for k in range(3)]
    J = np.empty(cshape + (3, 3))
    for k in range(3):          # component (x, y, z)
        for j in range(3):      # derivative wrt (x, y, z)
            J[..., k, j] = g[k][2 - j] + (1.0 if k == j else 0.0)
    det = np.linalg.det(J)
    if det.min() <= 0:
        raise ValueError("generated field is not diffeomorphic")
    if step == 1:
        return d
    # linear interpolation onto the full grid (positions i / step of the coarse grid)
    coords = np.meshgrid(*[np.minimum(np.arange(s) / step, cs - 1) for s, cs in zip(shape, cshape)],
                         indexing="ij")
    return np.stack([ndimage.map_coordinates(d[..., k], coords, order=1, mode="nearest")
                     for k in range(3)], -1).astype(np.float32)


def warp_field(vol: np.ndarray, d: np.ndarray) -> np.ndarray:
    """moving(x) = vol(x + d(x)) with linear interpolation."""
    zz, yy, xx = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in vol.shape], indexing="ij")
    coords = np.stack([zz + d[..., 2], yy + d[..., 1], xx + d[..., 0]], 0)
    return ndimage.map_coordinates(vol, coords, order=1, mode="nearest").astype(np.float32)


def warped_pair(seed=2000, shape=(160, 192, 160)):
    """Config 2: (fixed, moving, field)."""
    fixed = tissue(shape, seed)
    d = rbf_field(shape, 30, 6.0, seed)
    moving = warp_field(fixed, d)
    return fixed, moving, d


def serving_pair(seed, shape, nonlinear=False, dtype=np.float32):
    """Config 4 request: (fixed, moving, truth_param) of `shape` and `dtype`
    (float32 or int16): the tissue phantom, and as moving its monotone remap,
    for affine+CDM requests warped by an RBF field (max 4 voxels), then moved
    by a random affine."""
    rng = np.random.default_rng(seed)
    fixed = tissue(shape, seed)
    moving = monotone_remap(fixed, rng)
    if nonlinear:
        moving = warp_field(moving, rbf_field(shape, 30, 4.0, seed + 1, step=4))
    truth = random_affine(rng)
    moving = warp_affine(moving, truth)
    if np.dtype(dtype) == np.int16:
        fixed = np.clip(np.rint(fixed), -32768, 32767).astype(np.int16)
        moving = np.clip(np.rint(moving), -32768, 32767).astype(np.int16)
    return fixed.astype(dtype), moving.astype(dtype), truth


def quantize(v: np.ndarray, upper: float = 255.99) -> np.ndarray:
    """normalize_upper_lower2 in float32 (the product does this on the device)."""
    v = v.astype(np.float32)
    mn, mx = np.float32(v.min()), np.float32(v.max())
    rng = np.float32(mx - mn)
    if rng == 0:
        return np.zeros(v.shape, np.uint8)
    coef = np.float32(np.float32(upper) / rng)
    return ((v - mn) * coef).astype(np.uint8)


def candidate_params(rng: np.random.Generator, K: int, around=None, t=8.0, r_deg=10.0, s=0.05,
                     a=0.03) -> np.ndarray:
    """K affine_param candidates (a line-search-like batch) around `around`."""
    base = np.array(around if around is not None else [0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0],
                    np.float32)
    out = np.repeat(base[None, :], K, 0)
    for k in range(K):
        out[k] += random_affine(rng, t, r_deg, s, a) - np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0],
                                                                np.float32)
    return out.astype(np.float32)

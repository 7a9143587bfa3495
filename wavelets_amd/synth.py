"""Seeded synthetic cubelets for the benchmark workloads (numpy PCG64, so the
same seed gives the same bytes on every machine).  Arrays are indexed
[z, y, x] (x fastest), the layout of a cubelet's raw data.

  smooth_field_u8   2048 x 2048 x 1 uint8: 64 Gaussian blobs (sigma 20-300
                    px) plus fractal value noise, rescaled to 0..255
  ellipsoids_f32    512^3 float32: 40 rotated ellipsoids, intensities
                    U[0,4000] painted in order, plus N(0, 20^2) noise
  lattice_i16       256^3 int16 values (stored as int32 cubelets): trilinear
                    interpolation of a 17^3 U[500,3500] lattice plus
                    Laplace(b=30) noise, clipped to 0..4095
  mixed_stream      the sizes / dtypes / wavelets of a mixed cubelet stream
  serving_requests  the shapes / dtypes / kept fractions of serving requests
"""
import numpy as np


def smooth_field_u8(nx=2048, ny=2048, seed=1, blobs=64):
    rng = np.random.default_rng(seed)
    y = np.arange(ny, dtype=np.float32)[:, None]
    x = np.arange(nx, dtype=np.float32)[None, :]
    img = np.zeros((ny, nx), np.float32)
    for _ in range(blobs):
        cx, cy = rng.uniform(0, nx), rng.uniform(0, ny)
        s = rng.uniform(20, 300)
        a = rng.uniform(-1, 1)
        gx = np.exp(-((x - cx) ** 2) / (2 * s * s)).astype(np.float32)
        gy = np.exp(-((y - cy) ** 2) / (2 * s * s)).astype(np.float32)
        img += np.float32(a) * (gy * gx)
    # fractal value noise: octaves of bilinearly interpolated lattices with
    # amplitude proportional to the cell size (a 1/f spectrum)
    noise = np.zeros_like(img)
    cell = 512
    while cell >= 2:
        ly, lx = ny // cell + 2, nx // cell + 2
        lat = rng.standard_normal((ly, lx)).astype(np.float32)
        fy = (np.arange(ny, dtype=np.float32) / cell)
        fx = (np.arange(nx, dtype=np.float32) / cell)
        iy, ix = fy.astype(np.int64), fx.astype(np.int64)
        ty, tx = (fy - iy)[:, None], (fx - ix)[None, :]
        v00 = lat[iy][:, ix]
        v01 = lat[iy][:, ix + 1]
        v10 = lat[iy + 1][:, ix]
        v11 = lat[iy + 1][:, ix + 1]
        noise += (cell / 512.0) * ((1 - ty) * ((1 - tx) * v00 + tx * v01) +
                                   ty * ((1 - tx) * v10 + tx * v11))
        cell //= 2
    rng_img = img.max() - img.min()
    noise *= 0.1 * rng_img / max(1e-6, float(np.abs(noise).max()))
    img += noise
    img = (img - img.min()) / max(1e-12, float(img.max() - img.min())) * 255.0
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)[None, :, :]


def _rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.diag(r))


def ellipsoids_f32(n=512, seed=2, count=40, noise_sigma=20.0):
    rng = np.random.default_rng(seed)
    vol = np.zeros((n, n, n), np.float32)
    for _ in range(count):
        c = rng.uniform(0, n, 3)
        ax = rng.uniform(16, 128, 3)
        R = _rotation(rng)
        val = np.float32(rng.uniform(0, 4000))
        # half extent along world axis i of p = c + R diag(ax) u, |u| <= 1;
        # grid points a unit or more beyond it are well outside
        r = np.sqrt(((R * ax[None, :]) ** 2).sum(axis=1))
        lo = np.maximum(np.floor(c - r), 0).astype(int)
        hi = np.minimum(np.ceil(c + r) + 1, n).astype(int)
        if (hi <= lo).any():
            continue
        zz = np.arange(lo[2], hi[2], dtype=np.float32) - c[2]
        yy = np.arange(lo[1], hi[1], dtype=np.float32) - c[1]
        xx = np.arange(lo[0], hi[0], dtype=np.float32) - c[0]
        # body coordinates u = R^T (p - c) / axes, summed as squares
        M = (R.T / ax[:, None]).astype(np.float32)       # rows: body axes
        acc = None
        for k in range(3):
            u = (M[k, 2] * zz[:, None, None] + M[k, 1] * yy[None, :, None] +
                 M[k, 0] * xx[None, None, :])
            acc = u * u if acc is None else acc + u * u
        sub = vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub[acc <= 1.0] = val
    vol += rng.standard_normal(vol.shape, dtype=np.float32) * np.float32(noise_sigma)
    return vol


def lattice_i16(n=256, seed=3, cells=16, b=30.0):
    rng = np.random.default_rng(seed)
    lat = rng.uniform(500, 3500, (cells + 1,) * 3).astype(np.float32)
    f = np.arange(n, dtype=np.float32) * (cells / n)
    i = np.minimum(f.astype(np.int64), cells - 1)
    t = (f - i).astype(np.float32)
    # separable trilinear interpolation along x, then y, then z
    a = lat[:, :, i] * (1 - t) + lat[:, :, i + 1] * t
    a = a[:, i, :] * (1 - t)[None, :, None] + a[:, i + 1, :] * t[None, :, None]
    a = a[i, :, :] * (1 - t)[:, None, None] + a[i + 1, :, :] * t[:, None, None]
    a += rng.laplace(0, b, a.shape).astype(np.float32)
    return np.clip(np.rint(a), 0, 4095).astype(np.int16)


MIXED_SIZES = [(32, 32, 32), (64, 64, 64), (128, 128, 128), (256, 256, 64),
               (512, 512, 16)]       # (x, y, z)


def mixed_stream(count=2000, seed=4):
    """List of (size_xyz, dtype, wavelet) for the mixed-size stream."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        size = MIXED_SIZES[int(rng.integers(len(MIXED_SIZES)))]
        dtype = ("uint8", "int16", "float32")[int(rng.integers(3))]
        wavelet = 2 if rng.random() < 0.5 else 1
        out.append((size, dtype, wavelet))
    return out


def serving_requests(count, seed=5):
    """The serving workload's requests: (size_xyz, dtype, keep) with edges
    lognormal (median 128, sigma 0.6) clipped to [32, 512], depth lognormal
    (median 48, sigma 1.0) clipped to [1, 512], dtypes u8 / i16 / f32 and
    the kept fraction uniform on [0.02, 0.2] (threshold = 1 - keep)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        ex = int(np.clip(np.round(np.exp(rng.normal(np.log(128), 0.6))), 32, 512))
        ey = int(np.clip(np.round(np.exp(rng.normal(np.log(128), 0.6))), 32, 512))
        ez = int(np.clip(np.round(np.exp(rng.normal(np.log(48), 1.0))), 1, 512))
        dt = ("uint8", "int16", "float32")[int(rng.integers(3))]
        keep = float(rng.uniform(0.02, 0.2))
        out.append(((ex, ey, ez), dt, keep))
    return out


def cubelet(size_xyz, dtype, seed):
    """Contents for one stream cubelet: a smooth field (blobs) with noise,
    scaled to the dtype's range."""
    rng = np.random.default_rng(seed)
    x, y, z = size_xyz
    g = [np.arange(n, dtype=np.float32) for n in (z, y, x)]
    v = np.zeros((z, y, x), np.float32)
    for _ in range(4):
        c = [rng.uniform(0, n) for n in (z, y, x)]
        s = rng.uniform(4, max(8, max(x, y, z) / 3))
        w = [np.exp(-((g[k] - c[k]) ** 2) / (2 * s * s)) for k in range(3)]
        v += np.float32(rng.uniform(0.2, 1.0)) * (w[0][:, None, None] * w[1][None, :, None] *
                                                   w[2][None, None, :])
    v += rng.standard_normal(v.shape, dtype=np.float32) * np.float32(0.03)
    if dtype == "uint8":
        return np.clip(np.rint(v * 200 + 20), 0, 255).astype(np.uint8)
    if dtype == "int16":
        return np.clip(np.rint(v * 3000 + 300), 0, 4095).astype(np.int16)
    return (v * 4000).astype(np.float32)


def cubelet_device(size_xyz, dtype, seed, device="cuda"):
    """cubelet() generated on the device: the blob parameters come from the
    same numpy generator, the noise from torch's generator seeded with seed
    (deterministic on one platform).  Used for the mixed stream, whose 2000
    cubelets hold about 4.3 G samples."""
    import torch
    rng = np.random.default_rng(seed)
    x, y, z = size_xyz
    g = [torch.arange(n, dtype=torch.float32, device=device) for n in (z, y, x)]
    v = torch.zeros((z, y, x), dtype=torch.float32, device=device)
    for _ in range(4):
        c = [rng.uniform(0, n) for n in (z, y, x)]
        s = rng.uniform(4, max(8, max(x, y, z) / 3))
        w = [torch.exp(-((g[k] - float(c[k])) ** 2) / float(2 * s * s)) for k in range(3)]
        v += float(rng.uniform(0.2, 1.0)) * (w[0][:, None, None] * w[1][None, :, None] *
                                             w[2][None, None, :])
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    v += torch.randn(v.shape, generator=gen, device=device) * 0.03
    if dtype == "uint8":
        return torch.clamp(torch.round(v * 200 + 20), 0, 255).to(torch.uint8)
    if dtype == "int16":
        return torch.clamp(torch.round(v * 3000 + 300), 0, 4095).to(torch.int16)
    return v * 4000


def stream_steps(size_xyz):
    """min(3, ceilLog2(dim)) steps per axis."""
    out = []
    for d in size_xyz:
        m = 0
        while (1 << m) < d:
            m += 1
        out.append(min(3, m))
    return out

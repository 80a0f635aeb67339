"""Deterministic synthetic fields for the benchmark workloads and tests.

Every generator draws its parameters from ``numpy.random.default_rng(seed)``
and evaluates the field in binary64 with torch (on the GPU when one is given,
else on the CPU), rounding to the requested dtype at the end.  Fields are
HDF5-ordered arrays f[X][Y][Z] (Z fastest), the order CubismZ's hdf2cz reads
(Tools/Test_IO_Compressed.h:184-212).

* ``bumps_field``   -- 1.0 + 16 Gaussian bumps (centres U[0,1]^3, widths
  0.02-0.1, amplitudes U[-0.5,0.5]) + one tanh plane interface (2 cells
  thick, jump 0.3) + uniform noise in [-1e-4, 1e-4].
* ``bubble_field``  -- cavitation-like pressure: background 1.0 plus spherical
  bubbles (radius U[4,40] cells, centres uniform), tanh walls 2 cells thick,
  interior pressure log-uniform in [0.01, 100], and outside each bubble a
  1/r halo (p_b - 1) R/r.  ``screened=True`` multiplies the halo by
  exp(-(r-R)/R), so that each bubble's influence stays local.
"""
import math

import numpy as np


def _torch():
    import torch

    return torch


def _grid(torch, shape, device, x0, nx):
    """Cell-centre coordinates of the X-slab [x0, x0+nx) as (X, Y, Z) tensors."""
    NX, NY, NZ = shape
    X = torch.arange(x0, x0 + nx, dtype=torch.float64, device=device).view(nx, 1, 1) + 0.5
    Y = torch.arange(NY, dtype=torch.float64, device=device).view(1, NY, 1) + 0.5
    Z = torch.arange(NZ, dtype=torch.float64, device=device).view(1, 1, NZ) + 0.5
    return X, Y, Z


def _slabs(shape, budget=1 << 24):
    NX, NY, NZ = shape
    step = max(1, budget // (NY * NZ))
    for x0 in range(0, NX, step):
        yield x0, min(step, NX - x0)


def bumps_field(shape=(128, 128, 128), seed=0, dtype=None, device="cpu"):
    torch = _torch()
    dtype = dtype or torch.float32
    rng = np.random.default_rng(seed)
    NX, NY, NZ = shape
    L = float(max(shape))
    cen = rng.uniform(0.0, 1.0, (16, 3)) * L
    wid = rng.uniform(0.02, 0.1, 16) * L
    amp = rng.uniform(-0.5, 0.5, 16)
    nrm = rng.normal(size=3)
    nrm /= np.linalg.norm(nrm)
    off = float(np.dot(nrm, np.array(shape, float) / 2) + rng.uniform(-0.1, 0.1) * L)
    noise = torch.from_numpy(rng.uniform(-1e-4, 1e-4, shape)).to(device)
    out = torch.empty(shape, dtype=dtype, device=device)
    for x0, nx in _slabs(shape):
        X, Y, Z = _grid(torch, shape, device, x0, nx)
        f = torch.ones((nx, NY, NZ), dtype=torch.float64, device=device)
        for k in range(16):
            r2 = (X - cen[k, 0]) ** 2 + (Y - cen[k, 1]) ** 2 + (Z - cen[k, 2]) ** 2
            f += amp[k] * torch.exp(-r2 / (2.0 * wid[k] ** 2))
        d = nrm[0] * X + nrm[1] * Y + nrm[2] * Z - off
        f += 0.15 * (1.0 + torch.tanh(d / 2.0))
        f += noise[x0:x0 + nx]
        out[x0:x0 + nx] = f.to(dtype)
    return out


def bubble_params(shape, seed, nbubbles):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0.0, 1.0, (nbubbles, 3)) * np.array(shape, float)
    rad = rng.uniform(4.0, 40.0, nbubbles)
    pin = np.exp(rng.uniform(math.log(0.01), math.log(100.0), nbubbles))
    return cen, rad, pin


def bubble_field(shape=(512, 512, 512), seed=1, nbubbles=200, dtype=None, device="cpu", screened=False):
    torch = _torch()
    dtype = dtype or torch.float32
    NX, NY, NZ = shape
    cen, rad, pin = bubble_params(shape, seed, nbubbles)
    out = torch.empty(shape, dtype=dtype, device=device)
    for x0, nx in _slabs(shape):
        X, Y, Z = _grid(torch, shape, device, x0, nx)
        f = torch.ones((nx, NY, NZ), dtype=torch.float64, device=device)
        for k in range(nbubbles):
            R, dp = float(rad[k]), float(pin[k] - 1.0)
            r = torch.sqrt((X - cen[k, 0]) ** 2 + (Y - cen[k, 1]) ** 2 + (Z - cen[k, 2]) ** 2)
            w = 0.5 * (1.0 - torch.tanh((r - R) / 2.0))
            halo = R / torch.clamp(r, min=R)
            if screened:
                halo = halo * torch.exp(-torch.clamp(r - R, min=0.0) / R)
            f += dp * (w + (1.0 - w) * halo)
        out[x0:x0 + nx] = f.to(dtype)
    return out


def bubbles_for(shape, per_512cube=200):
    """Bubble count of a field at the 512^3 configuration's density (200 per
    512^3 cells), at least one."""
    return max(1, int(round(per_512cube * float(np.prod(shape)) / 512.0 ** 3)))


def batched_shapes(nfields=64, seed=7):
    """Edges drawn from {64, 96, ..., 384} per axis, one shape per field."""
    rng = np.random.default_rng(seed)
    edges = np.arange(64, 385, 32)
    shapes = [tuple(int(e) for e in rng.choice(edges, 3)) for _ in range(nfields)]
    eps = np.exp(rng.uniform(math.log(1e-4), math.log(1e-2), nfields)).astype(np.float32)
    return shapes, eps

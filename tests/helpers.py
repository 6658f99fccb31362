"""Shared helpers of the test suite."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def smooth_flow(dims, amp, seed):
    """Smooth 3-plane flow field (voxel units) from a few random sinusoids."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((3, nz, ny, nx), np.float64)
    for c in range(3):
        for _ in range(3):
            k = rng.integers(0, 3, 3)
            ph = rng.uniform(0, 2 * np.pi)
            out[c] += rng.uniform(-1, 1) * np.sin(2 * np.pi * (k[0] * x / nx + k[1] * y / ny + k[2] * z / nz) + ph)
    out *= amp / max(np.abs(out).max(), 1e-9)
    return out.astype(np.float32).reshape(-1)

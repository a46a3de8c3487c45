"""Seeded synthetic scene-linear frames for parity tests and benchmarks.

Frame model (shared by every workload):
  * luminance ramp L(x) = 2^(-6 + 8 x / (W-1)) across the width;
  * a smooth hue field: hue in [0,1) and saturation in [0,0.8] drawn on an
    8x8 seeded grid and upsampled bicubically (hue wraps circularly), turned
    into a Rec.709-linear chroma multiplier normalized to unit luminance
    (Y = 0.2126 R + 0.7152 G + 0.0722 B);
  * 64 neutral flat 32x32 patches at seeded positions with log-uniform
    luminance in [2^-6, 4];
  * optionally a fraction of pixels set to a specular value (RGB = 64).
Alpha is 1. Rows are contiguous RGBA float32.
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 20261203
_Y709 = np.array([0.2126, 0.7152, 0.0722])


def _cubic_upsample(grid: np.ndarray, size: int) -> np.ndarray:
    """Catmull-Rom upsampling of a 1D control grid (edge-clamped) to `size`."""
    n = grid.shape[0]
    t = np.linspace(0.0, n - 1.0, size)
    i = np.clip(np.floor(t).astype(np.int64), 0, n - 2)
    f = t - i
    idx = [np.clip(i + k, 0, n - 1) for k in (-1, 0, 1, 2)]
    p0, p1, p2, p3 = (grid[j] for j in idx)
    f = f.reshape((-1,) + (1,) * (grid.ndim - 1))
    return 0.5 * ((2 * p1) + (-p0 + p2) * f + (2 * p0 - 5 * p1 + 4 * p2 - p3) * f * f
                  + (-p0 + 3 * p1 - 3 * p2 + p3) * f * f * f)


def _field(rng: np.random.Generator, w: int, h: int, wrap: bool) -> np.ndarray:
    grid = rng.random((8, 8))
    if wrap:
        ang = grid * 2 * np.pi
        parts = [np.cos(ang), np.sin(ang)]
        up = []
        for part in parts:
            rows = _cubic_upsample(part, h)            # (h, 8)
            up.append(_cubic_upsample(rows.T, w).T)    # (h, w)
        return (np.arctan2(up[1], up[0]) / (2 * np.pi)) % 1.0
    rows = _cubic_upsample(grid, h)
    return np.clip(_cubic_upsample(rows.T, w).T, 0.0, 1.0)


def _hsv_to_rgb(h, s):
    i = np.floor(h * 6.0).astype(np.int64) % 6
    f = h * 6.0 - np.floor(h * 6.0)
    v = np.ones_like(h)
    p, q, t = v * (1 - s), v * (1 - s * f), v * (1 - s * (1 - f))
    r = np.choose(i, [v, q, p, p, t, v])
    g = np.choose(i, [t, v, v, q, p, p])
    b = np.choose(i, [p, p, t, v, v, q])
    return np.stack([r, g, b], axis=-1)


def frame(width: int, height: int, seed: int = DEFAULT_SEED, specular_fraction: float = 0.0,
          components: int = 4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    hue = _field(rng, width, height, wrap=True)
    sat = 0.8 * _field(rng, width, height, wrap=False)
    chroma = _hsv_to_rgb(hue, sat)
    chroma /= (chroma @ _Y709)[..., None]
    x = np.arange(width, dtype=np.float64)
    lum = np.exp2(-6.0 + 8.0 * x / max(width - 1, 1))[None, :]
    rgb = chroma * lum[..., None]
    n_patch = 64
    size = 32
    if width >= size and height >= size:
        px = rng.integers(0, width - size + 1, n_patch)
        py = rng.integers(0, height - size + 1, n_patch)
        lv = np.exp2(rng.uniform(-6.0, 2.0, n_patch))
        for k in range(n_patch):
            rgb[py[k]:py[k] + size, px[k]:px[k] + size, :] = lv[k]
    if specular_fraction > 0.0:
        count = int(round(specular_fraction * width * height))
        flat = rng.choice(width * height, size=count, replace=False)
        rgb.reshape(-1, 3)[flat] = 64.0
    out = np.ones((height, width, components), dtype=np.float32)
    out[..., :3] = rgb.astype(np.float32)
    return out

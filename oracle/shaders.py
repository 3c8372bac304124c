"""CPU restatement of the user stage kernels of the reference's extension
tests (``lib/zosimos/tests/custom.rs``): the fragment shaders
``lib/std/src/mandelbrot.frag``, ``crt.frag`` and ``flat_field.frag`` in IEEE
f32, each operation rounded on its own in the shader's order.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker of the
package's own HIP forms of these shaders; nothing here imports the product.
Pinned by the reference's blessed hashes of the three tests
(``tests/golden/reference_hashes.json``: mandelbrot, crt, flat_field) in
``tests/test_oracle.py``.
"""
from __future__ import annotations

import numpy as np

f = np.float32


def uv_grid(w: int, h: int):
    """Texture coordinates of the texel centres, ((x + 0.5) / w, (y + 0.5) / h)."""
    u = ((np.arange(w, dtype=f) + f(0.5)) / f(w))[None, :]
    v = ((np.arange(h, dtype=f) + f(0.5)) / f(h))[:, None]
    return np.broadcast_to(u, (h, w)).astype(f), np.broadcast_to(v, (h, w)).astype(f)


def _fma(a, b, c):
    """a * b + c with the product exact (f32 x f32 fits a double) and the sum
    rounded in double, then to f32 (one rounding but for double-rounding
    ties)."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f)


def mandelbrot(w: int, h: int, scale=(3.0, 3.0), position=(0.6, 0.5), steps: int = 2048):
    """``mandelbrot.frag:17-45``: c = (uv - position) * scale; ``steps``
    iterations of sum += z; z = (z.x^2 - z.y^2 + c.x, fma(2 z.x, z.y, c.y));
    then sum = sum / steps - c and the colour (clamp(2 - |z|, 0, 0.7),
    sum.x / 2, sum.y / 2, 1): Oklab (L, a, b) with alpha.  Returns (h, w, 4)
    float32."""
    u, v = uv_grid(w, h)
    cx = (u - f(position[0])) * f(scale[0])
    cy = (v - f(position[1])) * f(scale[1])
    zx = np.zeros((h, w), f)
    zy = np.zeros((h, w), f)
    sx = np.zeros((h, w), f)
    sy = np.zeros((h, w), f)
    with np.errstate(over="ignore", invalid="ignore"):
        for _ in range(steps):
            sx = sx + zx
            sy = sy + zy
            re = zx * zx - zy * zy
            im = _fma(f(2.0) * zx, zy, cy)
            zx = re + cx
            zy = im
        sx = sx / f(steps) - cx
        sy = sy / f(steps) - cy
        ln = np.sqrt(zx * zx + zy * zy)
        light = np.minimum(np.maximum(f(2.0) - ln, f(0.0)), f(0.7))
    # GLSL clamp of NaN: the hardware's max/min return the non-NaN operand.
    light = np.where(np.isnan(ln), f(0.0), light).astype(f)
    return np.stack([light, sx * f(0.5), sy * f(0.5), np.ones_like(light)], -1).astype(f)


def nearest(img, u, v):
    """A nearest, clamp-to-edge sample (``encoder.rs:1533-1537``) of (H, W, 4)
    ``img`` at texture coordinates (u, v)."""
    h, w = img.shape[:2]
    tx = np.clip(np.floor(u * f(w)), 0, w - 1).astype(np.int64)
    ty = np.clip(np.floor(v * f(h)), 0, h - 1).astype(np.int64)
    return img[ty, tx]


def crt(src, w: int, h: int, scale=None):
    """``crt.frag:17-40`` over a (w, h) result: screen = uint(uv * scale)
    (scale = the result's size), the operand sampled at uv, one channel
    (screen.x mod 3) kept and scaled by 0.16 cell (5 - cell) with
    cell = (screen.y + 3 ((screen.x / 3) mod 2)) mod 6, alpha kept.  ``src``
    is the operand's linear (H, W, 4) float32."""
    sw, sh = scale if scale is not None else (w, h)
    u, v = uv_grid(w, h)
    sx = (u * f(sw)).astype(np.uint32)
    sy = (v * f(sh)).astype(np.uint32)
    c = nearest(np.asarray(src, f), u, v)
    bias = ((sx // 3) % 2) * 3
    cell = (sy + bias) % 6
    k = f(0.16) * (cell.astype(f) * (5 - cell).astype(f))
    out = np.zeros((h, w, 4), f)
    ch = sx % 3
    for j in range(3):
        out[..., j] = np.where(ch == j, c[..., j] * k, f(0.0))
    out[..., 3] = c[..., 3]
    return out


def flat_field(src, field, mean: float, w: int = None, h: int = None):
    """``flat_field.frag:19-26``: the colour of ``src`` scaled by
    mean / field.x, both sampled at the texel's uv; alpha kept."""
    src = np.asarray(src, f)
    h = src.shape[0] if h is None else h
    w = src.shape[1] if w is None else w
    u, v = uv_grid(w, h)
    c = nearest(src, u, v)
    fx = nearest(np.asarray(field, f), u, v)[..., 0]
    with np.errstate(divide="ignore", invalid="ignore"):
        eff = f(mean) / fx
    out = c.copy()
    for j in range(3):
        out[..., j] = c[..., j] * eff
    return out

"""Blockhash perceptual hash -- TEST INFRASTRUCTURE ONLY (see
``oracle/__init__.py``).

The reference's integration tests compare every output image against
blessed perceptual hashes (``tests/util.rs:14-60``: ``blockhash::blockhash256``
of the output, exact string match against any line of
``tests/reference/<name>.crc.png``).  The ``blockhash`` crate (1.0.0) is an
un-vendored dependency; this restates the published block-mean-value
algorithm it implements (Yang, Gu and Niu 2006, as specified by the
blockhash.io reference implementation):

* the image is cut into bits x bits blocks; a pixel's value is r + g + b, or
  765 (white) when its alpha is 0; each block sums its pixels (the "quick"
  method when the sides divide evenly, otherwise each pixel is spread over the
  blocks it overlaps in proportion to the overlap -- the "precise" method,
  which also covers images narrower than the grid);
* the blocks are taken in four horizontal bands; in each band a block's bit
  is 1 when its sum exceeds the band's median (or equals it within 1 while
  the median is above half of the block's maximum);
* bits are written most significant first, four to a hex digit.

Pinned by ``test_oracle.py`` against the reference's own fixtures: the hash
of the unmodified background image equals the blessed hash of the reference's
red/green-swapped background (a channel swap leaves every r + g + b sum
unchanged).
"""
from __future__ import annotations

import numpy as np


def _median(v):
    s = np.sort(np.asarray(v, dtype=np.float64))
    n = len(s)
    return (s[n // 2 - 1] + s[n // 2]) / 2.0 if n % 2 == 0 else s[n // 2]


def _bits(blocks, pixels_per_block):
    blocks = np.asarray(blocks, dtype=np.float64)
    n = len(blocks)
    half = pixels_per_block * 256 * 3 / 2
    band = n // 4
    out = np.zeros(n, dtype=np.uint8)
    for i in range(4):
        seg = blocks[i * band:(i + 1) * band]
        m = _median(seg)
        out[i * band:(i + 1) * band] = (seg > m) | ((np.abs(seg - m) < 1) & (m > half))
    return out


def _hex(bits):
    return "".join("%x" % (int(bits[i]) << 3 | int(bits[i + 1]) << 2 | int(bits[i + 2]) << 1 | int(bits[i + 3]))
                   for i in range(0, len(bits), 4))


def pixel_values(rgba_u8):
    a = np.asarray(rgba_u8, dtype=np.uint8)
    v = a[..., 0].astype(np.float64) + a[..., 1] + a[..., 2]
    return np.where(a[..., 3] == 0, 765.0, v)


def blockhash(rgba_u8, bits: int = 16) -> str:
    """Hex blockhash of an (H, W, 4) RGBA8 image with bits x bits blocks."""
    v = pixel_values(rgba_u8)
    h, w = v.shape
    if w % bits == 0 and h % bits == 0:
        bw, bh = w // bits, h // bits
        blocks = v.reshape(bits, bh, bits, bw).sum(axis=(1, 3)).reshape(-1)
        return _hex(_bits(blocks, bw * bh))
    # precise: a pixel straddling a block boundary is split between the two
    # blocks in proportion to its overlap; the weights separate by axis, so
    # blocks = Ay^T . V . Ax with Ay[y, b] the share of row y in block row b.
    ay = _axis_weights(h, bits)
    ax = _axis_weights(w, bits)
    blocks = ay.T @ v @ ax
    return _hex(_bits(blocks.reshape(-1), (w / bits) * (h / bits)))


def _axis_weights(n: int, bits: int):
    """Share of pixel i (the interval [i, i + 1)) in block b (the interval
    [b n / bits, (b + 1) n / bits)): the length of their overlap.  A pixel
    that straddles one block boundary splits in proportion, as the
    published algorithm does; a pixel wider than a block (n < bits) covers
    several blocks, each with the part it holds."""
    edges = np.arange(bits + 1, dtype=np.float64) * (n / bits)
    lo = np.arange(n, dtype=np.float64)[:, None]
    return np.clip(np.minimum(lo + 1.0, edges[None, 1:]) - np.maximum(lo, edges[None, :-1]), 0.0, None)


def blockhash256(rgba_u8) -> str:
    return blockhash(rgba_u8, 16)


def lumaa16_to_rgba8(raw_bytes, width: int, height: int):
    """An (H, W, 4) RGBA8 view of LumaA16 texels (little-endian u16 luma,
    u16 alpha), as the reference's test hashes a LumaA16 image
    (``blend.rs:340-375``): grey replicated to r, g, b, each 16-bit channel
    rounded to 8 bits, round(v * 255 / 65535).  The ``image`` crate (0.25) that
    converts it is un-vendored; the rule is pinned by the reference's blessed
    ``transmute`` hash (``test_oracle.py``)."""
    v = np.frombuffer(np.ascontiguousarray(raw_bytes).tobytes(), dtype="<u2").reshape(height, width, 2)
    c = ((v.astype(np.uint32) * 255 + 32767) // 65535).astype(np.uint8)
    return np.stack([c[..., 0], c[..., 0], c[..., 0], c[..., 1]], -1)

"""The five compositing workloads of the benchmark configurations, as
CommandBuffer programs plus their seeded synthetic inputs.

C1  two 1024x768 RGBA8 sRGB images: premultiply, 'over' B onto A, affine
    bilinear resample (rotate 15 deg about the centre, scale 0.75,
    transparent border), un-premultiply, encode RGBA8 sRGB.
C2  7680x4320 RGBA8 sRGB -> linear -> Rec.709-to-Display-P3 -> Oklab ->
    chroma x1.2 -> back to linear P3 -> 'over' an RGBA16F layer -> RGBA8.
C3  256 images, sides log-uniform in 64-4096, half RGBA8 sRGB and half
    RGBA16F linear: decode -> 3x3 colour matrix -> bilinear resample by a
    uniform 0.5-2x scale -> encode (one tile-queue launch).
C4  16 x 3840x2160 RGBA8: sRGB decode, linear->XYZ, XYZ->Oklab, hue 30 deg,
    chroma x1.2, back, sRGB encode.
C5  8 layers of 3840x2160 RGBA16F, each with its own affine (rotation
    +-45 deg, scale 0.5-1.5) resampled bilinearly and composited 'over' in
    order, then the C4 colour chain and RGBA8 sRGB encode.

Inputs come from ``numpy.random.default_rng(seed)`` (PCG64), base seed 0,
one stream per image or layer.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import (Affine, AffineSample, Color, ColorModel, CommandBuffer, Descriptor, Primaries, Rectangle,
               SampleBits, SampleParts, Texel, Transfer, Whitepoint)

P3_LINEAR = Color(primaries=Primaries.DisplayP3, transfer=Transfer.Linear)
P3_SRGB = Color(primaries=Primaries.DisplayP3, transfer=Transfer.Srgb)
LINEAR_709 = Color(primaries=Primaries.Bt709, transfer=Transfer.Linear)
RGBA8 = Texel(SampleBits.UInt8x4, SampleParts.RgbA)
RGBA16F = Texel(SampleBits.Float16x4, SampleParts.RgbA)
LABA16F = Texel(SampleBits.Float16x4, SampleParts.LabA)


@dataclass
class Workload:
    name: str
    commands: CommandBuffer
    inputs: Dict[int, Descriptor]               # register -> descriptor
    outputs: Dict[int, Descriptor]              # output register -> descriptor
    make_inputs: Callable[[], Dict[int, np.ndarray]]  # register -> host bytes
    pixels: int                                 # output pixels per launch
    meta: dict = field(default_factory=dict)


def random_rgba8(rng: np.random.Generator, w: int, h: int) -> np.ndarray:
    return rng.integers(0, 256, size=(h, w * 4), dtype=np.uint8)


def random_rgba16f(rng: np.random.Generator, w: int, h: int, alpha: Optional[np.ndarray] = None) -> np.ndarray:
    v = rng.random((h, w, 4), dtype=np.float32)
    if alpha is not None:
        v[..., 3] = alpha
    return v.astype(np.float16).view(np.uint8).reshape(h, w * 8)


def smooth_noise(rng: np.random.Generator, w: int, h: int, cells: int = 8) -> np.ndarray:
    """Alpha in [0, 1] from bilinearly interpolated random values on a coarse
    (cells+1)^2 lattice (seeded smooth noise)."""
    g = rng.random((cells + 1, cells + 1))
    ys = np.linspace(0, cells, h, endpoint=False) + 0.5 * cells / h
    xs = np.linspace(0, cells, w, endpoint=False) + 0.5 * cells / w
    y0 = np.minimum(np.floor(ys).astype(int), cells - 1)
    x0 = np.minimum(np.floor(xs).astype(int), cells - 1)
    fy = (ys - y0)[:, None]
    fx = (xs - x0)[None, :]
    a = (g[y0][:, x0] * (1 - fx) * (1 - fy) + g[y0][:, x0 + 1] * fx * (1 - fy)
         + g[y0 + 1][:, x0] * (1 - fx) * fy + g[y0 + 1][:, x0 + 1] * fx * fy)
    return a.astype(np.float32)


def about_centre(w: int, h: int, scale: float, rotate_rad: float, sampling=AffineSample.BiLinear,
                 out_wh=None) -> Affine:
    """Scale and rotate about the image centre (``Affine`` builders,
    command.rs:3424-3485); ``out_wh`` moves the centre to another canvas."""
    ow, oh = out_wh or (w, h)
    return (Affine.new(sampling).shift(-w / 2.0, -h / 2.0).scale(scale, scale).rotate(rotate_rad)
            .shift(ow / 2.0, oh / 2.0))


# --------------------------------------------------------------------------
def c1(width: int = 1024, height: int = 768, seed: int = 0) -> Workload:
    cmd = CommandBuffer()
    d = Descriptor.rgba8_srgb(width, height)
    a = cmd.input(d)
    b = cmd.input(d)
    ab = cmd.blend(a, Rectangle(0, 0, width, height), b)
    canvas = cmd.solid_rgba(d, (0.0, 0.0, 0.0, 0.0))
    aff = about_centre(width, height, 0.75, math.radians(15.0))
    out = cmd.affine(canvas, aff, ab)
    o, od = cmd.output(out)

    def make():
        rng_a, rng_b = (np.random.default_rng([seed, k]) for k in range(2))
        return {a: random_rgba8(rng_a, width, height), b: random_rgba8(rng_b, width, height)}

    return Workload("c1_over_affine_1024x768", cmd, {a: d, b: d}, {o: od}, make, width * height,
                    dict(affine=aff.transformation))


def c2(width: int = 7680, height: int = 4320, seed: int = 0) -> Workload:
    cmd = CommandBuffer()
    d8 = Descriptor.rgba8_srgb(width, height)
    dl = Descriptor(width, height, RGBA16F, P3_LINEAR)
    src = cmd.input(d8)
    layer = cmd.input(dl)
    p3 = cmd.color_convert(src, P3_LINEAR, RGBA16F)
    lab = cmd.color_convert(p3, Color.OKLAB, LABA16F)
    lab = cmd.oklab_adjust(lab, chroma=1.2, hue_rad=0.0)
    back = cmd.color_convert(lab, P3_LINEAR, RGBA16F)
    over = cmd.blend(back, Rectangle(0, 0, width, height), layer)
    enc = cmd.color_convert(over, P3_SRGB, RGBA8)
    o, od = cmd.output(enc)

    def make():
        r0, r1 = (np.random.default_rng([seed, k]) for k in range(2))
        return {src: random_rgba8(r0, width, height), layer: random_rgba16f(r1, width, height)}

    return Workload("c2_color_over_7680x4320", cmd, {src: d8, layer: dl}, {o: od}, make, width * height)


def c2_reordered(width: int = 7680, height: int = 4320, seed: int = 0) -> Workload:
    """C2 with the over-blend moved before the Oklab adjustment: the same
    operations in an order no specialised kernel is registered for, so the
    stage runs on the general pointwise interpreter."""
    cmd = CommandBuffer()
    d8 = Descriptor.rgba8_srgb(width, height)
    dl = Descriptor(width, height, RGBA16F, P3_LINEAR)
    src = cmd.input(d8)
    layer = cmd.input(dl)
    p3 = cmd.color_convert(src, P3_LINEAR, RGBA16F)
    over = cmd.blend(p3, Rectangle(0, 0, width, height), layer)
    lab = cmd.color_convert(over, Color.OKLAB, LABA16F)
    lab = cmd.oklab_adjust(lab, chroma=1.2, hue_rad=0.0)
    back = cmd.color_convert(lab, P3_LINEAR, RGBA16F)
    enc = cmd.color_convert(back, P3_SRGB, RGBA8)
    o, od = cmd.output(enc)

    def make():
        r0, r1 = (np.random.default_rng([seed, k]) for k in range(2))
        return {src: random_rgba8(r0, width, height), layer: random_rgba16f(r1, width, height)}

    return Workload("c2_reordered_7680x4320", cmd, {src: d8, layer: dl}, {o: od}, make, width * height)


def c3_shapes(n: int = 256, seed: int = 0, lo: int = 64, hi: int = 4096):
    rng = np.random.default_rng([seed, 3])
    w = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(int)
    h = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(int)
    fmt = rng.integers(0, 2, n)  # 0 = RGBA8 sRGB, 1 = RGBA16F linear
    scale = rng.uniform(0.5, 2.0, n)
    return [(int(w[i]), int(h[i]), int(fmt[i]), float(scale[i])) for i in range(n)]


def c3(n: int = 256, seed: int = 0, max_side: int = 4096) -> Workload:
    cmd = CommandBuffer()
    shapes = c3_shapes(n, seed, hi=max_side)
    inputs, outputs, regs = {}, {}, []
    px = 0
    for i, (w, h, fmt, s) in enumerate(shapes):
        ow, oh = max(1, int(round(w * s))), max(1, int(round(h * s)))
        if fmt == 0:
            din = Descriptor.rgba8_srgb(w, h)
            conv = Color(primaries=Primaries.DisplayP3, transfer=Transfer.Srgb)
            tex = RGBA8
        else:
            din = Descriptor(w, h, RGBA16F, LINEAR_709)
            conv = P3_LINEAR
            tex = RGBA16F
        r = cmd.input(din)
        c = cmd.color_convert(r, conv, tex)
        canvas = cmd.solid_rgba(Descriptor(ow, oh, tex, conv), (0.0, 0.0, 0.0, 0.0))
        out = cmd.affine(canvas, Affine.new(AffineSample.BiLinear).scale(ow / w, oh / h), c)
        o, od = cmd.output(out)
        inputs[r] = din
        outputs[o] = od
        regs.append((r, fmt, w, h))
        px += ow * oh

    def make():
        data = {}
        for k, (r, fmt, w, h) in enumerate(regs):
            rng = np.random.default_rng([seed, 300 + k])
            data[r] = random_rgba8(rng, w, h) if fmt == 0 else random_rgba16f(rng, w, h)
        return data

    return Workload(f"c3_batch_{n}_mixed", cmd, inputs, outputs, make, px, dict(shapes=shapes))


def color_chain(cmd: CommandBuffer, reg: int, out_texel: Texel = RGBA8, hue_deg: float = 30.0,
                chroma: float = 1.2) -> int:
    """The learned-chain workload's pointwise chain: linear RGB -> XYZ -> Oklab,
    hue rotation, chroma scale, and back to sRGB-encoded RGB."""
    lab = cmd.color_convert(reg, Color.OKLAB, LABA16F)
    lab = cmd.oklab_adjust(lab, chroma=1.0, hue_rad=math.radians(hue_deg))
    lab = cmd.oklab_adjust(lab, chroma=chroma, hue_rad=0.0)
    return cmd.color_convert(lab, Color.SRGB, out_texel)


def c4(images: int = 16, width: int = 3840, height: int = 2160, seed: int = 0) -> Workload:
    cmd = CommandBuffer()
    d = Descriptor.rgba8_srgb(width, height)
    inputs, outputs, regs = {}, {}, []
    for i in range(images):
        r = cmd.input(d)
        o, od = cmd.output(color_chain(cmd, r))
        inputs[r] = d
        outputs[o] = od
        regs.append(r)

    def make():
        return {r: random_rgba8(np.random.default_rng([seed, 400 + k]), width, height) for k, r in enumerate(regs)}

    return Workload(f"c4_chain_{images}x{width}x{height}", cmd, inputs, outputs, make, images * width * height)


def c5(layers: int = 8, width: int = 3840, height: int = 2160, seed: int = 0) -> Workload:
    cmd = CommandBuffer()
    dl = Descriptor(width, height, RGBA16F, LINEAR_709)
    rng = np.random.default_rng([seed, 5])
    rots = rng.uniform(-math.pi / 4, math.pi / 4, layers)
    scales = rng.uniform(0.5, 1.5, layers)
    acc = cmd.solid_rgba(dl, (0.0, 0.0, 0.0, 0.0))
    inputs, regs, mats = {}, [], []
    for i in range(layers):
        r = cmd.input(dl)
        aff = about_centre(width, height, float(scales[i]), float(rots[i]))
        acc = cmd.affine(acc, aff, r)
        inputs[r] = dl
        regs.append(r)
        mats.append(aff.transformation)
    o, od = cmd.output(color_chain(cmd, acc))

    def make():
        data = {}
        for k, r in enumerate(regs):
            rk = np.random.default_rng([seed, 500 + k])
            alpha = smooth_noise(rk, width, height)
            data[r] = random_rgba16f(rk, width, height, alpha)
        return data

    return Workload(f"c5_composite_{layers}x{width}x{height}", cmd, inputs, {o: od}, make, width * height,
                    dict(matrices=mats))


def copy_frame(width: int = 7680, height: int = 4320, seed: int = 0) -> Workload:
    """Calibration workload: a byte copy (transmute RGBA8 sRGB -> RGBA8
    linear) of one frame, width * height * 4 bytes read and written once;
    hardware byte counters read on it give their scale."""
    cmd = CommandBuffer()
    d8 = Descriptor.rgba8_srgb(width, height)
    a = cmd.input(d8)
    o, od = cmd.output(cmd.transmute(a, Descriptor(width, height, RGBA8, LINEAR_709)))

    def make():
        return {a: random_rgba8(np.random.default_rng([seed, 900]), width, height)}

    return Workload(f"copy_rgba8_{width}x{height}", cmd, {a: d8}, {o: od}, make, width * height)


CONFIGS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5}
# Workloads measured beside the five configurations (not headline lines).
EXTRA = {"c2r": c2_reordered}

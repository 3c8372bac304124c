"""User stage kernels on the GPU: the package's HIP forms of the reference's
extension shaders (``lib/zosimos/tests/custom.rs``: mandelbrot, crt,
flat_field) loaded from their code objects with hipModuleLoadData and run as
stages of a program.

- Each kernel's RGBA f32 result, read back through an RGBA f32 register,
  against the oracle's IEEE f32 restatement (``oracle/shaders.py``).
- The three custom.rs programs end to end: the blessed perceptual hash of
  each, and the 8-bit result against the oracle pipeline.
- A program with a user kernel is captured as a hipGraph and its replays
  equal the eager launches; the allocation-free entry point refuses it; a
  missing kernel name is an error, not a launch.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import zosimos_amd as z
from oracle import blockhash as B
from oracle import color as C
from oracle import pipelines as PL
from oracle import shaders as SH
from oracle import stage as S

from gpu_util import report

pytestmark = pytest.mark.gpu

SB, SP = z.SampleBits, z.SampleParts
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HASHES = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))
F32 = lambda w, h: z.Descriptor(w, h, z.Texel(SB.Float32x4, SP.RgbA),  # noqa: E731
                                z.Color(model=z.ColorModel.Scalars, transfer=z.Transfer.Linear))
LCH8 = lambda w, h: z.Descriptor(w, h, z.Texel(SB.UInt8x4, SP.LchA), z.Color(model=z.ColorModel.Oklab))  # noqa: E731
LUMA8 = lambda w, h: z.Descriptor(w, h, z.Texel(SB.UInt8, SP.Luma), z.Color.SRGB)  # noqa: E731
NOISE = z.FractalNoise(8, 0.1, 0.2, (10.0, 10.0))


class Kernel(z.ShaderCommand):
    def __init__(self, name, desc, data=(), dtype=np.float32):
        self.src = z.ShaderSource.bundled(name)
        self.desc, self.values, self.dtype = desc, data, dtype

    def source(self):
        return self.src

    def data(self, d):
        if len(self.values):
            d.set_data(self.values, self.dtype)
        return self.desc


def rgba(name):
    from PIL import Image
    return np.ascontiguousarray(np.asarray(Image.open(os.path.join(GOLDEN, name + ".png")).convert("RGBA")))


def decode_srgb8(img):
    """The backend's decode of native sRGB RGBA8: the f32 rounding of the
    exact sRGB value for colour, k * fl(1/255) for alpha."""
    t = C.srgb_u8_decode_table().astype(np.float32)
    out = np.empty(img.shape, np.float32)
    out[..., :3] = t[img[..., :3]]
    out[..., 3] = img[..., 3].astype(np.float32) * np.float32(1.0 / 255.0)
    return out


def run(ctx, cmd, inputs, out_reg, out_desc, flags=0):
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx, flags)
    imgs = {}
    for reg, (desc, data) in inputs.items():
        imgs[reg] = ctx.image(desc)
        imgs[reg].upload(np.ascontiguousarray(data).reshape(-1))
    out = ctx.image(out_desc)
    exe.launch(imgs, {out_reg: out})
    ctx.sync()
    return out.download(), exe


def f32_of(raw, w, h):
    return raw.view(np.float32).reshape(h, w, 4)


def compare_f32(name, got, want):
    """Bit-exact where both are numbers; NaN where the oracle has NaN."""
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    bad = int((~same).sum())
    report(f"dynamic_{name}_f32", texels=got.shape[0] * got.shape[1], mismatched_components=bad)
    assert bad == 0, np.argwhere(~same)[:8]


def test_mandelbrot_kernel_matches_oracle(ctx):
    c = z.CommandBuffer()
    r = c.construct_dynamic(Kernel("mandelbrot", F32(256, 256), [3.0, 3.0, 0.6, 0.5]))
    o, od = c.output(r)
    raw, _ = run(ctx, c, {}, o, od)
    compare_f32("mandelbrot", f32_of(raw, 256, 256), SH.mandelbrot(256, 256))


def test_crt_kernel_matches_oracle(ctx):
    bg = rgba("background")
    dbg = z.Descriptor.rgba8_srgb(512, 512)
    c = z.CommandBuffer()
    i = c.input(dbg)
    r = c.unary_dynamic(i, Kernel("crt", F32(1536, 1536), [1536, 1536], np.uint32))
    o, od = c.output(r)
    raw, _ = run(ctx, c, {i: (dbg, bg)}, o, od)
    src = decode_srgb8(bg)
    compare_f32("crt", f32_of(raw, 1536, 1536), SH.crt(src, 1536, 1536))


def test_flat_field_kernel_matches_oracle(ctx):
    """Two RGBA f32 inputs of known values: the division mean / field.x is
    correctly rounded on both sides, so the result is bit-exact."""
    rng = np.random.default_rng(3)
    src = rng.random((96, 128, 4), dtype=np.float32)
    field = (rng.random((96, 128, 4), dtype=np.float32) * np.float32(0.2) + np.float32(0.01)).astype(np.float32)
    d = F32(128, 96)
    c = z.CommandBuffer()
    i, j = c.input(d), c.input(d)
    r = c.binary_dynamic(i, j, Kernel("flat_field", F32(128, 96), [0.124 / 2.0]))
    o, od = c.output(r)
    raw, _ = run(ctx, c, {i: (d, src.view(np.uint8)), j: (d, field.view(np.uint8))}, o, od)
    compare_f32("flat_field", f32_of(raw, 128, 96), SH.flat_field(src, field, np.float32(0.124 / 2.0)))


def check_hash(name, img):
    h = B.blockhash256(img)
    report(f"reference_hash_{name}", got=h, blessed=",".join(HASHES[name]))
    assert h in HASHES[name]


def test_custom_mandelbrot_program(ctx):
    """custom.rs:14-96 as written: Oklab LchA 8-bit register, then sRGB."""
    c = z.CommandBuffer()
    b = c.construct_dynamic(Kernel("mandelbrot", LCH8(256, 256), [3.0, 3.0, 0.6, 0.5]))
    o, od = c.output(c.color_convert(b, z.Color.SRGB, z.Texel(SB.UInt8x4, SP.RgbA)))
    raw, _ = run(ctx, c, {}, o, od)
    got = raw.reshape(256, 256, 4)
    check_hash("mandelbrot", got)
    lch = S.stage_encode(SH.mandelbrot(256, 256).reshape(-1, 4), 10, 20, 0x100).view(np.uint8).reshape(256, 256, 4)
    want = PL.oklab_lch8_to_srgb(lch)[1]
    diff = np.abs(got.astype(int) - want.astype(int))
    report("dynamic_mandelbrot_srgb", max_code_diff=int(diff.max()), over_1=int((diff > 1).sum()))
    assert diff.max() <= 1


def test_custom_crt_program(ctx):
    """custom.rs:98-185: the background magnified 3x through the mask, kept
    in the background's sRGB RGBA8 register and converted to sRGB."""
    bg = rgba("background")
    dbg = z.Descriptor.rgba8_srgb(512, 512)
    c = z.CommandBuffer()
    i = c.input(dbg)
    r = c.unary_dynamic(i, Kernel("crt", dbg.with_(width=1536, height=1536), [1536, 1536], np.uint32))
    o, od = c.output(c.color_convert(r, z.Color.SRGB, z.Texel(SB.UInt8x4, SP.RgbA)))
    raw, _ = run(ctx, c, {i: (dbg, bg)}, o, od)
    got = raw.reshape(1536, 1536, 4)
    check_hash("crt", got)
    src = decode_srgb8(bg)
    want = C.encode_u8(SH.crt(src, 1536, 1536), 5)
    diff = np.abs(got.astype(int) - want.astype(int))
    report("dynamic_crt_srgb", max_code_diff=int(diff.max()))
    assert diff.max() <= 1


@pytest.mark.parametrize("flags", [0, 1])
def test_custom_flat_field_program(ctx, flags):
    """custom.rs:187-285, the noise register also read back.  By default the
    noise stays FP16 between its stage and the copy for the kernel; strict
    registers (flags 1) store it in its Luma8 sRGB register as the reference
    does.  The oracle applies the shader to that field (the f16 rounding of
    the oracle's noise, or the GPU's own Luma8 codes decoded by stage.frag),
    and the 8-bit result is within one code of it; the Luma8 codes themselves
    are compared with the oracle's and reported."""
    bg = rgba("background")
    dbg = z.Descriptor.rgba8_srgb(512, 512)
    c = z.CommandBuffer()
    i = c.input(dbg)
    noise = c.distribution_fractal_noise(LUMA8(512, 512), NOISE)
    r = c.binary_dynamic(i, noise, Kernel("flat_field", dbg, [0.124 / 2.0]))
    o, od = c.output(c.color_convert(r, z.Color.SRGB, z.Texel(SB.UInt8x4, SP.RgbA)))
    on, ond = c.output(noise)
    exe = z.Linker.from_included().compile(c).lower_to(ctx, flags)
    img = ctx.image(dbg)
    img.upload(bg.reshape(-1))
    out, nout = ctx.image(od), ctx.image(ond)
    exe.launch({i: img}, {o: out, on: nout})
    ctx.sync()
    got = out.download().reshape(512, 512, 4)
    codes = nout.download().reshape(512, 512)
    check_hash("flat_field", got)
    nz = PL.fractal_noise(512, 512, 8, 0.1, 0.2, scale=(10.0, 10.0))
    want_codes = S.stage_encode(nz.reshape(-1, 4), 0, 4, 5).reshape(512, 512)
    if flags:
        field = S.stage_decode(codes.astype(np.uint32).ravel(), 0, 4, 5).reshape(512, 512, 4)
    else:
        field = nz.astype(np.float32).astype(np.float16).astype(np.float32)
    want = C.encode_u8(SH.flat_field(decode_srgb8(bg), field, np.float32(0.124 / 2.0)), 5)
    diff = np.abs(got.astype(int) - want.astype(int))
    report(f"dynamic_flat_field_srgb_flags{flags}", max_code_diff=int(diff.max()),
           luma8_codes_off=int((codes != want_codes).sum()))
    assert diff.max() <= 1


def test_graph_replays_a_user_kernel(ctx):
    """A user kernel between two of the backend's stages, captured once and
    replayed on rebound images: every replay equals the eager program."""
    d = z.Descriptor.rgba8_srgb(96, 64)
    c = z.CommandBuffer()
    i = c.input(d)
    r = c.unary_dynamic(i, Kernel("crt", d.with_(width=288, height=192), [288, 192], np.uint32))
    o, od = c.output(c.color_convert(r, z.Color(transfer=z.Transfer.Linear), z.Texel(SB.Float16x4, SP.RgbA)))
    prog = z.Linker.from_included().compile(c)
    exe = prog.lower_to(ctx)
    eager = prog.lower_to(ctx, z.PLAN_NO_GRAPH)
    rng = np.random.default_rng(5)
    src_a, src_b = ctx.image(d), ctx.image(d)
    out_a, out_b, ref = ctx.image(od), ctx.image(od), ctx.image(od)
    for k in range(3):
        src, out = (src_a, out_a) if k % 2 == 0 else (src_b, out_b)
        src.upload(rng.integers(0, 256, d.nbytes, dtype=np.uint8))
        exe.launch({i: src}, {o: out})
        eager.launch({i: src}, {o: ref})
        ctx.sync()
        assert np.array_equal(out.download(), ref.download()), k
    g = exe.graph_info()
    report("dynamic_graph", **g)
    assert g["state"] == "captured" and g["replays"] >= 1


def test_static_entry_refuses_user_kernels(ctx):
    d = z.Descriptor.rgba8_srgb(16, 16)
    c = z.CommandBuffer()
    i = c.input(d)
    o, od = c.output(c.unary_dynamic(i, Kernel("crt", d, [16, 16], np.uint32)))
    prog = z.Linker.from_included().compile(c)
    src, out = ctx.image(d), ctx.image(od)
    with pytest.raises(z.ZosimosError) as e:
        prog.launch_static(ctx, {i: src}, {o: out})
    assert e.value.code == z.Status.UNSUPPORTED


def test_missing_kernel_name_is_an_error(ctx):
    with open(os.path.join(z.SHADER_DIR, "crt.co"), "rb") as f:
        code = f.read()

    class Wrong(Kernel):
        def __init__(self):
            self.src = z.ShaderSource(code, "no_such_kernel")
            self.desc, self.values, self.dtype = z.Descriptor.rgba8_srgb(16, 16), (), np.float32

    c = z.CommandBuffer()
    i = c.input(z.Descriptor.rgba8_srgb(16, 16))
    c.output(c.unary_dynamic(i, Wrong()))
    prog = z.Linker.from_included().compile(c)
    with pytest.raises(z.ZosimosError) as e:
        prog.lower_to(ctx)
    assert e.value.code == z.Status.INVALID_ARGUMENT and "no_such_kernel" in e.value.message
    # the failed lookup leaves no error behind for the next launch to report
    d = z.Descriptor.rgba8_srgb(64, 16)
    c2 = z.CommandBuffer()
    i2 = c2.input(d)
    o2, od2 = c2.output(c2.color_convert(i2, z.Color(transfer=z.Transfer.Linear),
                                         z.Texel(SB.UInt16x4, SP.RgbA)))
    raw, _ = run(ctx, c2, {i2: (d, np.full(d.nbytes, 255, np.uint8))}, o2, od2)
    assert np.all(raw.view(np.uint16) == 65535)

"""GPU parity of the HIP backend against the FP64 oracle and the FP32
stage.frag restatement.  Every call goes through the C ABI.

Bars: integer layout work (copies, transmute, inscribe, crop, same-format
round trips of u8/u16/packed texels) bit-exact; 8-bit float outputs within
1 LSB of the FP64 oracle (colour compared where the 8-bit alpha is non-zero:
un-premultiplying divides by alpha); f16 outputs within 1 f16 ulp; linear f32
outputs within 1e-5 relative.
"""
import math

import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import workloads as W
from oracle import color as C
from oracle import pipelines as PL
from oracle import stage as S

from gpu_util import compare_u8, delta_e_u8, report, run

pytestmark = pytest.mark.gpu

RGBA8 = z.Texel(z.SampleBits.UInt8x4, z.SampleParts.RgbA)


def rand_u8(shape, seed):
    return np.random.default_rng(seed).integers(0, 256, size=shape, dtype=np.uint8)


# ---- byte-exact layout paths ----------------------------------------------------------
@pytest.mark.parametrize("desc", [z.Descriptor.rgba8_srgb(37, 13),
                                  z.Descriptor(20, 9, z.Texel(z.SampleBits.UInt16x4, z.SampleParts.RgbA)),
                                  z.Descriptor.rgba16f_linear(33, 7)])
def test_passthrough_bit_exact(ctx, desc):
    """tests/color_support.rs: input -> output is a byte copy."""
    cmd = z.CommandBuffer()
    a = cmd.input(desc)
    o, _ = cmd.output(a)
    data = rand_u8((desc.height, desc.width * desc.texel.bytes), 1)
    out, _ = run(ctx, cmd, {a: desc}, {a: data}, {o: desc})
    assert np.array_equal(out[o], data)


def test_transmute_rgba8_to_lumaa16_bit_exact(ctx):
    """tests/blend.rs:340-375 (assert at 372): the transmuted bytes equal the input."""
    d = z.Descriptor.rgba8_srgb(64, 40)
    t = z.Descriptor(64, 40, z.Texel(z.SampleBits.UInt16x2, z.SampleParts.LumaA), z.Color.SRGB)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    r = cmd.transmute(a, t)
    o, _ = cmd.output(r)
    data = rand_u8((40, 256), 2)
    out, exe = run(ctx, cmd, {a: d}, {a: data}, {o: t})
    assert "copy" in exe.describe()
    assert np.array_equal(out[o], data)


def test_srgb_u8_round_trip_through_float(ctx):
    """decode -> identity conversion -> encode returns every code."""
    d = z.Descriptor.rgba8_srgb(64, 16)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    c = cmd.color_convert(a, z.Color.SRGB, RGBA8)
    o, _ = cmd.output(c)
    data = np.tile(np.arange(256, dtype=np.uint8), 16).reshape(16, 256)
    out, exe = run(ctx, cmd, {a: d}, {a: data}, {o: d})
    assert np.array_equal(out[o], data)


@pytest.mark.parametrize("bits,parts,transfer,nbits", [
    (z.SampleBits.UInt8, z.SampleParts.Luma, z.Transfer.Srgb, 8),
    (z.SampleBits.UInt16, z.SampleParts.Luma, z.Transfer.Linear, 16),
    (z.SampleBits.UInt565, z.SampleParts.Rgb, z.Transfer.Linear, 16),
    (z.SampleBits.UInt4x4, z.SampleParts.RgbA, z.Transfer.Linear, 16),
    (z.SampleBits.UInt332, z.SampleParts.Rgb, z.Transfer.Linear, 8),
    (z.SampleBits.UInt8x2, z.SampleParts.LumaA, z.Transfer.Linear, 16),
    (z.SampleBits.UInt1010102, z.SampleParts.RgbA, z.Transfer.Linear, 32),
])
def test_staged_codec_round_trip_and_decode(ctx, bits, parts, transfer, nbits):
    """Staged texels (stage.frag decode/encode) to RGBA32F and back: the
    decoded values equal the FP32 C restatement bit for bit, and the encode
    of the decode returns every code."""
    n = 1 << min(nbits, 16)
    bpp = z.BITS_BYTES[bits]
    w = 256
    h = n // w
    color = z.Color(transfer=transfer)
    src = z.Descriptor(w, h, z.Texel(bits, parts), color)
    f32 = z.Descriptor(w, h, z.Texel(z.SampleBits.Float32x4, z.SampleParts.RgbA), z.Color(transfer=z.Transfer.Linear))
    if nbits <= 16:
        codes = np.arange(n, dtype=np.uint32)
    else:
        codes = np.random.default_rng(5).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    raw = codes.astype(np.uint32).view(np.uint8).reshape(n, 4)[:, :bpp].reshape(h, w * bpp)
    # decode to f32 (the conversion keeps primaries; only the texel changes)
    cmd = z.CommandBuffer()
    a = cmd.input(src)
    c = cmd.color_convert(a, f32.color, f32.texel)
    o, _ = cmd.output(c)
    out, _ = run(ctx, cmd, {a: src}, {a: raw}, {o: f32})
    got = out[o].view(np.float32).reshape(-1, 4)
    want = S.stage_decode(codes, bits, parts, transfer)
    if transfer == z.Transfer.Srgb:
        # the device pow and the C library's powf differ in the last ulps
        assert np.allclose(got, want, rtol=2e-6, atol=1e-7)
        return
    assert np.array_equal(got, want)
    # encode back (linear transfers are exact round trips)
    cmd = z.CommandBuffer()
    a = cmd.input(f32)
    c = cmd.color_convert(a, src.color, src.texel)
    o, _ = cmd.output(c)
    out, _ = run(ctx, cmd, {a: f32}, {a: got.view(np.uint8).reshape(h, w * 16)}, {o: src})
    assert np.array_equal(out[o], raw)


def test_f16_codec_exhaustive(ctx):
    """RGBA16F -> RGBA32F over all 65536 half values, then f32 -> f16 of
    random floats, against numpy's IEEE conversions."""
    halfs = np.arange(65536, dtype=np.uint16)
    w, h = 128, 128
    d16 = z.Descriptor.rgba16f_linear(w, h)
    d32 = z.Descriptor.rgba32f_linear(w, h)
    cmd = z.CommandBuffer()
    a = cmd.input(d16)
    c = cmd.color_convert(a, d32.color, d32.texel)
    o, _ = cmd.output(c)
    out, _ = run(ctx, cmd, {a: d16}, {a: halfs.view(np.uint8).reshape(h, w * 8)}, {o: d32})
    got = out[o].view(np.float32).ravel()
    want = halfs.view(np.float16).astype(np.float32)
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    assert same.all()
    vals = (np.random.default_rng(3).standard_normal(w * h * 4) * 300).astype(np.float32)
    cmd = z.CommandBuffer()
    a = cmd.input(d32)
    c = cmd.color_convert(a, d16.color, d16.texel)
    o, _ = cmd.output(c)
    out, _ = run(ctx, cmd, {a: d32}, {a: vals.view(np.uint8).reshape(h, w * 16)}, {o: d16})
    assert np.array_equal(out[o].view(np.uint16).ravel(), vals.astype(np.float16).view(np.uint16))


# ---- paints ------------------------------------------------------------------------
def test_inscribe_bit_exact(ctx):
    """tests/blend.rs:83-120 ('composed'): a 157x151 foreground inscribed at the
    origin of a 512x512 background, with the reference's normalize quirk."""
    bg_d, fg_d = z.Descriptor.rgba8_srgb(512, 512), z.Descriptor.rgba8_srgb(157, 151)
    cmd = z.CommandBuffer()
    bg = cmd.input(bg_d)
    fg = cmd.input(fg_d)
    r = cmd.inscribe(bg, z.Rectangle(0, 0, 157, 151), fg)
    o, _ = cmd.output(r)
    bgd, fgd = rand_u8((512, 2048), 11), rand_u8((151, 628), 12)
    out, _ = run(ctx, cmd, {bg: bg_d, fg: fg_d}, {bg: bgd, fg: fgd}, {o: bg_d})
    want = PL.paint_nearest(PL.u8_image(bgd, 512, 512).astype(np.float64), PL.u8_image(fgd, 157, 151).astype(np.float64),
                            (0, 0, 157, 151), PL.inscribe_quad((0, 0, 157, 151)), (0, 0, 512, 512))
    assert np.array_equal(out[o].reshape(512, 512, 4), want.astype(np.uint8))


def test_inscribe_square_is_a_slice_copy(ctx):
    """A square foreground at the origin (the only rect inscribe accepts,
    command.rs:1196-1198; normalize is the identity for squares) is a plain
    slice copy, including texels with alpha = 0."""
    bg_d, fg_d = z.Descriptor.rgba8_srgb(100, 80), z.Descriptor.rgba8_srgb(24, 24)
    cmd = z.CommandBuffer()
    bg = cmd.input(bg_d)
    fg = cmd.input(fg_d)
    r = cmd.inscribe(bg, z.Rectangle(0, 0, 24, 24), fg)
    o, _ = cmd.output(r)
    bgd, fgd = rand_u8((80, 400), 13), rand_u8((24, 96), 14)
    out, _ = run(ctx, cmd, {bg: bg_d, fg: fg_d}, {bg: bgd, fg: fgd}, {o: bg_d})
    want = bgd.reshape(80, 100, 4).copy()
    want[0:24, 0:24] = fgd.reshape(24, 24, 4)
    assert np.array_equal(out[o].reshape(80, 100, 4), want)


def test_crop_stretch_bit_exact(ctx):
    d = z.Descriptor.rgba8_srgb(64, 48)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    r = cmd.crop(a, z.Rectangle(8, 4, 40, 28))
    o, _ = cmd.output(r)
    data = rand_u8((48, 256), 15)
    out, _ = run(ctx, cmd, {a: d}, {a: data}, {o: d})
    want = PL.crop(PL.u8_image(data, 64, 48).astype(np.float64), (8, 4, 40, 28))
    assert np.array_equal(out[o].reshape(48, 64, 4), want.astype(np.uint8))


def test_nearest_affine_rotate45(ctx):
    """tests/blend.rs:122-167 ('affine'): foreground centred, rotated 45 deg,
    painted over the background with nearest sampling."""
    bg_d, fg_d = z.Descriptor.rgba8_srgb(512, 512), z.Descriptor.rgba8_srgb(157, 151)
    aff = (z.Affine.new(z.AffineSample.Nearest).shift(-(157 // 2), -(151 // 2)).rotate(math.pi / 4)
           .shift(256, 256))
    cmd = z.CommandBuffer()
    bg = cmd.input(bg_d)
    fg = cmd.input(fg_d)
    r = cmd.affine(bg, aff, fg)
    o, _ = cmd.output(r)
    bgd, fgd = rand_u8((512, 2048), 16), rand_u8((151, 628), 17)
    out, _ = run(ctx, cmd, {bg: bg_d, fg: fg_d}, {bg: bgd, fg: fgd}, {o: bg_d})
    quad = PL.affine_quad(np.float32(aff.transformation), 157, 151)
    want = PL.paint_nearest(PL.u8_image(bgd, 512, 512).astype(np.float64),
                            PL.u8_image(fgd, 157, 151).astype(np.float64), (0, 0, 157, 151), quad, (0, 0, 512, 512))
    got = out[o].reshape(512, 512, 4)
    mismatch = np.any(got != want.astype(np.uint8), axis=-1)
    report("nearest_affine_rotate45", mismatched_pixels=int(mismatch.sum()))
    assert mismatch.sum() == 0


def test_blend_at_offset(ctx):
    bg_d, fg_d = z.Descriptor.rgba8_srgb(50, 40), z.Descriptor.rgba8_srgb(20, 10)
    cmd = z.CommandBuffer()
    bg = cmd.input(bg_d)
    fg = cmd.input(fg_d)
    r = cmd.blend(bg, z.Rectangle(12, 20, 32, 30), fg)
    o, _ = cmd.output(r)
    bgd, fgd = rand_u8((40, 200), 18), rand_u8((10, 80), 19)
    out, _ = run(ctx, cmd, {bg: bg_d, fg: fg_d}, {bg: bgd, fg: fgd}, {o: bg_d})
    A = C.premultiply(C.decode_u8(PL.u8_image(bgd, 50, 40), 5))
    B = C.premultiply(C.decode_u8(PL.u8_image(fgd, 20, 10), 5))
    res = A.copy()
    res[20:30, 12:32] = C.over(B, A[20:30, 12:32])
    want = C.encode_u8(C.unpremultiply(res), 5)
    mr, ma, _ = compare_u8("blend_at_offset", out[o], want, alpha_ok=want[..., 3] > 0)
    assert mr <= 1 and ma <= 1


# ---- float chains against the FP64 oracle --------------------------------------------
def test_c2_chain_small(ctx):
    wl = W.c2(256, 96, seed=1)
    host = wl.make_inputs()
    (o, od), = wl.outputs.items()
    out, _ = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    src_r, layer_r = list(wl.inputs)
    lin, want = PL.c2(PL.u8_image(host[src_r], 256, 96), PL.f16_image(host[layer_r], 256, 96))
    mr, ma, exact = compare_u8("c2_chain", out[o], want, alpha_ok=want[..., 3] > 0)
    report("c2_chain", max_delta_e_ok=f"{delta_e_u8(out[o].reshape(96, 256, 4), lin, 7):.5f}")
    assert mr <= 1 and ma <= 1 and exact > 0.99


def test_c4_chain_small(ctx):
    wl = W.c4(images=2, width=160, height=90, seed=2)
    host = wl.make_inputs()
    out, _ = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    for (ri, ro) in zip(wl.inputs, wl.outputs):
        lin, want = PL.c4(PL.u8_image(host[ri], 160, 90))
        mr, ma, exact = compare_u8("c4_chain", out[ro], want)
        report("c4_chain", max_delta_e_ok=f"{delta_e_u8(out[ro].reshape(90, 160, 4), lin):.5f}")
        assert mr <= 1 and ma == 0 and exact > 0.99


def test_pointwise_batch_ragged(ctx):
    """Independent images of one pointwise signature run as one launch of the
    batched kernel, in chunks of 16,384 pixels: sizes below, at and above a
    chunk, with partial groups of 4 pixels.  The same executable then runs
    with one image at an address that is not 16-byte aligned (each stage
    launches on its own), and aligned again (bit-identical to the first)."""
    sizes = [(1, 1), (3, 1), (7, 5), (128, 128), (129, 127), (300, 61), (16384, 1), (16385, 2)]
    cmd = z.CommandBuffer()
    ins, outs, regs = {}, {}, []
    for w, h in sizes:
        d = z.Descriptor.rgba8_srgb(w, h)
        r = cmd.input(d)
        o, od = cmd.output(W.color_chain(cmd, r))
        ins[r], outs[o] = d, od
        regs.append((r, o, w, h))
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    assert "as one launch" in exe.describe() and exe.stats()[0] == 1
    host = {r: rand_u8((h, w * 4), 100 + k) for k, (r, o, w, h) in enumerate(regs)}
    imgs = {}
    for r, d in ins.items():
        imgs[r] = ctx.image(d)
        imgs[r].upload(host[r])
    out_imgs = {o: ctx.image(d) for o, d in outs.items()}

    def check(tag):
        for r, o, w, h in regs:
            _, want = PL.c4(PL.u8_image(host[r], w, h))
            mr, ma, _ = compare_u8(f"{tag}_{w}x{h}", out_imgs[o].download(), want)
            assert mr <= 1 and ma == 0

    exe.launch(imgs, out_imgs)
    ctx.sync()
    check("pointwise_batch")
    first = {o: img.download() for o, img in out_imgs.items()}

    r0, _, w0, h0 = regs[4]
    big = ctx.image(z.Descriptor.rgba8_srgb(w0 * h0 + 1, 1))
    mis = ctx.wrap(ins[r0], big.device_ptr + 4, w0 * h0 * 4)
    mis.upload(host[r0])
    for img in out_imgs.values():
        img.upload(np.zeros_like(img.download()))
    exe.launch({**imgs, r0: mis}, out_imgs)
    ctx.sync()
    check("pointwise_batch_unaligned")
    exe.launch(imgs, out_imgs)
    ctx.sync()
    for o, img in out_imgs.items():
        assert np.array_equal(img.download(), first[o])
    mis.free()
    big.free()


def test_bilinear_generator(ctx):
    """command.rs:1615-1633 / bilinear.frag at the pixel centres against the
    FP64 oracle: RGBA32F linear within 1e-5 relative (plus 1e-6), and a
    staged RGBA8 store (Scalars colour: truncating, stage.frag mux) equal to
    the oracle's code, or one off where the value sits within rounding of a
    code edge."""
    w, h = 37, 23
    params = np.random.default_rng(21).uniform(-0.5, 1.0, 24)
    b = z.Bilinear(*(tuple(params[4 * k:4 * k + 4]) for k in range(6)))
    want = PL.bilinear_grid(w, h, params)
    cmd = z.CommandBuffer()
    f32 = z.Descriptor.rgba32f_linear(w, h)
    o, od = cmd.output(cmd.bilinear(f32, b))
    out, exe = run(ctx, cmd, {}, {}, {o: od})
    assert "grid" in exe.describe()
    got = PL.f32_image(out[o], w, h)
    err = np.abs(got - want)
    report("bilinear_f32", max_abs=f"{err.max():.2e}")
    assert np.all(err <= 1e-5 * np.abs(want) + 1e-6)

    scal = z.Descriptor.rgba8_srgb(w, h).with_(color=z.Color(model=z.ColorModel.Scalars, transfer=z.Transfer.Linear))
    cmd = z.CommandBuffer()
    o, od = cmd.output(cmd.bilinear(scal, b))
    out, _ = run(ctx, cmd, {}, {}, {o: od})
    got8 = PL.u8_image(out[o], w, h).astype(np.int32)
    want8 = PL.staged_u8_store(want).astype(np.int32)
    d = want8 - got8
    edge = np.abs(np.clip(want, 0, 1) * 255 - np.rint(np.clip(want, 0, 1) * 255)) < 1e-3
    report("bilinear_u8_staged", exact=f"{float(np.mean(d == 0)):.5f}", max_diff=int(np.abs(d).max()))
    assert np.all((d == 0) | ((np.abs(d) == 1) & edge))


def test_derivative_height_f32_ragged(ctx):
    """box3.frag along the height (transposed weights) of a ragged RGBA32F
    image (partial groups of 4, clamp-to-edge at every border) after a
    colour matrix (its own stage), against the FP64 oracle within 1e-5
    relative plus 1e-6."""
    w, h = 45, 13
    d = z.Descriptor.rgba32f_linear(w, h)
    rng = np.random.default_rng(33)
    host = rng.uniform(-0.2, 1.2, (h, w, 4)).astype(np.float32)
    cmd = z.CommandBuffer()
    i = cmd.input(d)
    p3 = cmd.color_convert(i, z.Color(primaries=z.Primaries.DisplayP3, transfer=z.Transfer.Linear), d.texel)
    o, od = cmd.output(cmd.derivative(p3, z.DerivativeMethod.Sobel, z.Direction.Height))
    out, exe = run(ctx, cmd, {i: d}, {i: host}, {o: od})
    assert "box3" in exe.describe()
    m = C.convert_matrix(3, 7, 5)
    lin = host.astype(np.float64).copy()
    lin[..., :3] = np.einsum("ij,...j->...i", m, lin[..., :3])
    want = PL.box3(lin, PL.derivative_weights(z.DerivativeMethod.Sobel, width=False))
    got = PL.f32_image(out[o], w, h)
    err = np.abs(got - want)
    report("derivative_height_f32", max_abs=f"{err.max():.2e}")
    assert np.all(err <= 1e-5 * np.abs(want) + 1e-6)

    # Of a constant image (materialised for the stencil): zero colour, alpha 1.
    d8 = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    s = cmd.solid_rgba(d8, (0.2, 0.4, 0.6, 0.5))
    o, od = cmd.output(cmd.derivative(s, z.DerivativeMethod.Scharr3, z.Direction.Width))
    out, _ = run(ctx, cmd, {}, {}, {o: od})
    px = PL.u8_image(out[o], w, h)
    assert np.all(px[..., :3] == 0) and np.all(px[..., 3] == 255)


def test_linear_f32_output_within_1e5(ctx):
    """A linear RGBA32F chain (P3 conversion, Oklab round trip with chroma and
    hue) against FP64: <= 1e-5 relative (plus 1e-6 absolute near zero)."""
    w, h = 128, 64
    d = z.Descriptor.rgba32f_linear(w, h)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    r = W.color_chain(cmd, a, out_texel=z.Texel(z.SampleBits.Float32x4, z.SampleParts.RgbA))
    # color_chain ends in sRGB-encoded RGB; convert back to linear f32
    lin_out = z.Color(transfer=z.Transfer.Linear)
    r = cmd.color_convert(r, lin_out, z.Texel(z.SampleBits.Float32x4, z.SampleParts.RgbA))
    o, od = cmd.output(r)
    vals = np.random.default_rng(4).random((h, w, 4), dtype=np.float32)
    out, _ = run(ctx, cmd, {a: d}, {a: vals.view(np.uint8).reshape(h, w * 16)}, {o: od})
    got = out[o].view(np.float32).reshape(h, w, 4).astype(np.float64)
    want = PL.chain(vals.astype(np.float64))
    # Relative to the magnitude of each pixel's colour: components that
    # cancel to ~0 inside the matrix chain carry the pixel's absolute error.
    err = np.abs(got - want)
    scale = np.max(np.abs(want[..., :3]), axis=-1, keepdims=True)
    rel = float(np.max(err[..., :3] / np.maximum(scale, 1e-3)))
    report("linear_f32_chain", max_rel_to_pixel=f"{rel:.3e}", max_abs=f"{err.max():.3e}")
    assert np.all(err[..., :3] <= 1e-5 * np.maximum(scale, 1e-3))
    assert np.array_equal(got[..., 3], want[..., 3])


def test_c1_over_affine_full_size(ctx):
    """The CPU-runnable configuration at its real size (2 x 1024x768)."""
    wl = W.c1(seed=3)
    host = wl.make_inputs()
    (o, od), = wl.outputs.items()
    out, exe = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    a, b = list(wl.inputs)
    lin, want = PL.c1(PL.u8_image(host[a], 1024, 768), PL.u8_image(host[b], 1024, 768), wl.meta["affine"], (768, 1024))
    mr, ma, exact = compare_u8("c1_over_affine", out[o], want, alpha_ok=want[..., 3] > 0)
    report("c1_over_affine", max_delta_e_ok=f"{delta_e_u8(out[o].reshape(768, 1024, 4), lin):.5f}")
    assert mr <= 1 and ma <= 1 and exact > 0.99


def test_c3_batch_small(ctx):
    wl = W.c3(n=10, seed=4, max_side=300)
    host = wl.make_inputs()
    out, exe = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    assert "tile queue" in exe.describe()
    for k, ((ri, din), (ro, dout)) in enumerate(zip(wl.inputs.items(), wl.outputs.items())):
        w, h, fmt, s = wl.meta["shapes"][k]
        lin, want = PL.c3_image(host[ri], fmt, w, h, dout.width, dout.height)
        if fmt == 0:
            mr, ma, _ = compare_u8(f"c3_img{k}_u8", out[ro], want, alpha_ok=want[..., 3] > 0)
            assert mr <= 1 and ma <= 1
        else:
            got = out[ro].view(np.float16).reshape(dout.height, dout.width, 4)
            a_ok = want[..., 3:4] > 1e-3
            ulp = np.abs(got.view(np.uint16).astype(np.int32) - want.view(np.uint16).astype(np.int32))
            ok = np.broadcast_to(a_ok, ulp.shape) | (np.arange(4) == 3)
            report(f"c3_img{k}_f16", max_ulp=int(ulp[ok].max(initial=0)))
            assert ulp[ok].max(initial=0) <= 1


@pytest.mark.parametrize("w,h,ow,oh", [(1, 7, 3, 21), (7, 1, 4, 1), (1, 1, 3, 3), (2, 2, 1, 1), (5, 3, 2, 1)])
def test_resample_degenerate_sizes(ctx, w, h, ow, oh):
    """The per-image program of the resampling batch on degenerate shapes:
    sources one texel wide or tall (too narrow for the tap-pair kernels, so
    the general resampler runs), single pixels, and outputs of one pixel;
    within 1 code of the FP64 oracle where alpha is non-zero."""
    p3 = z.Color(primaries=z.Primaries.DisplayP3, transfer=z.Transfer.Srgb)
    din = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    r = cmd.input(din)
    c = cmd.color_convert(r, p3, RGBA8)
    canvas = cmd.solid_rgba(z.Descriptor(ow, oh, RGBA8, p3), (0.0, 0.0, 0.0, 0.0))
    out_r = cmd.affine(canvas, z.Affine.new(z.AffineSample.BiLinear).scale(ow / w, oh / h), c)
    o, od = cmd.output(out_r)
    data = rand_u8((h, w * 4), 70 + w + 3 * h)
    out, _ = run(ctx, cmd, {r: din}, {r: data}, {o: od})
    lin, want = PL.c3_image(data, 0, w, h, ow, oh)
    mr, ma, _ = compare_u8(f"degenerate_{w}x{h}_to_{ow}x{oh}", out[o], want, alpha_ok=want[..., 3] > 0)
    assert mr <= 1 and ma <= 1


def test_c5_layers_small(ctx):
    wl = W.c5(layers=8, width=192, height=128, seed=5)
    host = wl.make_inputs()
    (o, od), = wl.outputs.items()
    out, exe = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    layers = [PL.f16_image(host[r], 192, 128) for r in wl.inputs]
    lin, want = PL.c5(layers, wl.meta["matrices"], (128, 192))
    mr, ma, exact = compare_u8("c5_layers", out[o], want, alpha_ok=want[..., 3] > 0)
    report("c5_layers", max_delta_e_ok=f"{delta_e_u8(out[o].reshape(128, 192, 4), lin):.5f}")
    assert mr <= 1 and ma <= 1


# ---- runtime behaviour ------------------------------------------------------------
def test_knob_matches_rebuilt_program(ctx):
    """tests/knobs.rs: one executable, several parameterisations through a
    knob, each equal to a program built with that parameter."""
    d = z.Descriptor.rgba8_srgb(64, 32)
    data = rand_u8((32, 256), 20)

    def build(mat):
        cmd = z.CommandBuffer()
        a = cmd.input(d)
        canvas = cmd.solid_rgba(d, (0, 0, 0, 0))
        r = cmd.affine(canvas, z.Affine(tuple(mat), z.AffineSample.BiLinear), a)
        k = cmd.knob(r)
        o, _ = cmd.output(r)
        return cmd, a, o, k

    m0 = z.workloads.about_centre(64, 32, 0.8, 0.3).transformation
    cmd, a, o, k = build(m0)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    ia = ctx.image(d)
    ia.upload(data)
    io = ctx.image(d)
    for s, r in [(0.8, 0.3), (1.3, -0.7), (0.6, 1.1)]:
        m = z.workloads.about_centre(64, 32, s, r).transformation
        f = np.asarray(m, np.float32)
        blob = np.array([f[0], f[3], f[6], 0, f[1], f[4], f[7], 0, f[2], f[5], f[8], 0], np.float32).tobytes()
        exe.knob(k, blob)
        exe.launch({a: ia}, {o: io})
        got = io.download()
        cmd2, a2, o2, _ = build(m)
        ref, _ = run(ctx, cmd2, {a2: d}, {a2: data}, {o2: d})
        assert np.array_equal(got, ref[o2])


@pytest.mark.parametrize("bits,parts,w,h,tol", [(z.SampleBits.UInt8, z.SampleParts.Luma, 400, 400, 1),
                                                  (z.SampleBits.UInt8, z.SampleParts.Luma, 37, 23, 1),
                                                  (z.SampleBits.UInt16x2, z.SampleParts.LumaA, 61, 17, 2)])
def test_distribution_normal2d(ctx, bits, parts, w, h, tol):
    """distribution_normal2d.frag into staged grey formats (blend.rs:285-311
    for the 400x400 Luma8 case) against the FP64 oracle through the FP32
    stage.frag encode: within `tol` codes."""
    d = z.Descriptor(w, h, z.Texel(bits, parts), z.Color.SRGB)
    for p in (z.DistributionNormal2d.with_direction([0.04998, 0.0501]),
              z.DistributionNormal2d.with_diagonal(0.05, 0.3)):
        cmd = z.CommandBuffer()
        o, od = cmd.output(cmd.distribution_normal2d(d, p))
        out, _ = run(ctx, cmd, {}, {}, {o: od})
        want = S.stage_encode(PL.normal2d(w, h, p.params()).reshape(-1, 4).astype(np.float32), bits, parts, 5)
        if parts == z.SampleParts.Luma:
            got, want = out[o].reshape(-1).astype(np.int64), want.astype(np.int64)
        else:
            got = np.frombuffer(out[o].tobytes(), "<u2").reshape(-1, 2).astype(np.int64)
            want = np.stack([want & 0xFFFF, want >> 16], -1).astype(np.int64)
        report("normal2d", size=(w, h), bits=bits, max_code_diff=int(np.abs(got - want).max()))
        assert np.abs(got - want).max() <= tol


@pytest.mark.parametrize("w,h,octaves,damping,scale", [(61, 37, 4, 1.0, (100.0, 100.0)),
                                                         (256, 19, 9, 0.6, (13.5, 40.0)),
                                                         (5, 3, 1, 1.0, (5.0, 3.0))])
def test_fractal_noise_f32_bit_exact(ctx, w, h, octaves, damping, scale):
    """fractal_noise.frag into RGBA32F: the kernel rounds every product and
    sum on its own in the shader's order, as the f32 oracle does, so the
    values agree bit for bit (negative cells after the rotations included)."""
    d = z.Descriptor.rgba32f_linear(w, h)
    n = z.FractalNoise.with_octaves(octaves).set_damping(damping)
    n.grid_scale = scale
    cmd = z.CommandBuffer()
    o, od = cmd.output(cmd.distribution_fractal_noise(d, n))
    out, _ = run(ctx, cmd, {}, {}, {o: od})
    got = out[o].view(np.float32).reshape(h, w, 4)
    want = PL.fractal_noise(w, h, octaves, n.initial_amplitude, n.amplitude_damping, scale).astype(np.float32)
    diff = np.abs(got.astype(np.float64) - want)
    report("fractal_noise_f32", size=(w, h), octaves=octaves, max_abs=float(diff.max()),
           unequal=int((got != want).sum()))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("primaries,wp", [(z.Primaries.Bt709, z.Whitepoint.D65),
                                           (z.Primaries.DisplayP3, z.Whitepoint.D50)])
def test_srlab2_encode_decode(ctx, primaries, wp):
    """ColorConversion::XyzToSrLab2 / SrLab2ToXyz (srlab2.frag) on ragged
    linear RGBA32F images: encode within 2e-5 of the FP64 oracle; decode
    (white point scaling and clamp included) within 2e-5."""
    w, h = 53, 11
    rgb_c = z.Color(primaries=primaries, transfer=z.Transfer.Linear, whitepoint=wp)
    src = z.Descriptor(w, h, z.Texel(z.SampleBits.Float32x4, z.SampleParts.RgbA), rgb_c)
    lab_c = z.Color(model=z.ColorModel.SrLab2, primaries=0, transfer=z.Transfer.Linear, whitepoint=z.Whitepoint.D65)
    lab_t = z.Texel(z.SampleBits.Float32x4, z.SampleParts.LabA)
    rng = np.random.default_rng(12)
    data = rng.uniform(0.0, 1.0, (h, w, 4)).astype(np.float32)
    cmd = z.CommandBuffer()
    a = cmd.input(src)
    lab = cmd.color_convert(a, lab_c, lab_t)
    ol, old_ = cmd.output(lab)
    ob, obd = cmd.output(cmd.color_convert(lab, rgb_c, src.texel))
    out, _ = run(ctx, cmd, {a: src}, {a: data.view(np.uint8).reshape(h, -1)}, {ol: old_, ob: obd})
    got_lab = out[ol].view(np.float32).reshape(h, w, 4)
    got_rgb = out[ob].view(np.float32).reshape(h, w, 4)
    want_lab = C.rgb_to_srlab2(data[..., :3].astype(np.float64), primaries, wp)
    want_rgb = C.srlab2_to_rgb(got_lab[..., :3].astype(np.float64), primaries, wp, z.Whitepoint.D65)
    e1 = np.abs(got_lab[..., :3] - want_lab).max()
    e2 = np.abs(got_rgb[..., :3] - want_rgb).max()
    report("srlab2", primaries=primaries, wp=wp, encode_err=float(e1), decode_err=float(e2))
    assert e1 <= 2e-5 and e2 <= 2e-5
    assert np.array_equal(got_lab[..., 3], data[..., 3]) and np.array_equal(got_rgb[..., 3], data[..., 3])


@pytest.mark.parametrize("pal_desc", [z.Descriptor.rgba8_srgb(37, 23), z.Descriptor.rgba16f_linear(256, 1)])
def test_palette_lookup_bit_exact(ctx, pal_desc):
    """palette.frag: every output texel is the palette texel the oracle picks
    from the same f32 index colours (R across, B down; values outside
    [0, 1], infinities and NaN clamp to the edge), byte for byte (finite
    RGBA16F palettes: a texel passes through its decode and encode)."""
    w, h = 45, 29
    idx_d = z.Descriptor(w, h, z.Texel(z.SampleBits.Float32x4, z.SampleParts.RgbA), z.Color(transfer=z.Transfer.Linear))
    rng = np.random.default_rng(33)
    vals = rng.uniform(-0.2, 1.2, (h, w, 4)).astype(np.float32)
    vals[0, :4, 0] = [np.nan, np.inf, -np.inf, 1.0]
    if pal_desc.texel.bits == z.SampleBits.Float16x4:  # finite halves: a NaN's payload is not kept
        pal = rng.uniform(-4, 4, (pal_desc.height, pal_desc.width * 4)).astype(np.float16).view(np.uint8)
    else:
        pal = rand_u8((pal_desc.height, pal_desc.width * pal_desc.texel.bytes), 34)
    cmd = z.CommandBuffer()
    p = cmd.input(pal_desc)
    i = cmd.input(idx_d)
    o, od = cmd.output(cmd.palette(p, z.Palette(width=z.ColorChannel.R, height=z.ColorChannel.B), i))
    out, _ = run(ctx, cmd, {p: pal_desc, i: idx_d}, {p: pal, i: vals.view(np.uint8).reshape(h, -1)}, {o: od})
    texels = pal.reshape(pal_desc.height, pal_desc.width, pal_desc.texel.bytes)
    want = PL.palette_lookup(texels, vals, [1, 0, 0, 0], [0, 0, 1, 0])
    report("palette", palette=(pal_desc.width, pal_desc.height))
    assert np.array_equal(out[o].reshape(h, w, -1), want)


def test_invoke_matches_inline_program(ctx):
    """A function (colour conversion, affine resample, over) invoked twice
    gives the bytes of the same ops written inline."""
    d = z.Descriptor.rgba8_srgb(48, 32)
    m = tuple(z.workloads.about_centre(48, 32, 0.9, 0.3).transformation)

    def body_ops(cmd, a, b):
        c = cmd.color_convert(b, z.Color(primaries=z.Primaries.DisplayP3), RGBA8)
        c = cmd.color_convert(c, z.Color.SRGB, RGBA8)
        return cmd.affine(a, z.Affine(m, z.AffineSample.BiLinear), c)

    body = z.CommandBuffer()
    ba, bb = body.input(d), body.input(d)
    body.output(body_ops(body, ba, bb))
    cmd = z.CommandBuffer()
    x, y = cmd.input(d), cmd.input(d)
    (r,) = cmd.invoke(body, [x, y])
    (r,) = cmd.invoke(body, [r, y])
    o, od = cmd.output(r)
    ref = z.CommandBuffer()
    x2, y2 = ref.input(d), ref.input(d)
    o2, od2 = ref.output(body_ops(ref, body_ops(ref, x2, y2), y2))
    data = {k: rand_u8((32, 192), 60 + k) for k in range(2)}
    got, _ = run(ctx, cmd, {x: d, y: d}, {x: data[0], y: data[1]}, {o: od})
    want, _ = run(ctx, ref, {x2: d, y2: d}, {x2: data[0], y2: data[1]}, {o2: od2})
    assert np.array_equal(got[o], want[o2])


@pytest.mark.parametrize("out_wh", [(83, 41), (20, 11), (50, 30)])
def test_resize_nearest(ctx, out_wh):
    """CommandBuffer::resize (command.rs:1675-1705): the coordinate ramp's
    pixel-centre values, stored in the ramp register's RGBA8 Scalars texel
    (a staged format: truncating ``uint(c * 255)``, command.rs:1677), index
    the source as a palette with half a ramp texel added (palette.frag's
    bias): equal to the oracle byte for byte, in the default plan (the fused
    ramp is rounded through its format) and with strict registers."""
    w, h = 50, 30
    ow, oh = out_wh
    d = z.Descriptor.rgba8_srgb(w, h)
    data = rand_u8((h, w * 4), 50)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    o, od = cmd.output(cmd.resize(a, ow, oh))
    assert (od.width, od.height) == (ow, oh)
    out, exe = run(ctx, cmd, {a: d}, {a: data}, {o: od})
    strict, _ = run(ctx, cmd, {a: d}, {a: data}, {o: od}, flags=z.PLAN_STRICT_REGISTERS)
    f = np.float32
    ramp = np.zeros((oh, ow, 4), f)
    ramp[..., 0] = ((np.arange(ow, dtype=f) + f(0.5)) / f(ow))[None, :]
    ramp[..., 1] = ((np.arange(oh, dtype=f) + f(0.5)) / f(oh))[:, None]
    ramp[..., 3] = 1.0
    ramp8 = (np.clip(ramp, 0, 1) * f(255.0)).astype(np.uint32).astype(f) / f(255.0)
    want = PL.palette_lookup(data.reshape(h, w, 4), ramp8, [1, 0, 0, 0], [0, 1, 0, 0])
    assert "quant" in exe.describe()
    assert np.array_equal(out[o].reshape(oh, ow, 4), want)
    assert np.array_equal(strict[o], out[o])


@pytest.mark.parametrize("channel", [z.ColorChannel.R, z.ColorChannel.G, z.ColorChannel.B])
def test_extract_to_single_channel(ctx, channel):
    """extract (command.rs:1221-1265) stored as a single-channel sRGB texel:
    within 1 code of the oracle's stage.frag encode of the decoded channel.
    The staged 8-bit store truncates OETF(EOTF(k / 255)) * 255, which lies
    within 1e-5 of k: an ulp of the power function decides whether it lands
    on k or k - 1, so device and oracle agree on most codes, not all."""
    w, h = 37, 9
    d = z.Descriptor.rgba8_srgb(w, h)
    data = rand_u8((h, w * 4), 40 + channel)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    o, od = cmd.output(cmd.extract(a, channel))
    out, _ = run(ctx, cmd, {a: d}, {a: data}, {o: od})
    lin = C.decode_u8(data.reshape(h, w, 4), 5).astype(np.float32)
    want = S.stage_encode(lin.reshape(-1, 4), 0, 1 + channel, 5).astype(np.int64).reshape(h, w)
    got = out[o].reshape(h, w).astype(np.int64)
    report("extract", channel=channel, equal=f"{float(np.mean(got == want)):.3f}")
    assert np.abs(got - want).max() <= 1


def _pitched(data, pitch, rows):
    """Host bytes laid out with `pitch`-byte rows (Descriptor::to_aligned)."""
    buf = np.zeros(pitch * rows, np.uint8)
    buf.reshape(rows, pitch)[:, :data.shape[1]] = data
    return buf


@pytest.mark.parametrize("desc", [z.Descriptor.rgba8_srgb(37, 5), z.Descriptor.rgba16f_linear(40, 3),
                                  z.Descriptor(8, 8, z.Texel(z.SampleBits.UInt8, z.SampleParts.Luma),
                                               z.Color.SRGB)])
def test_from_buffer_bit_exact(ctx, desc):
    """command.rs:937-968: from_buffer reads rows padded to 256 bytes; the
    tight output holds each row's bytes."""
    row = desc.width * desc.texel.bytes
    pitch = (row + 255) // 256 * 256
    data = rand_u8((desc.height, row), 31)
    cmd = z.CommandBuffer()
    buf = cmd.buffer_init(_pitched(data, pitch, desc.height).tobytes() + b"\x5a" * 13)
    o, od = cmd.output(cmd.from_buffer(buf, desc))
    out, exe = run(ctx, cmd, {}, {}, {o: od})
    report("from_buffer", shape=(desc.width, desc.height), pitch=pitch)
    assert "copy" in exe.describe()
    assert np.array_equal(out[o].reshape(desc.height, row), data)


def test_from_buffer_feeds_stages_like_an_input(ctx):
    """A buffer image read in place by a colour conversion and an affine
    resample equals the same program over a bound input of those texels."""
    d = z.Descriptor.rgba8_srgb(45, 30)
    data = rand_u8((30, 180), 32)
    p3 = z.Color(primaries=z.Primaries.DisplayP3, transfer=z.Transfer.Srgb)

    def body(cmd, src):
        c = cmd.color_convert(src, p3, RGBA8)
        canvas = cmd.solid_rgba(z.Descriptor(45, 30, RGBA8, p3), (0, 0, 0, 0))
        m = z.workloads.about_centre(45, 30, 0.9, 0.4).transformation
        return cmd.output(cmd.affine(canvas, z.Affine(tuple(m), z.AffineSample.BiLinear), c))

    cmd = z.CommandBuffer()
    o, od = body(cmd, cmd.from_buffer(cmd.buffer_init(_pitched(data, 256, 30).tobytes()), d))
    got, _ = run(ctx, cmd, {}, {}, {o: od})
    cmd2 = z.CommandBuffer()
    a = cmd2.input(d)
    o2, od2 = body(cmd2, a)
    want, _ = run(ctx, cmd2, {a: d}, {a: data}, {o2: od2})
    assert np.array_equal(got[o], want[o2])


def test_buffer_knob_replaces_bytes(ctx):
    """A knob on a buffer register (WithKnob::buffer_init) swaps the bytes a
    from_buffer image reads, launch after launch, through the captured graph."""
    d = z.Descriptor.rgba8_srgb(20, 6)
    cmd = z.CommandBuffer()
    buf = cmd.buffer_zero(256 * 6)
    k = cmd.knob(buf)
    o, od = cmd.output(cmd.color_convert(cmd.from_buffer(buf, d), z.Color.SRGB, RGBA8))
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    out = ctx.image(od)
    exe.launch({}, {o: out})
    ctx.sync()
    assert not out.download().any()
    for seed in (40, 41, 42):
        data = rand_u8((6, 80), seed)
        exe.knob(k, _pitched(data, 256, 6).tobytes())
        for _ in range(2):
            exe.launch({}, {o: out})
        ctx.sync()
        assert np.array_equal(out.download().reshape(6, 80), data)
    with pytest.raises(z.ZosimosError):
        exe.knob(k, b"\0" * 100)


def test_bind_errors(ctx):
    d = z.Descriptor.rgba8_srgb(16, 16)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    o, _ = cmd.output(cmd.color_convert(a, z.Color(primaries=z.Primaries.DisplayP3), RGBA8))
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    good, wrong, out = ctx.image(d), ctx.image(z.Descriptor.rgba8_srgb(16, 8)), ctx.image(d)
    with pytest.raises(z.ZosimosError) as e:
        exe.launch({a: wrong}, {o: out})
    assert e.value.code == z.Status.MISMATCHED_DESCRIPTOR
    with pytest.raises(z.ZosimosError) as e:
        exe.launch({}, {o: out})
    assert e.value.code == z.Status.MISSING_KEY
    exe.launch({a: good}, {o: out})
    ctx.sync()

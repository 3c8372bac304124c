"""The backend's outputs against the reference's own blessed perceptual
hashes: the reference's integration tests (``lib/zosimos/tests/blend.rs``)
hash their output images with ``blockhash256`` and require an exact match
with a blessed line (``tests/util.rs:14-60``).  Same inputs (the reference's
``tests/input`` images, copied with the blessed lines into ``tests/golden``
by ``make_reference_fixtures.py``), same programs, same check.

Covered: ``run_blending`` (inscribe, ``composed``), ``run_affine`` (nearest
affine rotated 45 degrees about the centres, ``affine``),
``run_conversion`` (BT.709 transfer to sRGB, ``convert_bt709``),
``run_transmute`` (RGBA8 to LumaA16 byte copy, ``transmute``),
``run_oklab`` (a bilinear grid read as Oklab LCh and converted to sRGB,
``oklab``), ``run_adaptation`` (von Kries adaptation from D65 to D50,
``adapted``), ``run_solid`` (``solid``) and ``run_derivative`` (five 3x3
derivative filters, ``derived_<method>``); from ``knobs.rs`` one bilinear
executable re-parameterised through a knob five times
(``bilinear-knob-0`` .. ``4``); from ``buffer.rs`` ``run_bilinear``
(``bilinear_from_buffer``, its six parameters given directly rather than
through a buffer register), ``run_from_buffer`` (an 8x8 Luma8 image read
from a buffer with 256-byte rows, ``from_buffer``, both as a bound input of
that layout and through ``buffer_init``/``from_buffer``) and
``run_from_buffer_knob`` (the same through a buffer knob,
``from_buffer-with-knob``); ``run_distribution`` and
``run_distribution_normal1d`` (``distribution_normal2d``, ``_normal1d``),
``run_fractal_noise`` (``distribution_fractal2d``), ``run_srlab2``
(``srlab2``), ``run_swap`` (``swapped``).
``run_distribution_u8`` is checked against the oracle only
(``test_gpu_parity.py``): its blessed hash needs a rounding 8-bit store,
where stage.frag truncates.
"""
import json
import math
import os

import numpy as np
import pytest

import zosimos_amd as z
from oracle import blockhash as B
from oracle import color as C
from oracle import pipelines as PL
from oracle import stage as S

from gpu_util import report

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HASHES = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))


def rgba(name):
    from PIL import Image
    return np.ascontiguousarray(np.asarray(Image.open(os.path.join(GOLDEN, name + ".png")).convert("RGBA")))


def run_program(ctx, cmd, inputs, out_reg, out_desc):
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    imgs = {}
    for reg, (desc, data) in inputs.items():
        imgs[reg] = ctx.image(desc)
        imgs[reg].upload(data.reshape(-1))
    out = ctx.image(out_desc)
    exe.launch(imgs, {out_reg: out})
    ctx.sync()
    return out.download().reshape(out_desc.height, out_desc.width, 4)


def check(name, image):
    got = B.blockhash256(image)
    report(f"reference_hash_{name}", got=got, blessed=",".join(HASHES[name]))
    assert got in HASHES[name]


def test_composed_matches_reference_hash(ctx):
    bg, fg = rgba("background"), rgba("foreground")
    dbg = z.Descriptor.rgba8_srgb(bg.shape[1], bg.shape[0])
    dfg = z.Descriptor.rgba8_srgb(fg.shape[1], fg.shape[0])
    cmd = z.CommandBuffer()
    b, f = cmd.input(dbg), cmd.input(dfg)
    r = cmd.inscribe(b, z.Rectangle(0, 0, fg.shape[1], fg.shape[0]), f)
    o, od = cmd.output(r)
    check("composed", run_program(ctx, cmd, {b: (dbg, bg), f: (dfg, fg)}, o, od))


def test_affine_matches_reference_hash(ctx):
    bg, fg = rgba("background"), rgba("foreground")
    dbg = z.Descriptor.rgba8_srgb(bg.shape[1], bg.shape[0])
    dfg = z.Descriptor.rgba8_srgb(fg.shape[1], fg.shape[0])
    aff = (z.Affine.new(z.AffineSample.Nearest)
           .shift(-float(fg.shape[1] // 2), -float(fg.shape[0] // 2))
           .rotate(math.pi / 4)
           .shift(float(bg.shape[1] // 2), float(bg.shape[0] // 2)))
    cmd = z.CommandBuffer()
    b, f = cmd.input(dbg), cmd.input(dfg)
    o, od = cmd.output(cmd.affine(b, aff, f))
    check("affine", run_program(ctx, cmd, {b: (dbg, bg), f: (dfg, fg)}, o, od))


def test_convert_bt709_matches_reference_hash(ctx):
    bg = rgba("background")
    srgb = z.Descriptor.rgba8_srgb(bg.shape[1], bg.shape[0])
    bt = srgb.with_(color=z.Color.BT709_RGB)
    cmd = z.CommandBuffer()
    i = cmd.input(bt)
    o, od = cmd.output(cmd.color_convert(i, srgb.color, srgb.texel))
    check("convert_bt709", run_program(ctx, cmd, {i: (bt, bg)}, o, od))


def test_transmute_matches_reference_hash(ctx):
    """blend.rs:340-375: the transmuted bytes equal the input, and the LumaA16
    image hashes to the blessed line."""
    bg = rgba("background")
    h, w = bg.shape[:2]
    src = z.Descriptor.rgba8_srgb(w, h)
    dst = z.Descriptor(w, h, z.Texel(z.SampleBits.UInt16x2, z.SampleParts.LumaA), z.Color.SRGB)
    cmd = z.CommandBuffer()
    i = cmd.input(src)
    o, od = cmd.output(cmd.transmute(i, dst))
    out = run_program(ctx, cmd, {i: (src, bg)}, o, od)
    assert np.array_equal(out.reshape(-1), bg.reshape(-1))
    check("transmute", B.lumaa16_to_rgba8(out, w, h))


def test_oklab_matches_reference_hash(ctx):
    """blend.rs:449-508: lightness 0.4 + 0.4, chroma 0 to 1 down, hue 0 to 1
    across, generated into RGBA8 with a Scalars colour, transmuted to an LchA
    Oklab texel and converted to sRGB.  Against the FP64 oracle within 1 code
    where the oracle's value is not within rounding of a code edge, and the
    blessed hash."""
    w = h = 400
    srgb = z.Descriptor.rgba8_srgb(w, h)
    scal = srgb.with_(color=z.Color(model=z.ColorModel.Scalars, transfer=z.Transfer.Linear))
    lch = z.Descriptor(w, h, z.Texel(z.SampleBits.UInt8x4, z.SampleParts.LchA), z.Color.OKLAB)
    plane = z.Bilinear(u_min=(0.4, 0.0, 0.0, 1.0), u_max=(0.4, 0.0, 1.0, 1.0),
                       v_min=(0.4, 0.0, 0.0, 1.0), v_max=(0.4, 1.0, 0.0, 1.0))
    cmd = z.CommandBuffer()
    g = cmd.bilinear(scal, plane)
    t = cmd.transmute(g, lch)
    o, od = cmd.output(cmd.color_convert(t, srgb.color, srgb.texel))
    out = run_program(ctx, cmd, {}, o, od)
    staged = PL.staged_u8_store(PL.bilinear_grid(w, h, plane.params()))
    _, want = PL.oklab_lch8_to_srgb(staged)
    d = np.abs(out.astype(np.int32) - want.astype(np.int32))
    report("oklab_plane", max_code=int(d.max()), exact=f"{float(np.mean(d == 0)):.5f}")
    assert d[..., 3].max() == 0 and d.max() <= 1
    check("oklab", out)


def test_srlab2_matches_reference_hash(ctx):
    """blend.rs:510-575: the grid of run_oklab transmuted to an LchA SrLab2
    (D65) texel and converted to sRGB; within 1 code of the FP64 oracle
    where it is not within rounding of a code edge, and the blessed hash."""
    w = h = 400
    srgb = z.Descriptor.rgba8_srgb(w, h)
    scal = srgb.with_(color=z.Color(model=z.ColorModel.Scalars, transfer=z.Transfer.Linear))
    lch = z.Descriptor(w, h, z.Texel(z.SampleBits.UInt8x4, z.SampleParts.LchA), z.Color.SRLAB2_D65)
    plane = z.Bilinear(u_min=(0.4, 0.0, 0.0, 1.0), u_max=(0.4, 0.0, 1.0, 1.0),
                       v_min=(0.4, 0.0, 0.0, 1.0), v_max=(0.4, 1.0, 0.0, 1.0))
    cmd = z.CommandBuffer()
    t = cmd.transmute(cmd.bilinear(scal, plane), lch)
    o, od = cmd.output(cmd.color_convert(t, srgb.color, srgb.texel))
    out = run_program(ctx, cmd, {}, o, od)
    lab = C.decode_u8(PL.staged_u8_store(PL.bilinear_grid(w, h, plane.params())), 0x100)
    lin = lab.copy()
    lin[..., :3] = C.srlab2_to_rgb(lab[..., :3])
    want = C.encode_u8(lin, 5)
    d = np.abs(out.astype(np.int32) - want.astype(np.int32))
    report("srlab2_plane", max_code=int(d.max()), exact=f"{float(np.mean(d == 0)):.5f}")
    assert d[..., 3].max() == 0 and d.max() <= 1
    check("srlab2", out)


def test_adapted_matches_reference_hash(ctx):
    """blend.rs:169-197: the background adapted from D65 to D50 by von Kries,
    within 1 LSB of the FP64 oracle and hashing to the blessed line."""
    bg = rgba("background")
    h, w = bg.shape[:2]
    src = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    i = cmd.input(src)
    a = cmd.chromatic_adaptation(i, z.ChromaticAdaptationMethod.VonKries, z.Whitepoint.D50)
    o, od = cmd.output(a)
    assert od.color.whitepoint == z.Whitepoint.D50
    out = run_program(ctx, cmd, {i: (src, bg)}, o, od)
    lin = C.decode_u8(bg, 5)
    want = lin.copy()
    want[..., :3] = np.einsum("ij,...j->...i", C.adapt_rgb_matrix(3, 5, 3, 1), lin[..., :3])
    d = np.abs(out.astype(np.int32) - C.encode_u8(want, 5).astype(np.int32))
    report("adapted", max_code=int(d.max()), exact=f"{float(np.mean(d == 0)):.5f}")
    assert d.max() <= 1
    check("adapted", out)


def test_solid_matches_reference_hash(ctx):
    """blend.rs:620-643: a 400x400 solid (0.5, 0.5, 1, 1) in RGBA8 sRGB."""
    d = z.Descriptor.rgba8_srgb(400, 400)
    cmd = z.CommandBuffer()
    o, od = cmd.output(cmd.solid_rgba(d, (0.5, 0.5, 1.0, 1.0)))
    out = run_program(ctx, cmd, {}, o, od)
    want = C.encode_u8(np.array([0.5, 0.5, 1.0, 1.0]), 5)
    assert np.all(out == want)
    check("solid", out)


KNOB_PLANES = [((0.0, 0.0, 0.5, 0.5), (0.5, 0.5, 0.5, 0.5)), ((0.2, 0.0, 0.5, 0.5), (0.5, 0.5, 0.5, 0.5)),
               ((0.4, 0.0, 0.5, 0.5), (0.5, 0.5, 0.5, 0.5)), ((0.5, 0.0, 0.4, 0.5), (0.5, 0.5, 0.5, 0.5)),
               ((0.5, 0.0, 0.2, 0.5), (0.5, 0.5, 0.5, 0.5))]


def test_bilinear_knob_matches_reference_hashes(ctx):
    """knobs.rs:46-167: one executable of a 512x512 bilinear (RGBA8 sRGB), its
    parameters overridden through a knob for each of five launches; each
    output within 1 LSB of the FP64 oracle and hashing to its blessed line."""
    d = z.Descriptor.rgba8_srgb(512, 512)
    cmd = z.CommandBuffer()
    g = cmd.bilinear(d, z.Bilinear(u_min=(0, 0, 1, 1), u_max=(1, 1, 1, 1), v_min=(0, 0, 1, 1), v_max=(1, 1, 1, 1)))
    k = cmd.knob(g)
    o, od = cmd.output(g)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    out = ctx.image(od)
    for idx, (lo, hi) in enumerate(KNOB_PLANES):
        b = z.Bilinear(u_min=lo, u_max=hi, v_min=lo, v_max=hi)
        exe.knob(k, np.asarray(b.params(), np.float32).tobytes())
        exe.launch({}, {o: out})
        ctx.sync()
        got = out.download().reshape(512, 512, 4)
        want = C.encode_u8(PL.bilinear_grid(512, 512, b.params()), 5)
        assert np.abs(got.astype(np.int32) - want.astype(np.int32)).max() <= 1
        check(f"bilinear-knob-{idx}", got)


def test_bilinear_from_buffer_matches_reference_hash(ctx):
    """buffer.rs:121-149: a 256x256 bilinear (RGBA8 sRGB) with the buffer's six
    vec4 as its parameters."""
    p = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.7, 1.0, 0.0, 0.0, 0.3, 1.0, 0.0, 1.0, 0.3, 1.0] + [0.0, 0.0, 0.0, 1.0] * 2
    b = z.Bilinear(*(tuple(p[4 * i:4 * i + 4]) for i in range(6)))
    d = z.Descriptor.rgba8_srgb(256, 256)
    cmd = z.CommandBuffer()
    o, od = cmd.output(cmd.bilinear(d, b))
    out = run_program(ctx, cmd, {}, o, od)
    want = C.encode_u8(PL.bilinear_grid(256, 256, p), 5)
    assert np.abs(out.astype(np.int32) - want.astype(np.int32)).max() <= 1
    check("bilinear_from_buffer", out)


def test_from_buffer_matches_reference_hash(ctx):
    """buffer.rs:39-65: 2,048 buffer bytes read as an 8x8 Luma8 sRGB image
    with 256-byte rows (row 1 black, the rest white), output tight; the bytes
    of the output equal the buffer's and the image hashes to the blessed
    line."""
    a = np.full(8 * 256, 0xFF, np.uint8)
    a[256:264] = 0
    buf = ctx.image(z.Descriptor.rgba8_srgb(512, 1))  # 2,048 bytes of device memory
    buf.upload(a)
    luma = z.Descriptor(8, 8, z.Texel(z.SampleBits.UInt8, z.SampleParts.Luma), z.Color.SRGB)
    img = ctx.wrap(luma.with_(row_stride=256), buf.device_ptr, a.nbytes)
    cmd = z.CommandBuffer()
    i = cmd.input(luma.with_(row_stride=256))
    o, od = cmd.output(i)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    out = ctx.image(od)
    exe.launch({i: img}, {o: out})
    ctx.sync()
    got = out.download().reshape(8, 8)
    assert np.array_equal(got, a.reshape(8, 256)[:, :8])
    check("from_buffer", np.stack([got, got, got, np.full_like(got, 255)], -1))


@pytest.mark.parametrize("name,dist", [
    ("distribution_normal2d", lambda: z.DistributionNormal2d.with_diagonal(0.2, 0.2)),
    ("distribution_normal1d", lambda: z.DistributionNormal2d.with_direction([0.04998, 0.0501]))])
def test_distribution_normal2d_matches_reference_hash(ctx, name, dist):
    """blend.rs:226-282: normal distributions generated into a 400x400 sRGB
    LumaA16 image (staged: sRGB transfer, truncating 16-bit store); within 2
    codes of the oracle and hashing to the blessed line."""
    d = z.Descriptor(400, 400, z.Texel(z.SampleBits.UInt16x2, z.SampleParts.LumaA), z.Color.SRGB)
    p = dist()
    cmd = z.CommandBuffer()
    o, od = cmd.output(cmd.distribution_normal2d(d, p))
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    out = ctx.image(od)
    exe.launch({}, {o: out})
    ctx.sync()
    raw = out.download()
    want = S.stage_encode(PL.normal2d(400, 400, p.params()).reshape(-1, 4).astype(np.float32), 11, 5, 5)
    got = np.frombuffer(raw.tobytes(), "<u2").reshape(-1, 2)
    w16 = np.stack([want & 0xFFFF, want >> 16], -1)
    report(name, max_code_diff=int(np.abs(got.astype(np.int64) - w16).max()))
    assert np.abs(got.astype(np.int64) - w16).max() <= 2
    check(name, B.lumaa16_to_rgba8(raw, 400, 400))


def test_fractal_noise_matches_reference_hash(ctx):
    """blend.rs:314-338: four octaves of fractal noise into 400x400 RGBA8
    sRGB; within 1 code of the oracle and hashing to the blessed line."""
    d = z.Descriptor.rgba8_srgb(400, 400)
    cmd = z.CommandBuffer()
    o, od = cmd.output(cmd.distribution_fractal_noise(d, z.FractalNoise.with_octaves(4)))
    out = run_program(ctx, cmd, {}, o, od)
    want = C.encode_u8(PL.fractal_noise(400, 400, 4, 0.25, 1.0), 5)
    report("fractal_noise", max_code_diff=int(np.abs(out.astype(int) - want.astype(int)).max()))
    assert np.abs(out.astype(int) - want.astype(int)).max() <= 1
    check("distribution_fractal2d", out)


def test_palette_reference_program(ctx):
    """blend.rs:377-421 (run_palette) with strict registers: the background
    sampled at the coordinates of an 8-bit sRGB red-green ramp.  The ramp is
    computed in FP32 on the device and in FP64 by the oracle, then stored in
    its 8-bit register; every output texel equals the oracle's lookup.  The
    blessed ``palette`` hash is not reproduced exactly: its two blessed lines
    differ from each other in 6 of 256 bits (it depends on the machine that
    rendered it), and this output lies no farther than that from one of
    them."""
    bg = rgba("background")
    h, w = bg.shape[:2]
    ramp_p = z.Bilinear(u_min=(0.0, 0.0, 0.0, 1.0), v_min=(0.0, 0.0, 0.0, 1.0), uv_min=(0.0, 0.0, 0.0, 1.0),
                        u_max=(0.7, 0.0, 0.0, 1.0), v_max=(0.0, 0.7, 0.0, 1.0), uv_max=(0.3, 0.3, 0.0, 1.0))
    d = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    i = cmd.input(d)
    ramp = cmd.bilinear(z.Descriptor.rgba8_srgb(400, 400), ramp_p)
    o, od = cmd.output(cmd.palette(i, z.Palette(width=z.ColorChannel.R, height=z.ColorChannel.G), ramp))
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx, z.PLAN_STRICT_REGISTERS)
    img = ctx.image(d)
    img.upload(bg.reshape(-1))
    out_img = ctx.image(od)
    exe.launch({i: img}, {o: out_img})
    ctx.sync()
    out = out_img.download().reshape(400, 400, 4)
    ramp_lin = C.decode_u8(C.encode_u8(PL.bilinear_grid(400, 400, ramp_p.params()), 5), 5)
    want = PL.palette_lookup(bg, ramp_lin, [1, 0, 0, 0], [0, 1, 0, 0])
    bad = np.argwhere(np.any(out != want, axis=-1))
    report("palette_program", differing=len(bad))
    assert len(bad) == 0
    got = B.blockhash256(out)
    spread = bin(int(HASHES["palette"][0], 16) ^ int(HASHES["palette"][1], 16)).count("1")
    dist = min(bin(int(got, 16) ^ int(b, 16)).count("1") for b in HASHES["palette"])
    report("palette_hash", hamming=dist, blessed_spread=spread)
    assert dist <= spread


@pytest.mark.parametrize("flags", [0, 1])
def test_swapped_matches_reference_hash(ctx, flags):
    """blend.rs:423-447 (run_swap): R and G of the background extracted and
    injected into each other's channel; the result is the background with R
    and G swapped byte for byte (fused, and with strict registers, where the
    channels pass through single-channel sRGB texels), and hashes to the
    blessed line."""
    bg = rgba("background")
    h, w = bg.shape[:2]
    d = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    i = cmd.input(d)
    r = cmd.extract(i, z.ColorChannel.R)
    g = cmd.extract(i, z.ColorChannel.G)
    s = cmd.inject(cmd.inject(i, z.ColorChannel.G, r), z.ColorChannel.R, g)
    o, od = cmd.output(s)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx, flags)
    img = ctx.image(d)
    img.upload(bg.reshape(-1))
    out_img = ctx.image(od)
    exe.launch({i: img}, {o: out_img})
    ctx.sync()
    out = out_img.download().reshape(h, w, 4)
    d_ = np.abs(out.astype(int) - bg[..., [1, 0, 2, 3]].astype(int))
    report("swapped", flags=flags, max_code=int(d_.max()), exact=f"{float(np.mean(d_ == 0)):.5f}")
    if flags == 0:
        assert d_.max() == 0
    else:  # an 8-bit sRGB channel store truncates: OETF(EOTF(k)) may land a hair below k
        assert d_.max() <= 1
    check("swapped", out)


LUMA8 = z.Descriptor(8, 8, z.Texel(z.SampleBits.UInt8, z.SampleParts.Luma), z.Color.SRGB)


def test_buffer_init_from_buffer_matches_reference_hash(ctx):
    """buffer.rs:39-65 through the buffer register itself: buffer_init of the
    2,048 bytes, from_buffer as an 8x8 Luma8 image, output tight."""
    a = np.full(8 * 256, 0xFF, np.uint8)
    a[256:264] = 0
    cmd = z.CommandBuffer()
    buf = cmd.buffer_init(a.tobytes())
    o, od = cmd.output(cmd.from_buffer(buf, LUMA8))
    assert od.row_stride in (0, 8)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    out = ctx.image(od)
    exe.launch({}, {o: out})
    ctx.sync()
    got = out.download().reshape(8, 8)
    assert np.array_equal(got, a.reshape(8, 256)[:, :8])
    check("from_buffer", np.stack([got, got, got, np.full_like(got, 255)], -1))


def test_from_buffer_knob_matches_reference_hash(ctx):
    """buffer.rs:66-119: a knobbed buffer of 0xff bytes read as an 8x8 Luma8
    image; the launch patches row 1 to black through the knob."""
    a = np.full(8 * 256, 0xFF, np.uint8)
    cmd = z.CommandBuffer()
    buf = cmd.buffer_init(a.tobytes())
    k = cmd.knob(buf)
    o, od = cmd.output(cmd.from_buffer(buf, LUMA8))
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    out = ctx.image(od)
    exe.launch({}, {o: out})
    ctx.sync()
    assert np.all(out.download() == 0xFF)
    a[256:264] = 0
    exe.knob(k, a.tobytes())
    exe.launch({}, {o: out})
    ctx.sync()
    got = out.download().reshape(8, 8)
    assert np.array_equal(got, a.reshape(8, 256)[:, :8])
    check("from_buffer-with-knob", np.stack([got, got, got, np.full_like(got, 255)], -1))


@pytest.mark.parametrize("method,name", [(z.DerivativeMethod.Scharr3, "Scharr3"),
                                         (z.DerivativeMethod.Scharr3To4Bit, "Scharr3To4Bit"),
                                         (z.DerivativeMethod.Scharr3To8Bit, "Scharr3To8Bit"),
                                         (z.DerivativeMethod.Prewitt, "Prewitt"),
                                         (z.DerivativeMethod.Sobel, "Sobel")])
def test_derivative_matches_reference_hash(ctx, method, name):
    """blend.rs:577-618: the background's derivative along the width, within
    1 LSB of the FP64 oracle and hashing to the method's blessed line."""
    bg = rgba("background")
    h, w = bg.shape[:2]
    src = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    i = cmd.input(src)
    o, od = cmd.output(cmd.derivative(i, method, z.Direction.Width))
    out = run_program(ctx, cmd, {i: (src, bg)}, o, od)
    want = C.encode_u8(PL.box3(C.decode_u8(bg, 5), PL.derivative_weights(method)), 5)
    d = np.abs(out.astype(np.int32) - want.astype(np.int32))
    report(f"derived_{name}", max_code=int(d.max()), exact=f"{float(np.mean(d == 0)):.5f}")
    assert d.max() <= 1
    check(f"derived_{name}", out)



def test_generic_palette_reference_program(ctx):
    """tests/generic.rs: a generic function -- its input's descriptor is a
    generic parameter -- samples its input as a palette at a 2048x2048 ramp
    of u v (``uv_max = (1, 1, 0, 0)`` into an RGBA8 Scalars texel, whose
    staged store truncates); the main buffer instantiates it with the
    background's descriptor and the linker monomorphizes the call
    (``Linker::link(&main, &[], &[f], &[&[1], &[]])``).  The output equals
    the oracle byte for byte (the ramp in IEEE f32, as the device computes
    it) and reproduces a blessed ``generic`` hash exactly."""
    bg = rgba("background")
    h, w = bg.shape[:2]
    f = z.CommandBuffer()
    a = f.generic(z.GenericDeclaration())
    img_in = f.input_generic(a)
    ramp_d = z.Descriptor(2048, 2048, z.Texel(z.SampleBits.UInt8x4, z.SampleParts.RgbA),
                          z.Color(model=z.ColorModel.Scalars, transfer=z.Transfer.Linear))
    idx = f.bilinear(ramp_d, z.Bilinear(uv_max=(1.0, 1.0, 0.0, 0.0)))
    f.output(f.palette(img_in, z.Palette(height=z.ColorChannel.R, width=z.ColorChannel.G), idx))
    main = z.CommandBuffer()
    conv = main.function(f.computed_signature())
    d = z.Descriptor.rgba8_srgb(w, h)
    i = main.input(d)
    (res,) = main.invoke(conv, z.InvocationArguments(generics=(main.register_descriptor(i),), arguments=(i,)))
    o, od = main.output(res)
    assert (od.width, od.height) == (2048, 2048)
    exe = z.Linker().link(main, [], [f], [[1], []]).lower_to(ctx)
    img = ctx.image(d)
    img.upload(bg.reshape(-1))
    out_img = ctx.image(od)
    exe.launch({i: img}, {o: out_img})
    ctx.sync()
    out = out_img.download().reshape(2048, 2048, 4)
    f32 = np.float32
    u = (np.arange(2048, dtype=f32) + f32(0.5)) / f32(2048)
    uv = (u[:, None] * u[None, :]).astype(f32)
    code = (uv * f32(255.0)).astype(np.uint32).astype(f32) / f32(255.0)
    ramp = np.zeros((2048, 2048, 4), f32)
    ramp[..., 0] = code
    ramp[..., 1] = code
    want = PL.palette_lookup(bg, ramp, [0, 1, 0, 0], [1, 0, 0, 0])
    diff = int(np.any(out != want, axis=-1).sum())
    got = B.blockhash256(out)
    dist = min(bin(int(got, 16) ^ int(b, 16)).count("1") for b in HASHES["generic"])
    report("generic_program", differing=diff, hamming=dist)
    assert diff == 0
    assert got in HASHES["generic"]

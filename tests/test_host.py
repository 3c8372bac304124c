"""Host-side logic of the backend (no GPU): the C ABI exports, the C++
front end's type checks (command.rs error classes), the High emission and
the fused plans."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import _abi
from zosimos_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    for h in ("zosimos_hip.h", "zosimos_command.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:zh_status|void\*|const char\*)\s+(z[hs]_\w+)\(", text, re.M)
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = header_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.exported_symbols())
    assert "gfx950" in z.version()


def rgba8(w=16, h=8):
    return z.Descriptor.rgba8_srgb(w, h)


def expect(code, fn, *args):
    with pytest.raises(z.ZosimosError) as e:
        fn(*args)
    assert e.value.code == code, e.value
    return e.value


def test_command_type_errors():
    cmd = z.CommandBuffer()
    a = cmd.input(rgba8(16, 8))
    b = cmd.input(rgba8(8, 8))
    c = cmd.input(z.Descriptor.rgba16f_linear(8, 8))
    # inscribe: rect must be the above image (command.rs:1196-1198)
    expect(z.Status.OTHER, cmd.inscribe, a, z.Rectangle(0, 0, 4, 4), b)
    # ... inside the below image
    expect(z.Status.OTHER, cmd.inscribe, a, z.Rectangle(10, 0, 18, 8), b)
    # conflicting texels
    err = expect(z.Status.TYPE, cmd.inscribe, a, z.Rectangle(0, 0, 8, 8), c)
    assert err.is_type_err()
    # transmute with mismatched size / texel bytes
    expect(z.Status.TYPE, cmd.transmute, a, rgba8(8, 8))
    expect(z.Status.TYPE, cmd.transmute, a, z.Descriptor.rgba16f_linear(16, 8))
    # singular affine (command.rs:1648-1656)
    expect(z.Status.OTHER, cmd.affine, a, z.Affine.new(z.AffineSample.Nearest).scale(0.0, 1.0), a)
    # no conversion Oklab -> Oklab
    lab = cmd.color_convert(a, z.Color.OKLAB, z.Texel(z.SampleBits.Float16x4, z.SampleParts.LabA))
    expect(z.Status.TYPE, cmd.color_convert, lab, z.Color.OKLAB, z.Texel(z.SampleBits.Float16x4, z.SampleParts.LabA))
    # blend rect size / position
    expect(z.Status.OTHER, cmd.blend, a, z.Rectangle(0, 0, 4, 8), b)
    expect(z.Status.OTHER, cmd.blend, a, z.Rectangle(12, 0, 20, 8), b)
    # bad register
    expect(z.Status.OTHER, cmd.crop, 99, z.Rectangle(0, 0, 1, 1))
    # inconsistent descriptor
    expect(z.Status.TYPE, cmd.input, z.Descriptor(0, 8, z.Texel(10, 8)))
    # render needs a texel with a texture form (command.rs:1719-1736): the
    # 3-byte UInt8x3 and 6-byte UInt16x3 texels have no StageKind
    # (shaders/stage.rs:63-72); an output of them is fine
    for bits in (z.SampleBits.UInt8x3, z.SampleBits.UInt16x3):
        r = cmd.color_convert(a, z.Color.SRGB, z.Texel(bits, z.SampleParts.Rgb))
        expect(z.Status.TYPE, cmd.render, r)
        cmd.output(r)
    for t in (z.Texel(z.SampleBits.UInt8x4, z.SampleParts.RgbA), z.Texel(z.SampleBits.Float16x4, z.SampleParts.RgbA),
              z.Texel(z.SampleBits.UInt565, z.SampleParts.Rgb)):
        cmd.render(cmd.color_convert(a, z.Color.SRGB, t))


def test_inscribe_high_emission_keeps_normalize_quirk():
    cmd = z.CommandBuffer()
    bg = cmd.input(rgba8(32, 32))
    fg = cmd.input(rgba8(10, 6))
    r = cmd.inscribe(bg, z.Rectangle(0, 0, 10, 6), fg)
    cmd.output(r)
    tex, ops = z.Linker.from_included().compile(cmd).high()
    draws = [o for o in ops if o.high == 2]
    assert len(draws) == 2
    assert draws[0].target == 0 and draws[1].target == 1  # Discard then Load
    # placement quad is 10 x 10 (Rectangle::normalize, command.rs:3541)
    assert list(draws[1].quad) == [0, 0, 10, 0, 10, 10, 0, 10]
    assert (draws[1].selection.max_x, draws[1].selection.max_y) == (10, 6)


def plan(wl_or_cmd, flags=0):
    cmd = wl_or_cmd.commands if hasattr(wl_or_cmd, "commands") else wl_or_cmd
    return z.Linker.from_included().compile(cmd).lower_to(None, flags)


def test_plan_c2_is_one_fused_kernel():
    e = plan(W.c2(64, 32))
    text = e.describe()
    assert text.count("stage ") == 1
    assert "pointwise 64x32 flat: load#0 mat3 cbrt mat3 cube mat3 clamp01 premul load#1 premul over unpremul store#2" in text
    assert e.stats() == (1, 64 * 32 * 16)


def test_plan_c1_c5_resample_layers():
    t1 = plan(W.c1(64, 48)).describe()
    assert "resample 64x48" in t1 and "layer0(bilinear-over 64x48)[load#0 premul load#1 premul over]" in t1
    t5 = plan(W.c5(8, 64, 48)).describe()
    assert t5.count("stage ") == 1 and "layer7(bilinear-over" in t5
    assert "post[unpremul mat3 cbrt mat3 cube mat3 clamp01 store#8]" in t5


def test_plan_c3_batches_into_one_tile_queue_per_form():
    """Independent resample stages batch by kernel form: one persistent
    tile-queue launch per form, every stage in exactly one batch; under
    ZH_PLAN_PAIR the two forms share one grid that deals their queues
    alternately, first form's stages first."""
    e = plan(W.c3(12, max_side=300))
    text = e.describe()
    forms = re.findall(r"^stage .*\[kernel (\w+)", text, re.M)
    batches = re.findall(r"batch: stages ([\d,]+) as one tile queue", text)
    ids = sorted(int(i) for b in batches for i in b.split(","))
    assert ids == list(range(12)) and len(batches) == len(set(forms)) == 2
    assert e.stats()[0] == 2
    paired = plan(W.c3(12, max_side=300), z.PLAN_PAIR)
    text = paired.describe()
    assert paired.stats()[0] == 1
    assert "[kernel pair w_s8m_s8_sep + w_hm_h_sep] over 12 stages" in text
    order = re.findall(r"batch: stages ([\d,]+) as one grid dealing two tile queues alternately", text)
    ids = [int(i) for i in order[0].split(",")]
    assert sorted(ids) == list(range(12))
    split = sum(1 for f in forms if f == forms[ids[0]])
    assert {forms[i] for i in ids[:split]} == {forms[ids[0]]} and {forms[i] for i in ids[split:]} != {forms[ids[0]]}
    # no pairs without run-time kernels
    assert "pair" not in plan(W.c3(12, max_side=300), z.PLAN_PAIR | z.PLAN_NO_JIT).describe()


def test_plan_strict_registers_quantises():
    cmd = z.CommandBuffer()
    a = cmd.input(rgba8())
    lab = cmd.color_convert(a, z.Color.OKLAB, z.Texel(z.SampleBits.UInt8x4, z.SampleParts.LabA))
    back = cmd.color_convert(lab, z.Color.SRGB, z.Texel(10, 8))
    cmd.output(back)
    assert "quant" not in plan(cmd).describe()
    assert "quant" in plan(cmd, z.PLAN_STRICT_REGISTERS).describe()


def test_knob_length_and_replan_without_device():
    cmd = z.CommandBuffer()
    a = cmd.input(rgba8())
    c = cmd.color_convert(a, z.Color(primaries=z.Primaries.DisplayP3), z.Texel(10, 8))
    k = cmd.knob(c)
    cmd.output(c)
    e = plan(cmd)
    expect(z.Status.KNOB_LENGTH, e.knob, k, b"\0" * 4)
    ident = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    e.knob(k, bytes((ctypes.c_float * 12)(*ident)))
    assert "mat3" not in e.describe()  # identity folds away
    expect(z.Status.INVALID_ARGUMENT, e.knob, k + 7, b"\0" * 48)


def test_buffer_commands():
    """buffer_init / buffer_zero / from_buffer (command.rs:937-968,
    1777-1801): type checks, the 256-byte row requirement, and the plan."""
    cmd = z.CommandBuffer()
    a = cmd.input(rgba8(8, 8))
    luma = z.Descriptor(8, 8, z.Texel(z.SampleBits.UInt8, z.SampleParts.Luma), z.Color.SRGB)
    buf = cmd.buffer_init(b"\xff" * (8 * 256))
    assert cmd.buffer_size(buf) == 2048
    assert cmd.buffer_size(cmd.buffer_zero(17)) == 17
    # not a buffer / not an image
    expect(z.Status.TYPE, cmd.from_buffer, a, luma)
    expect(z.Status.TYPE, cmd.describe, buf)
    expect(z.Status.TYPE, cmd.color_convert, buf, z.Color.SRGB, z.Texel(10, 8))
    expect(z.Status.TYPE, cmd.buffer_size, a)
    # rows padded to 256 bytes: nine rows do not fit, eight do
    expect(z.Status.INVALID_ARGUMENT, cmd.from_buffer, buf, luma.with_(height=9))
    expect(z.Status.INVALID_ARGUMENT, cmd.from_buffer, buf, rgba8(65, 8))
    img = cmd.from_buffer(buf, luma)
    assert cmd.describe(img).row_stride in (0, 8)
    expect(z.Status.TYPE, cmd.knob, img)
    k = cmd.knob(buf)
    o, od = cmd.output(img)
    assert (od.width, od.height) == (8, 8)
    e = plan(cmd)
    assert "copy 64 bytes" in e.describe()
    expect(z.Status.KNOB_LENGTH, e.knob, k, b"\0" * 2047)
    e.knob(k, b"\0" * 2048)


def test_normal2d_parameters():
    """DistributionNormal2d::with_diagonal / with_direction
    (shaders/distribution_normal2d.rs) against the linear algebra they stand
    for: the inverse (pseudo-inverse) covariance and (2 pi)^k times its
    (pseudo-)determinant, in f32."""
    p = z.DistributionNormal2d.with_diagonal(0.2, 0.3).params()
    assert np.allclose(p, [0, 0, 5.0, 0, 0, 1 / 0.3, (2 * np.pi) ** 2 * 0.06, 0], rtol=1e-6)
    p = z.DistributionNormal2d.with_diagonal(0.0, 0.5).params()  # a zero variance drops its axis
    assert np.allclose(p, [0, 0, 0, 0, 0, 2.0, 2 * np.pi * 0.5, 0], rtol=1e-6)
    d = np.array([0.3, -0.3])
    p = z.DistributionNormal2d.with_direction(d).params()
    assert np.allclose(np.reshape(p[2:6], (2, 2)), np.linalg.pinv(np.outer(d, d)), rtol=1e-5)
    assert np.isclose(p[6], 2 * np.pi * (d @ d), rtol=1e-6)
    # |x| != |y|: the off-diagonal is x y / |d|^4; the diagonal is evaluated
    # as 1 / (4 x^2), 1 / (4 y^2), as the reference evaluates it
    x, y = 0.04998, 0.0501
    p = z.DistributionNormal2d.with_direction([x, y]).params()
    assert np.allclose(p[3:5], x * y / (x * x + y * y) ** 2, rtol=1e-5)
    assert np.allclose([p[2], p[5]], [1 / (4 * x * x), 1 / (4 * y * y)], rtol=1e-5)
    expect(z.Status.INVALID_ARGUMENT, z.DistributionNormal2d.with_direction, [float("inf"), 0.0])
    cmd = z.CommandBuffer()
    expect(z.Status.TYPE, cmd.distribution_normal2d, rgba8(8, 8), z.DistributionNormal2d.with_diagonal(1, 1))
    luma = z.Descriptor(8, 8, z.Texel(z.SampleBits.UInt8, z.SampleParts.Luma), z.Color.SRGB)
    g = cmd.distribution_normal2d(luma, z.DistributionNormal2d.with_diagonal(1, 1))
    cmd.output(g)
    assert "normal2d store" in plan(cmd).describe()


def test_fractal_noise_parameters():
    """shaders/fractal_noise.rs: with_octaves and set_damping, and the
    reference's own test_set_damping (8 octaves damped by 0.5: the
    amplitudes sum to 1)."""
    n = z.FractalNoise.with_octaves(4)
    assert (n.num_octaves, n.initial_amplitude, n.amplitude_damping, n.grid_scale) == (4, 0.25, 1.0, (100.0, 100.0))
    n = z.FractalNoise.with_octaves(8).set_damping(0.5)
    amp, total = n.initial_amplitude, 0.0
    for _ in range(8):
        total += amp
        amp *= n.amplitude_damping
    assert abs(1.0 - total) < 1e-6
    assert z.FractalNoise.with_octaves(3).set_damping(1.0).initial_amplitude == 1.0
    cmd = z.CommandBuffer()
    cmd.output(cmd.distribution_fractal_noise(rgba8(16, 16), z.FractalNoise.with_octaves(4)))
    assert "noise store" in plan(cmd).describe()
    cmd = z.CommandBuffer()
    cmd.output(cmd.distribution_fractal_noise(rgba8(16, 16), z.FractalNoise.with_octaves(5000)))
    expect(z.Status.UNSUPPORTED, plan, cmd)


def test_palette_command():
    """CommandBuffer::palette (command.rs:1442-1485): the result has the
    palette's texel and colour at the indices' size; channels other than R,
    G, B are type errors; the palette stage reads the materialised indices."""
    cmd = z.CommandBuffer()
    pal = cmd.input(z.Descriptor.rgba16f_linear(64, 8))
    idx = cmd.input(rgba8(30, 20))
    r = cmd.palette(pal, z.Palette(width=z.ColorChannel.R, height=z.ColorChannel.G), idx)
    d = cmd.describe(r)
    assert (d.width, d.height, d.texel.bits) == (30, 20, z.SampleBits.Float16x4)
    expect(z.Status.TYPE, cmd.palette, pal, z.Palette(width=z.ColorChannel.A), idx)
    cmd.output(r)
    text = plan(cmd).describe()
    assert "load#0 palette store#2" in text or "palette" in text
    # resize: a coordinate ramp of the target size, then the palette lookup
    cmd = z.CommandBuffer()
    rs = cmd.resize(cmd.input(rgba8(30, 20)), 61, 7)
    assert (cmd.describe(rs).width, cmd.describe(rs).height) == (61, 7)
    cmd.output(rs)
    text = plan(cmd).describe()
    assert "grid" in text and "palette" in text


def test_extract_inject_commands():
    """extract / inject (command.rs:1221-1265, 1360-1440): single-channel
    texels of the source's bits and colour; alpha and multi-channel images
    are refused; the swap of run_swap plans as channel and inject micro-ops."""
    cmd = z.CommandBuffer()
    i = cmd.input(rgba8(16, 8))
    r = cmd.extract(i, z.ColorChannel.R)
    d = cmd.describe(r)
    assert (d.texel.bits, d.texel.parts, d.width) == (z.SampleBits.UInt8, z.SampleParts.R, 16)
    expect(z.Status.OTHER, cmd.extract, i, z.ColorChannel.A)
    expect(z.Status.TYPE, cmd.inject, i, z.ColorChannel.G, i)
    h = cmd.input(z.Descriptor(16, 8, z.Texel(z.SampleBits.UInt16, z.SampleParts.G), z.Color.SRGB))
    expect(z.Status.TYPE, cmd.inject, i, z.ColorChannel.G, h)  # 16-bit channel into an 8-bit image
    g = cmd.extract(i, z.ColorChannel.G)
    s = cmd.inject(cmd.inject(i, z.ColorChannel.G, r), z.ColorChannel.R, g)
    cmd.output(s)
    text = plan(cmd).describe()
    assert "channel" in text and "inject" in text


def _over_p3(cmd, a, b):
    """A small function body: b converted to Display-P3 and blended over a."""
    p3 = cmd.color_convert(b, z.Color(primaries=z.Primaries.DisplayP3), z.Texel(10, 8))
    back = cmd.color_convert(p3, z.Color.SRGB, z.Texel(10, 8))
    return cmd.blend(a, z.Rectangle(0, 0, 16, 8), back)


def test_invoke_inlines_the_function():
    """CommandBuffer::function + invoke (command.rs:856-920) folded into the
    caller: two calls plan exactly as the same ops written inline; argument
    count and descriptors are checked."""
    body = z.CommandBuffer()
    ba, bb = body.input(rgba8(16, 8)), body.input(rgba8(16, 8))
    body.output(_over_p3(body, ba, bb))
    cmd = z.CommandBuffer()
    x, y = cmd.input(rgba8(16, 8)), cmd.input(rgba8(16, 8))
    (r1,) = cmd.invoke(body, [x, y])
    (r2,) = cmd.invoke(body, [r1, y])
    cmd.output(r2)
    ref = z.CommandBuffer()
    x2, y2 = ref.input(rgba8(16, 8)), ref.input(rgba8(16, 8))
    ref.output(_over_p3(ref, _over_p3(ref, x2, y2), y2))
    assert plan(cmd).describe() == plan(ref).describe()
    expect(z.Status.TYPE, cmd.invoke, body, [x])
    expect(z.Status.TYPE, cmd.invoke, body, [x, cmd.input(rgba8(8, 8))])
    expect(z.Status.TYPE, cmd.invoke, body, [x, y], 2)


def test_cube_root_sign_handling_only_where_needed():
    """The cube root skips its sign handling only where the planner proves the
    input non-negative: an 8-bit load through a matrix of non-negative
    entries (sRGB to Oklab's cone space).  An RGBA16F source may hold negative
    values, so its chain keeps the signed cube root."""
    def kernel(desc):
        cmd = z.CommandBuffer()
        lab = cmd.color_convert(cmd.input(desc), z.Color.OKLAB, z.Texel(z.SampleBits.Float16x4, z.SampleParts.LabA))
        cmd.output(lab)
        return re.findall(r"^stage .*\[kernel (\w+)", plan(cmd).describe(), re.M)
    assert kernel(rgba8(64, 8)) == ["oklab_enc_s8hp"]
    assert kernel(z.Descriptor.rgba16f_linear(64, 8)) != ["oklab_enc_s8hp"]


def test_launch_without_device_fails_loudly():
    e = plan(W.c2(8, 8))
    with pytest.raises(z.ZosimosError) as err:
        e.launch({}, {})
    assert err.value.code == z.Status.NO_DEVICE


def srgb_tables():
    import numpy as np
    buf = ctypes.create_string_buffer(6164)
    assert _abi.lib().zh_srgb_tables(buf, 6164) == 0
    raw = np.frombuffer(buf.raw, dtype=np.uint8)
    dec = raw[:1024].view(np.float32)
    thr = raw[1024:1024 + 1040].view(np.float32)
    idx = raw[2064:2064 + 4097]
    return dec, thr, idx


def table_encode(x, thr, idx):
    """The kernels' lookup: bucket code, then one threshold comparison."""
    import numpy as np
    c = np.minimum(np.maximum(np.nan_to_num(x.astype(np.float32), nan=0.0), np.float32(0)), np.float32(1))
    i = (c * np.float32(4096)).astype(np.int64)  # 0 .. 4096
    k = idx[i].astype(np.int64)
    return (k + (c >= thr[k + 1])).astype(np.uint8)


def test_srgb_tables_are_exact():
    """The table-based 8-bit sRGB encode equals round(255 * OETF(x)) in FP64
    for every tested f32 x: all thresholds +-3 ulps, all bucket edges +-1 ulp,
    and 4M random values; the decode table is the FP64 EOTF rounded once."""
    import numpy as np
    from oracle import color as C
    dec, thr, idx = srgb_tables()
    assert np.array_equal(dec, C.srgb_u8_decode_table().astype(np.float32))
    t = thr[1:256].view(np.uint32).astype(np.int64)
    near = np.concatenate([t + d for d in range(-3, 4)]).astype(np.uint32).view(np.float32)
    edges = (np.arange(4097, dtype=np.float32) / np.float32(4096)).view(np.uint32).astype(np.int64)
    edges = np.concatenate([edges + d for d in (-1, 0, 1)]).astype(np.uint32).view(np.float32)
    rnd = np.random.default_rng(0).random(4_000_000, dtype=np.float32)
    x = np.concatenate([near, edges, rnd, np.float32([0, 1, -1, 2, np.inf, -np.inf, np.nan])])
    want = C.encode_u8(np.stack([x, x, x, np.ones_like(x)], -1).astype(np.float64), 5)[:, 0]
    assert np.array_equal(table_encode(x, thr, idx), want)
    # thresholds are increasing and a 1/4096 bucket holds at most one of them
    assert np.all(np.diff(thr[1:256]) > 0)
    assert len(np.unique(np.floor(thr[1:256] * 4096))) == 255

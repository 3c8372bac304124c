"""User stage kernels on the host (no GPU): the HIP form of ShaderCommand
(command/dynamic.rs:7-51), construct/unary/binary_dynamic
(command.rs:2933-3048) and how the planner places them: one launch of the
user kernel over RGBA f32 copies of its operands, its result rounded through
the register's declared texel where it is read."""
import os

import numpy as np
import pytest

import zosimos_amd as z

SB, SP = z.SampleBits, z.SampleParts


def expect(code, fn, *args):
    with pytest.raises(z.ZosimosError) as e:
        fn(*args)
    assert e.value.code == code, e.value
    return e.value


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


LCH8_OKLAB = z.Descriptor(256, 256, z.Texel(SB.UInt8x4, SP.LchA), z.Color(model=z.ColorModel.Oklab))


def test_bundled_code_objects_are_built():
    for name in ("mandelbrot", "crt", "flat_field"):
        path = os.path.join(z.SHADER_DIR, name + ".co")
        assert os.path.getsize(path) > 1000
        with open(path, "rb") as f:
            assert f.read(24) == b"__CLANG_OFFLOAD_BUNDLE__"


def test_construct_dynamic_plans_one_user_launch():
    """custom.rs:76-90: the mandelbrot register, then an sRGB conversion."""
    c = z.CommandBuffer()
    b = c.construct_dynamic(Kernel("mandelbrot", LCH8_OKLAB, [3.0, 3.0, 0.6, 0.5]))
    assert c.describe(b) == LCH8_OKLAB
    o, od = c.output(c.color_convert(b, z.Color.SRGB, z.Texel(SB.UInt8x4, SP.RgbA)))
    text = z.Linker.from_included().compile(c).lower_to(None).describe()
    lines = text.splitlines()
    assert lines[0] == "stage 0: dynamic 256x256: 0 operand(s) [kernel dynamic:mandelbrot]"
    # the conversion reads the f32 result rounded through the LchA register
    assert lines[1].startswith("stage 1: pointwise 256x256 flat: load#0 quant#1")
    assert "intermediate 0: 256x256 bits 19" in text


def test_unary_and_binary_dynamic_copy_operands_to_f32():
    bg = z.Descriptor.rgba8_srgb(64, 48)
    big = bg.with_(width=192, height=144)
    c = z.CommandBuffer()
    i = c.input(bg)
    r = c.unary_dynamic(i, Kernel("crt", big, [192, 144], np.uint32))
    assert c.describe(r) == big
    c.output(r)
    text = z.Linker.from_included().compile(c).lower_to(None).describe()
    assert "stage 0: pointwise 64x48 flat: load#0 store#1" in text  # operand -> RGBA f32
    assert "stage 1: dynamic 192x144: 1 operand(s) [kernel dynamic:crt]" in text
    assert "stage 2: pointwise 192x144 flat: load#0 store#1" in text  # encode into the output

    flat = z.Descriptor(64, 48, z.Texel(SB.UInt8, SP.Luma), z.Color.SRGB)
    c2 = z.CommandBuffer()
    i2 = c2.input(bg)
    noise = c2.distribution_fractal_noise(flat, z.FractalNoise(8, 0.1, 0.2, (10.0, 10.0)))
    r2 = c2.binary_dynamic(i2, noise, Kernel("flat_field", bg, [0.062]))
    c2.output(r2)
    text2 = z.Linker.from_included().compile(c2).lower_to(None).describe()
    assert "dynamic 64x48: 2 operand(s) [kernel dynamic:flat_field]" in text2
    # the noise generator has a stage of its own: stored as FP16 by default,
    # in its declared Luma8 sRGB texel with strict registers
    assert "stage 0: pointwise 64x48: noise store#0" in text2 and "intermediate 0: 64x48 bits 18" in text2
    strict = z.Linker.from_included().compile(c2).lower_to(None, z.PLAN_STRICT_REGISTERS).describe()
    assert "intermediate 0: 64x48 bits 0\n" in strict


def test_dynamic_type_checks():
    c = z.CommandBuffer()
    k = Kernel("crt", z.Descriptor.rgba8_srgb(8, 8), [8, 8], np.uint32)
    buf = c.buffer_init(b"\0" * 64)
    # a buffer register is not a texture
    expect(z.Status.TYPE, c.unary_dynamic, buf, k)
    # a generic register has no concrete descriptor
    g = z.CommandBuffer()
    x = g.input_generic(g.generic())
    expect(z.Status.TYPE, g.unary_dynamic, x, k)
    # an invalid result descriptor
    bad = Kernel("crt", z.Descriptor(0, 8, z.Texel(SB.UInt8x4, SP.RgbA), z.Color.SRGB))
    i = c.input(z.Descriptor.rgba8_srgb(8, 8))
    expect(z.Status.TYPE, c.unary_dynamic, i, bad)
    # dynamic shaders take no knob (WithKnob has no dynamic ops)
    r = c.unary_dynamic(i, k)
    expect(z.Status.TYPE, c.knob, r)
    # an unknown register
    expect(z.Status.OTHER, c.unary_dynamic, 999, k)


def test_command_buffer_keeps_the_shader():
    """The command buffer and the linked program hold their own reference:
    the caller's source may go first."""
    import gc
    c = z.CommandBuffer()
    k = Kernel("mandelbrot", LCH8_OKLAB, [3.0, 3.0, 0.6, 0.5])
    r = c.construct_dynamic(k)
    c.output(r)
    del k
    gc.collect()
    prog = z.Linker.from_included().compile(c)
    del c
    gc.collect()
    assert "dynamic:mandelbrot" in prog.lower_to(None).describe()


def test_dynamic_inside_an_invoked_function():
    """A function body with a user kernel, called twice: two user launches."""
    d = z.Descriptor.rgba8_srgb(32, 32)
    body = z.CommandBuffer()
    x = body.input(d)
    body.output(body.unary_dynamic(x, Kernel("crt", d, [32, 32], np.uint32)))
    main = z.CommandBuffer()
    i = main.input(d)
    (a,) = main.invoke(body, [i])
    (b,) = main.invoke(body, [a])
    main.output(b)
    text = z.Linker.from_included().compile(main).lower_to(None).describe()
    assert text.count("[kernel dynamic:crt]") == 4  # two stage lines, two launch lines


def test_caller_built_dynamic_op_is_checked():
    """A zh_op list built by the caller (the -sys crate's path): the operand
    count must match the call, and the data must fit the blob."""
    import ctypes
    lib = z._abi.lib()
    sh = z.ShaderSource.bundled("crt")
    d = z.Descriptor.rgba8_srgb(8, 8)
    texs = (z._abi.DescriptorC * 2)(d.c(), d.c())

    class Call(ctypes.Structure):
        _fields_ = [("shader", ctypes.c_void_p), ("num_args", ctypes.c_uint32), ("data_len", ctypes.c_uint32)]

    def build(num_args, num_operands, data_len, blob_extra):
        blob = bytes(Call(sh._h.value, num_args, data_len)) + b"\0" * blob_extra
        buf = ctypes.create_string_buffer(blob, len(blob))
        ops = (z._abi.OpC * 3)()
        ops[0].high, ops[0].dst, ops[0].src, ops[0].knob = 0, 0, 0, -1
        ops[1].high, ops[1].dst, ops[1].shader, ops[1].knob = 2, 1, 21, -1
        ops[1].num_operands = num_operands
        ops[1].operands[0] = 0
        ops[1].params, ops[1].params_len = ctypes.cast(buf, ctypes.c_void_p), len(blob)
        ops[2].high, ops[2].dst, ops[2].src, ops[2].knob = 1, 0, 1, -1
        h = ctypes.c_void_p()
        return lib.zh_program_build(None, texs, 2, ops, 3, 0, ctypes.byref(h)), h

    code, h = build(1, 1, 8, 8)
    assert code == 0
    lib.zh_program_destroy(h)
    assert build(2, 1, 8, 8)[0] == z.Status.INVALID_ARGUMENT  # operand count
    assert build(1, 1, 8, 4)[0] == z.Status.INVALID_ARGUMENT  # data past the blob

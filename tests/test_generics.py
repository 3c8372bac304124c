"""Generic functions and the linker on the host (no GPU): generic
parameters, function symbols, invocation type checks, monomorphization at
link time (command.rs:753-923, 2083-2185) and knobs inside functions
(RegisterKnob, command.rs:2128-2140)."""
import ctypes

import pytest

import zosimos_amd as z

SB, SP = z.SampleBits, z.SampleParts
SCALARS8 = z.Descriptor(2048, 2048, z.Texel(SB.UInt8x4, SP.RgbA), z.Color(model=z.ColorModel.Scalars,
                                                                          transfer=z.Transfer.Linear))


def expect(code, fn, *args, **kw):
    with pytest.raises(z.ZosimosError) as e:
        fn(*args, **kw)
    assert e.value.code == code, e.value
    return e.value


def generic_palette():
    """The function of tests/generic.rs: its input's descriptor is a generic
    parameter; the input is sampled as a palette at a 2048x2048 ramp."""
    f = z.CommandBuffer()
    a = f.generic(z.GenericDeclaration())
    img = f.input_generic(a)
    idx = f.bilinear(SCALARS8, z.Bilinear(uv_max=(1.0, 1.0, 0.0, 0.0)))
    pal = f.palette(img, z.Palette(height=z.ColorChannel.R, width=z.ColorChannel.G), idx)
    f.output(pal)
    return f


def caller(f, descs):
    main = z.CommandBuffer()
    conv = main.function(f.computed_signature())
    regs = []
    for d in descs:
        i = main.input(d)
        (r,) = main.invoke(conv, z.InvocationArguments(generics=(main.register_descriptor(i),), arguments=(i,)))
        o, od = main.output(r)
        regs.append((i, o, od))
    return main, regs


def test_generic_palette_links_and_monomorphizes():
    f = generic_palette()
    sig = f.computed_signature()
    assert (sig.vars, sig.inputs, sig.outputs) == (1, 1, 1)
    bg = z.Descriptor.rgba8_srgb(512, 512)
    small = z.Descriptor(300, 200, z.Texel(SB.Float16x4, SP.RgbA), z.Color(transfer=z.Transfer.Linear))
    main, regs = caller(f, [bg, small])
    # the function is not linked: compiling the main buffer alone fails
    expect(z.Status.OTHER, z.Linker.from_included().compile, main)
    prog = z.Linker().link(main, [], [f], [[1], []])
    tex, ops = prog.high()
    # two instantiations, each a bilinear ramp and a palette draw
    draws = [o for o in ops if o.high == 2]
    assert [o.shader for o in draws] == [5, 11, 5, 11]
    for (i, o, od), pal_src in zip(regs, (bg, small)):
        out_tex = tex[[op.src for op in ops if op.high == 1][regs.index((i, o, od))]]
        # GenericDescriptor { chroma: the palette's, ..the indices' } (command.rs:1466-1470)
        assert (out_tex.width, out_tex.height) == (2048, 2048)
        assert (out_tex.texel.bits, out_tex.texel.parts) == (pal_src.texel.bits, pal_src.texel.parts)
        assert out_tex.color.transfer == pal_src.color.transfer
        assert (od.width, od.height) == (2048, 2048)
        assert prog.input_slot(i) in (0, 1) and prog.output_slot(o) in (0, 1)
    # both calls plan (the backend sees no call and no generic descriptor)
    assert prog.lower_to(None).describe().count("stage ") >= 2


def test_invocation_type_checks():
    f = generic_palette()
    main = z.CommandBuffer()
    conv = main.function(f.computed_signature())
    i = main.input(z.Descriptor.rgba8_srgb(64, 64))
    j = main.input(z.Descriptor.rgba8_srgb(32, 64))
    ti, tj = main.register_descriptor(i), main.register_descriptor(j)
    # arity (CommandError::INVALID_CALL)
    expect(z.Status.OTHER, main.invoke, conv, z.InvocationArguments(generics=(ti,), arguments=(i, j)))
    expect(z.Status.OTHER, main.invoke, conv, z.InvocationArguments(generics=(), arguments=(i,)))
    # the argument must equal the parameter rewritten with the generics
    expect(z.Status.OTHER, main.invoke, conv, z.InvocationArguments(generics=(tj,), arguments=(i,)))
    # a derived descriptor of another size does not match either
    tk = main.derive_descriptor(ti, (65, 64))
    expect(z.Status.OTHER, main.invoke, conv, z.InvocationArguments(generics=(tk,), arguments=(i,)))
    # the same size again does
    tl = main.derive_descriptor(ti, (64, 64))
    (r,) = main.invoke(conv, z.InvocationArguments(generics=(tl,), arguments=(i,)))
    assert main.describe(r).width == 2048
    # a generic register needs a concrete descriptor for a colour conversion
    g = z.CommandBuffer()
    v = g.generic()
    x = g.input_generic(v)
    expect(z.Status.TYPE, g.color_convert, x, z.Color.SRGB, z.Texel(SB.UInt8x4, SP.RgbA))
    expect(z.Status.TYPE, g.describe, x)
    # crop keeps a generic descriptor; outputs may be generic
    g.output(g.crop(x, z.Rectangle(0, 0, 4, 4)))


def test_link_checks_signature_and_arity():
    f = generic_palette()
    main, regs = caller(f, [z.Descriptor.rgba8_srgb(64, 64)])
    other = z.CommandBuffer()
    other.output(other.input(z.Descriptor.rgba8_srgb(64, 64)))
    # a definition of another signature is refused
    expect(z.Status.OTHER, z.Linker().link, main, [], [other], [[1], []])
    # links must name every symbol of a buffer
    expect(z.Status.OTHER, z.Linker().link, main, [], [f], [[1, 1], []])
    # a generic main buffer needs its descriptors
    expect(z.Status.OTHER, z.Linker().link, f, [], [], [])
    prog = z.Linker().link(f, [z.Descriptor.rgba8_srgb(16, 16)], [], [])
    assert prog.lower_to(None).describe().count("stage ") >= 1


def test_nested_generic_calls_and_knobs():
    """A generic function calling another generic function with its own
    parameter; a knob inside the inner body is reported per link index and
    keeps its length rule."""
    inner = z.CommandBuffer()
    v = inner.generic()
    x = inner.input_generic(v)
    grid = inner.bilinear(SCALARS8.with_(width=64, height=32), z.Bilinear(u_max=(1.0, 0.0, 0.0, 1.0)))
    kreg = grid
    kid_inner = inner.knob(kreg)
    inner.output(inner.palette(x, z.Palette(width=z.ColorChannel.R), grid))

    outer = z.CommandBuffer()
    w = outer.generic()
    y = outer.input_generic(w)
    fn = outer.function(inner.computed_signature())
    (r,) = outer.invoke(fn, z.InvocationArguments(generics=(w,), arguments=(y,)))
    outer.output(r)

    main = z.CommandBuffer()
    fo = main.function(outer.computed_signature())
    i = main.input(z.Descriptor.rgba8_srgb(40, 40))
    solid = main.solid_rgba(z.Descriptor.rgba8_srgb(8, 8), (0.5, 0.5, 0.5, 1.0))
    kid_main = main.knob(solid)
    (res,) = main.invoke(fo, z.InvocationArguments(generics=(main.register_descriptor(i),), arguments=(i,)))
    main.output(res)
    main.output(solid)
    prog = z.Linker().link(main, [], [outer, inner], [[1], [2], []])
    # main's knob keeps its id; the inner body's knob gets a fresh one
    assert prog.query_knob(0, solid) == kid_main
    k = prog.query_knob(2, kreg)
    assert k != kid_main and kid_inner == 0
    expect(z.Status.INVALID_ARGUMENT, prog.query_knob, 1, kreg)
    exe = prog.lower_to(None)
    expect(z.Status.KNOB_LENGTH, exe.knob, k, b"\0" * 4)
    exe.knob(k, bytes((ctypes.c_float * 24)(*([0.0] * 4 + [0.5, 0.0, 0.0, 1.0] + [0.0] * 16))))
    tex, ops = prog.high()
    assert [o.shader for o in ops if o.high == 2].count(11) == 1  # one palette draw


def test_inline_invoke_compares_the_whole_descriptor():
    body = z.CommandBuffer()
    body.output(body.input(z.Descriptor.rgba8_srgb(8, 8)))
    main = z.CommandBuffer()
    other_white = main.input(z.Descriptor(8, 8, z.Texel(SB.UInt8x4, SP.RgbA),
                                          z.Color(whitepoint=z.Whitepoint.D50)))
    expect(z.Status.TYPE, main.invoke, body, [other_white])
    ok = main.input(z.Descriptor.rgba8_srgb(8, 8))
    assert len(main.invoke(body, [ok])) == 1
    expect(z.Status.TYPE, main.invoke, generic_palette(), [ok])

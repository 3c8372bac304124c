"""The image pool on the host (no GPU): keys and their data
(pool.rs:19-156, 244-291), and the checks a launch from the pool makes
before anything moves (Executable::check_satisfiable, run.rs:1085-1167)."""
import numpy as np
import pytest

import zosimos_amd as z


def expect(code, fn, *args):
    with pytest.raises(z.ZosimosError) as e:
        fn(*args)
    assert e.value.code == code, e.value
    return e.value


def test_keys_and_host_bytes():
    pool = z.Pool()
    d = z.Descriptor.rgba8_srgb(5, 3)
    data = np.arange(5 * 3 * 4, dtype=np.uint8)
    k = pool.insert(d, data)
    assert k != 0
    got, kind = pool.info(k)
    assert kind == "host" and (got.width, got.height) == (5, 3)
    assert np.array_equal(pool.as_bytes(k), data)
    late = pool.declare(z.Descriptor.rgba16f_linear(7, 2))
    assert late not in (0, k) and pool.kind(late) == "late bound"
    expect(z.Status.INVALID_ARGUMENT, pool.as_bytes, late)
    expect(z.Status.INVALID_ARGUMENT, pool.device_ptr, k)
    # download of host bytes is a no-op; a late-bound image has nothing to move
    pool.download(k)
    expect(z.Status.INVALID_ARGUMENT, pool.download, late)
    pool.set_no_read(k)
    pool.remove(k)
    expect(z.Status.INVALID_ARGUMENT, pool.info, k)
    expect(z.Status.INVALID_ARGUMENT, pool.remove, k)
    with pytest.raises(ValueError):
        pool.insert(d, data[:-1])
    assert pool.cached_buffers() == 0


def test_insert_with_row_stride_keeps_tight_rows():
    """Rows given with a pitch are stored tight (ImageBuffer rows)."""
    import ctypes
    pool = z.Pool()
    d = z.Descriptor.rgba8_srgb(3, 2)
    pitched = np.zeros((2, 20), dtype=np.uint8)
    pitched[:, :12] = np.arange(24, dtype=np.uint8).reshape(2, 12)
    k = ctypes.c_uint64()
    z._check(z._abi.lib().zh_pool_insert_host(pool._h, ctypes.byref(d.c()), pitched.ctypes.data, 20,
                                              ctypes.byref(k)))
    assert np.array_equal(pool.as_bytes(k.value), np.arange(24, dtype=np.uint8))
    # a pitch shorter than a row is refused
    code = z._abi.lib().zh_pool_insert_host(pool._h, ctypes.byref(d.c()), pitched.ctypes.data, 8, ctypes.byref(k))
    assert code == z.Status.INVALID_ARGUMENT


def inscribe_program():
    bg, fg = z.Descriptor.rgba8_srgb(16, 16), z.Descriptor.rgba8_srgb(8, 8)
    cmd = z.CommandBuffer()
    b, f = cmd.input(bg), cmd.input(fg)
    o, od = cmd.output(cmd.inscribe(b, z.Rectangle(0, 0, 8, 8), f))
    return z.Linker.from_included().compile(cmd).lower_to(None), (b, f, o), (bg, fg, od)


def test_launch_checks_before_anything_moves():
    exe, (b, f, o), (bg, fg, od) = inscribe_program()
    pool = z.Pool()
    kb = pool.insert(bg, np.zeros(bg.nbytes, np.uint8))
    kf = pool.insert(fg, np.zeros(fg.nbytes, np.uint8))
    wrong = pool.insert(z.Descriptor.rgba8_srgb(8, 9), np.zeros(8 * 9 * 4, np.uint8))
    late = pool.declare(fg)
    # an unbound input (StartError::MissingKey)
    expect(z.Status.MISSING_KEY, exe.launch_pool, pool, {b: kb}, {o: 0})
    # an input whose image holds only a descriptor
    expect(z.Status.MISSING_KEY, exe.launch_pool, pool, {b: kb, f: late}, {o: 0})
    # a descriptor other than the register's (StartError::MismatchedDescriptor)
    expect(z.Status.MISMATCHED_DESCRIPTOR, exe.launch_pool, pool, {b: kb, f: wrong}, {o: 0})
    expect(z.Status.MISMATCHED_DESCRIPTOR, exe.launch_pool, pool, {b: kb, f: kf}, {o: kf})
    # the whole descriptor is compared, colour included (run.rs:1099-1106)
    d50 = pool.insert(fg.with_(color=z.Color(whitepoint=z.Whitepoint.D50)), np.zeros(fg.nbytes, np.uint8))
    expect(z.Status.MISMATCHED_DESCRIPTOR, exe.launch_pool, pool, {b: kb, f: d50}, {o: 0})
    # an input also written, an unknown key
    same = pool.insert(bg, np.zeros(bg.nbytes, np.uint8))
    expect(z.Status.INVALID_ARGUMENT, exe.launch_pool, pool, {b: kb, f: kf}, {o: kb})
    expect(z.Status.INVALID_ARGUMENT, exe.launch_pool, pool, {b: 12345, f: kf}, {o: 0})
    # everything checks out: the program has no device to run on
    expect(z.Status.NO_DEVICE, exe.launch_pool, pool, {b: kb, f: kf}, {o: same})
    expect(z.Status.NO_DEVICE, exe.launch_pool, pool, {b: kb, f: kf}, {o: pool.declare(od)})
    # nothing moved
    assert [pool.kind(k) for k in (kb, kf, late)] == ["host", "host", "late bound"]


def test_one_image_is_not_bound_twice():
    d = z.Descriptor.rgba8_srgb(8, 8)
    cmd = z.CommandBuffer()
    b, f = cmd.input(d), cmd.input(d)
    o, _ = cmd.output(cmd.inscribe(b, z.Rectangle(0, 0, 8, 8), f))
    exe = z.Linker.from_included().compile(cmd).lower_to(None)
    pool = z.Pool()
    k, k2 = pool.insert(d, np.zeros(d.nbytes, np.uint8)), pool.insert(d, np.zeros(d.nbytes, np.uint8))
    expect(z.Status.INVALID_ARGUMENT, exe.launch_pool, pool, {b: k, f: k}, {o: 0})
    expect(z.Status.NO_DEVICE, exe.launch_pool, pool, {b: k, f: k2}, {o: 0})


def test_dot_names_every_image_and_stage():
    """Executable::dot (run.rs:490-1010): images as boxes, stages as
    ellipses, an edge per read and write."""
    exe, (b, f, o), _ = inscribe_program()
    dot = exe.dot()
    assert dot.startswith("digraph program {") and dot.rstrip().endswith("}")
    # the inscribe of two RGBA8 inputs: one resample stage reading both
    assert "in0 -> s0;" in dot and "in1 -> s0;" in dot and "s0 -> out0;" in dot
    # a program with an intermediate: a shared value read by two outputs' stages
    d = z.Descriptor.rgba8_srgb(32, 32)
    c = z.CommandBuffer()
    i = c.input(d)
    lin = c.color_convert(i, z.Color(transfer=z.Transfer.Linear), z.Texel(z.SampleBits.Float16x4, z.SampleParts.RgbA))
    c.output(c.crop(lin, z.Rectangle(0, 0, 16, 16)))
    c.output(c.crop(lin, z.Rectangle(16, 16, 32, 32)))
    dot2 = z.Linker.from_included().compile(c).lower_to(None).dot()
    assert "s0 -> t0;" in dot2 and "t0 -> s1;" in dot2 and "t0 -> s2;" in dot2
    assert dot2.count("shape=ellipse") == 3


def test_stepped_start_needs_a_device():
    exe, (b, f, o), (bg, fg, od) = inscribe_program()
    pool = z.Pool()
    kb = pool.insert(bg, np.zeros(bg.nbytes, np.uint8))
    kf = pool.insert(fg, np.zeros(fg.nbytes, np.uint8))
    expect(z.Status.NO_DEVICE, exe.start_pool, pool, {b: kb, f: kf}, {o: 0})
    expect(z.Status.MISSING_KEY, exe.start_pool, pool, {b: kb}, {o: 0})


def test_environment_checks_before_launch():
    """Environment::bind (run.rs:1171-1205) refuses a pool image whose layout
    or texel differs from the register's and a register that is no input;
    bind_render wants a device image; knob data must have the knob's
    length (run.rs:1292-1298)."""
    d = z.Descriptor.rgba8_srgb(16, 8)
    cmd = z.CommandBuffer()
    i = cmd.input(d)
    c = cmd.color_convert(i, z.Color(primaries=z.Primaries.DisplayP3), d.texel)
    knob = cmd.knob(c)
    o, od = cmd.output(c)
    exe = z.Linker.from_included().compile(cmd).lower_to(None)
    pool = z.Pool()
    good = pool.insert(d, np.zeros(d.nbytes, np.uint8))
    taller = pool.insert(z.Descriptor.rgba8_srgb(16, 9), np.zeros(16 * 9 * 4, np.uint8))
    half = pool.declare(z.Descriptor.rgba16f_linear(16, 8))
    env = exe.from_pool(pool)
    env.bind(i, good)
    expect(z.Status.MISMATCHED_DESCRIPTOR, env.bind, i, taller)
    expect(z.Status.MISMATCHED_DESCRIPTOR, env.bind, i, half)
    expect(z.Status.INVALID_ARGUMENT, env.bind, o, good)  # an output register
    env.bind_output(o, pool.declare(od))
    expect(z.Status.INVALID_ARGUMENT, env.bind_render, o, good)  # host image
    assert len(exe.knob_data(knob)) == 48
    expect(z.Status.KNOB_LENGTH, env.knob, knob, b"\0" * 12)
    env.knob(knob, bytes(48))
    assert exe.knob_data(knob) != bytes(48)  # held for the run, not set yet

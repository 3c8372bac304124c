"""Launching from the image pool on the GPU: the reference's
``tests/loop.rs`` (one inscribe executable relaunched 200 times from a pool,
each output retired into the pool, the last one checked against the blessed
``composed`` hash), and the handoffs of pool.rs / run.rs: host bytes traded
in by upload, device images used in place and retired without a copy, an
output written into the pool's own device image, and a ``no_read`` input
moved out of the pool for the launch (pool.rs:900-941, run.rs:2791-2998).
Every result is compared byte for byte with the same program launched on
directly bound images."""
import json
import os

import numpy as np
import pytest

import zosimos_amd as z
from oracle import blockhash as B

from gpu_util import report

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HASHES = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))


def rgba(name):
    from PIL import Image
    return np.ascontiguousarray(np.asarray(Image.open(os.path.join(GOLDEN, name + ".png")).convert("RGBA")))


def inscribe(ctx, bg, fg):
    dbg = z.Descriptor.rgba8_srgb(bg.shape[1], bg.shape[0])
    dfg = z.Descriptor.rgba8_srgb(fg.shape[1], fg.shape[0])
    cmd = z.CommandBuffer()
    b, f = cmd.input(dbg), cmd.input(dfg)
    o, od = cmd.output(cmd.inscribe(b, z.Rectangle(0, 0, fg.shape[1], fg.shape[0]), f))
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    return exe, (b, f, o), (dbg, dfg, od)


def direct(ctx, exe, regs, descs, bg, fg):
    (b, f, o), (dbg, dfg, od) = regs, descs
    ib, iff, io = ctx.image(dbg), ctx.image(dfg), ctx.image(od)
    ib.upload(bg.reshape(-1))
    iff.upload(fg.reshape(-1))
    exe.launch({b: ib, f: iff}, {o: io})
    ctx.sync()
    return io.download().reshape(-1)


def test_loop_relaunch_from_pool_matches_reference_hash(ctx):
    bg, fg = rgba("background"), rgba("foreground")
    exe, (b, f, o), (dbg, dfg, od) = inscribe(ctx, bg, fg)
    want = direct(ctx, exe, (b, f, o), (dbg, dfg, od), bg, fg)
    pool = z.Pool()
    kb, kf = pool.insert(dbg, bg), pool.insert(dfg, fg)
    keys = set()
    for i in range(200):
        run = exe.launch_pool(pool, {b: kb, f: kf}, {o: 0})
        result = run.retire_output(o)
        run.finish()
        keys.add(result)
        # the two uploads and the output buffer return to the cache and are reused
        assert pool.cached_buffers() == 3, i
    assert len(keys) == 200 and pool.kind(result) == "host"
    got = pool.as_bytes(result)
    assert np.array_equal(got, want)
    h = B.blockhash256(got.reshape(od.height, od.width, 4))
    report("pool_loop_composed", got=h, iterations=200)
    assert h in HASHES["composed"]
    # the inputs stayed in the pool untouched
    assert np.array_equal(pool.as_bytes(kb), bg.reshape(-1)) and np.array_equal(pool.as_bytes(kf), fg.reshape(-1))
    pool.clear_cache()
    assert pool.cached_buffers() == 0


def test_device_handoffs(ctx):
    rng = np.random.default_rng(7)
    bg = rng.integers(0, 256, (96, 128, 4), dtype=np.uint8)
    fg = rng.integers(0, 256, (40, 56, 4), dtype=np.uint8)
    exe, (b, f, o), (dbg, dfg, od) = inscribe(ctx, bg, fg)
    want = direct(ctx, exe, (b, f, o), (dbg, dfg, od), bg, fg)
    pool = z.Pool()
    kb, kf = pool.insert(dbg, bg), pool.insert(dfg, fg)

    # retired as a device image: no copy, the buffer itself moves to the pool
    run = exe.launch_pool(pool, {b: kb, f: kf}, {o: 0})
    kdev = run.retire_output(o, to_host=False)
    run.finish()
    assert pool.kind(kdev) == "device"
    ptr = pool.device_ptr(kdev)

    # a device image used in place as the next launch's background
    cached = pool.cached_buffers()
    run = exe.launch_pool(pool, {b: kdev, f: kf}, {o: 0})
    k2 = run.retire_output(o)
    run.finish()
    assert pool.cached_buffers() == cached and pool.device_ptr(kdev) == ptr
    want2 = direct(ctx, exe, (b, f, o), (dbg, dfg, od), want.reshape(bg.shape), fg)
    assert np.array_equal(pool.as_bytes(k2), want2)

    # an output written straight into the pool's own device image
    own = ctx.image(od)
    kown = pool.insert_device(own)
    optr = pool.device_ptr(kown)
    run = exe.launch_pool(pool, {b: kb, f: kf}, {o: kown})
    assert run.retire_output(o, to_host=False) == kown
    run.finish()
    assert pool.device_ptr(kown) == optr
    pool.download(kown)
    assert pool.kind(kown) == "host" and np.array_equal(pool.as_bytes(kown), want)

    # an output retired into a declared (late-bound) image of its descriptor
    late = pool.declare(od)
    run = exe.launch_pool(pool, {b: kb, f: kf}, {o: late})
    assert run.retire_output(o) == late
    run.finish()
    assert np.array_equal(pool.as_bytes(late), want)

    # host bytes moved to the device ahead of the launch
    pool.upload(kf, ctx)
    assert pool.kind(kf) == "device"
    run = exe.launch_pool(pool, {b: kb, f: kf}, {o: 0})
    k3 = run.retire_output(o)
    run.finish()
    assert np.array_equal(pool.as_bytes(k3), want)
    report("pool_device_handoffs", bytes=want.size, equal=True)


def test_no_read_input_moves_out_for_the_launch(ctx):
    rng = np.random.default_rng(11)
    bg = rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)
    fg = rng.integers(0, 256, (32, 32, 4), dtype=np.uint8)
    exe, (b, f, o), (dbg, dfg, od) = inscribe(ctx, bg, fg)
    want = direct(ctx, exe, (b, f, o), (dbg, dfg, od), bg, fg)
    pool = z.Pool()
    kb, kf = pool.insert(dbg, bg), pool.insert(dfg, fg)
    pool.set_no_read(kf)
    run = exe.launch_pool(pool, {b: kb, f: kf}, {o: 0})
    # swapped out of the pool until retired (PoolImageMut::trade)
    assert pool.kind(kf) == "late bound" and pool.kind(kb) == "host"
    assert run.retire_input(f) == kf
    assert pool.kind(kf) == "host"
    k = run.retire_output(o)
    run.finish()
    assert np.array_equal(pool.as_bytes(k), want)
    # not retired: finish gives it back
    run = exe.launch_pool(pool, {b: kb, f: kf}, {o: 0})
    assert pool.kind(kf) == "late bound"
    run.finish()
    assert pool.kind(kf) == "host" and np.array_equal(pool.as_bytes(kf), fg.reshape(-1))
    # a launch refused by its checks leaves everything in place
    with pytest.raises(z.ZosimosError):
        exe.launch_pool(pool, {b: kb}, {o: 0})
    assert pool.kind(kf) == "host"


def crt_chain(ctx, w=96, h=64):
    """Three launches: the operand copy, a user kernel, the encode."""
    class Crt(z.ShaderCommand):
        def source(self):
            return z.ShaderSource.bundled("crt")

        def data(self, d):
            d.set_data([3 * w, 3 * h], np.uint32)
            return z.Descriptor.rgba8_srgb(3 * w, 3 * h)

    d = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    i = cmd.input(d)
    o, od = cmd.output(cmd.unary_dynamic(i, Crt()))
    prog = z.Linker.from_included().compile(cmd)
    return prog, prog.lower_to(ctx), i, o, d, od


def test_stepped_execution(ctx):
    """Execution::step_to: nothing runs before the first step, each step
    queues one launch, retiring waits for the last; the result equals the
    launch stepped at once."""
    prog, exe, i, o, d, od = crt_chain(ctx)
    rng = np.random.default_rng(2)
    pool = z.Pool()
    k = pool.insert(d, rng.integers(0, 256, d.nbytes, dtype=np.uint8))
    run = exe.launch_pool(pool, {i: k}, {o: 0})
    want = pool.as_bytes(run.retire_output(o))
    run.finish()

    run = exe.start_pool(pool, {i: k}, {o: 0})
    assert run.is_running()
    with pytest.raises(z.ZosimosError):
        run.retire_output(o)  # launches left to step
    with pytest.raises(z.ZosimosError):
        exe.launch_pool(pool, {i: k}, {o: 0})  # the program is busy stepping
    steps = []
    while True:
        n = run.step_to(1).launches
        if n == 0:
            break
        steps.append(n)
    assert steps == [1, 1, 1]
    got = pool.as_bytes(run.retire_output(o))
    run.finish()
    assert np.array_equal(got, want)
    # a larger limit queues the rest at once
    run = exe.start_pool(pool, {i: k}, {o: 0})
    assert run.step_to(100).launches == 3 and run.step_to(5).launches == 0
    assert np.array_equal(pool.as_bytes(run.retire_output(o)), want)
    run.finish()
    report("pool_stepped", steps=len(steps))


def test_resources_used_and_reuse(ctx):
    """Execution::resources_used: the first launch allocates its upload and
    output buffers, the next one reuses them from the pool's cache; a pruned
    execution gives nothing back to the cache."""
    prog, exe, i, o, d, od = crt_chain(ctx, 64, 32)
    pool = z.Pool()
    k = pool.insert(d, np.full(d.nbytes, 7, np.uint8))
    run = exe.launch_pool(pool, {i: k}, {o: 0})
    u1 = run.resources_used()
    run.finish()
    assert u1["buffers_allocated"] == 2 and u1["buffers_reused"] == 0
    assert u1["buffer_mem"] == d.nbytes + od.nbytes and u1["program_mem"] > 0
    run = exe.launch_pool(pool, {i: k}, {o: 0})
    u2 = run.resources_used()
    run.prune()
    run.finish()
    assert u2["buffers_allocated"] == 0 and u2["buffers_reused"] == 2 and u2["buffer_reused"] == u1["buffer_mem"]
    assert pool.cached_buffers() == 0
    report("pool_resources", first=u1, second=u2)


def test_finish_by_discarding_and_launch_once(ctx):
    prog, exe, i, o, d, od = crt_chain(ctx, 32, 32)
    pool = z.Pool()
    data = np.arange(d.nbytes, dtype=np.uint32).astype(np.uint8)
    k = pool.insert(d, data)
    run = exe.launch_pool(pool, {i: k}, {o: 0})
    want = pool.as_bytes(run.retire_output(o))
    run.finish()
    pool.set_no_read(k)
    run = exe.launch_pool(pool, {i: k}, {o: 0})
    cached = pool.cached_buffers()
    run.finish_by_discarding()
    # the moved input is dropped with the execution; no buffer went back
    assert pool.kind(k) == "late bound" and pool.cached_buffers() == cached
    # launch_once consumes the executable
    once = prog.lower_to(ctx)
    src, out = ctx.image(d), ctx.image(od)
    src.upload(data)
    once.launch_once({i: src}, {o: out})
    assert once._h is None
    assert np.array_equal(out.download().reshape(-1), want)


def run_once_with_output(ctx, commands, pool, binds, output, knobs=()):
    """``tests/util.rs:61-125`` (``run_executable_with_output``) through the
    same names: from_pool, knob, bind, recover_buffers, launch, clear_cache,
    step until done, retire_gracefully, the output callback,
    retire_buffers, finish."""
    exe = z.Linker.from_included().compile(commands).lower_to(ctx)
    env = exe.from_pool(pool)
    for knob, data in knobs:
        env.knob(knob, data)
    for reg, key in binds:
        env.bind(reg, key)
    env.recover_buffers()
    execution = exe.launch(env)
    pool.clear_cache()
    marks = []
    while execution.is_running():
        point = execution.step()
        point.block_on()
        marks.append(point.debug_mark())
        assert point.time_spent() > 0
    retire = execution.retire_gracefully(pool)
    result = output(retire)
    retire.retire_buffers()
    retire.finish()
    return result, marks, exe


def test_util_run_sequence_matches_reference_hash(ctx):
    """The reference's test driver sequence on the inscribe program of
    tests/blend.rs: the retired output's block hash is the blessed
    ``composed`` hash, and each step's debug mark names a launch."""
    bg, fg = rgba("background"), rgba("foreground")
    dbg = z.Descriptor.rgba8_srgb(bg.shape[1], bg.shape[0])
    dfg = z.Descriptor.rgba8_srgb(fg.shape[1], fg.shape[0])
    cmd = z.CommandBuffer()
    b, f = cmd.input(dbg), cmd.input(dfg)
    o, od = cmd.output(cmd.inscribe(b, z.Rectangle(0, 0, fg.shape[1], fg.shape[0]), f))
    pool = z.Pool()
    kb, kf = pool.insert(dbg, bg), pool.insert(dfg, fg)
    key, marks, exe = run_once_with_output(ctx, cmd, pool, [(b, kb), (f, kf)], lambda r: r.output(o))
    assert marks and all(m and m.startswith("launch ") for m in marks), marks
    got = pool.as_bytes(key)
    h = B.blockhash256(got.reshape(od.height, od.width, 4))
    report("util_run_sequence", hash=h, steps=len(marks))
    assert h in HASHES["composed"]


def test_environment_checks_and_run_scoped_knobs(ctx):
    """Environment::bind rejects an image of another layout, knob data of the
    wrong length is refused, and a run's knob data holds for that run only
    (the next run without it uses the program's own data again);
    Retire::output_key is the bound key."""
    w, h = 48, 32
    d = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    i = cmd.input(d)
    c = cmd.color_convert(i, z.Color(primaries=z.Primaries.DisplayP3), d.texel)
    knob = cmd.knob(c)
    o, od = cmd.output(c)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    pool = z.Pool()
    rng = np.random.default_rng(31)
    data = rng.integers(0, 256, d.nbytes, dtype=np.uint8)
    k = pool.insert(d, data)
    wrong = pool.insert(z.Descriptor.rgba8_srgb(w + 1, h), np.zeros((w + 1) * h * 4, np.uint8))

    env = exe.from_pool(pool)
    with pytest.raises(z.ZosimosError):
        env.bind(i, wrong)
    with pytest.raises(z.ZosimosError) as e:
        env.knob(knob, b"\0" * 7)
    assert e.value.code == z.Status.KNOB_LENGTH

    def run(with_identity):
        out = pool.declare(od)
        env = exe.from_pool(pool)
        env.bind(i, k)
        env.bind_output(o, out)
        if with_identity:
            ident = np.zeros(12, np.float32)
            ident[[0, 5, 10]] = 1.0  # std140 mat3: three vec4 columns
            env.knob(knob, ident.tobytes())
        run = exe.launch(env)
        while run.is_running():
            run.step().block_on()
        retire = run.retire_gracefully(pool)
        assert retire.output_key(o) == out
        key = retire.output(o)
        retire.finish()
        assert key == out
        return pool.as_bytes(key)

    converted = run(False)
    copied = run(True)
    again = run(False)
    # the identity matrix turns the conversion into a byte copy
    assert np.array_equal(copied, data)
    assert np.array_equal(again, converted) and not np.array_equal(converted, data)

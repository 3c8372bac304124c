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

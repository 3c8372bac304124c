"""zh_program_launch_static (no device allocation: the plan's tables live in
the library's static scratch) against zh_program_build + zh_program_launch on
the same High programs, through the C ABI.  Outputs must be byte-identical
(the same kernels run; only where their tables live differs).

Three calls in a row, each smaller than the one before, so every call
overwrites tables the previous call left in the scratch: a batch of two
tile queues (resample stages with counters and tile-layer masks), a
program with a materialised intermediate, and a one-stage pointwise
program.  A program whose tables exceed the scratch is refused before
anything launches, and the next call works.
"""
import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import workloads as W

from gpu_util import report, run

pytestmark = pytest.mark.gpu


def _with_intermediate(w, h, seed):
    """A value read by two consumers is materialised (an intermediate)."""
    cmd = z.CommandBuffer()
    d8 = z.Descriptor.rgba8_srgb(w, h)
    dl = z.Descriptor(w, h, W.RGBA16F, W.P3_LINEAR)
    a = cmd.input(d8)
    l = cmd.input(dl)
    c = cmd.color_convert(a, W.P3_LINEAR, W.RGBA16F)
    o1, od1 = cmd.output(cmd.color_convert(c, W.P3_SRGB, W.RGBA8))
    o2, od2 = cmd.output(cmd.blend(c, z.Rectangle(0, 0, w, h), l))

    def make():
        rng = np.random.default_rng(seed)
        return {a: rng.integers(0, 256, (h, w * 4), dtype=np.uint8), l: W.random_rgba16f(rng, w, h)}

    return W.Workload("intermediate", cmd, {a: d8, l: dl}, {o1: od1, o2: od2}, make, w * h)


def _static_run(ctx, wl, host):
    prog = z.Linker.from_included().compile(wl.commands)
    ins = {}
    for r, d in wl.inputs.items():
        ins[r] = ctx.image(d)
        ins[r].upload(host[r])
    outs = {o: ctx.image(d) for o, d in wl.outputs.items()}
    for img in outs.values():
        img.upload(np.zeros((img.desc.height, img.desc.width * img.desc.texel.bytes), np.uint8))
    prog.launch_static(ctx, ins, outs)
    ctx.sync()
    return {o: img.download() for o, img in outs.items()}, prog


def test_static_launch_matches_program_launch(ctx):
    cases = [W.c3(n=12, seed=31, max_side=700), _with_intermediate(300, 200, 33), W.c2(160, 90, seed=32)]
    for wl in cases:
        host = wl.make_inputs()
        ref, exe = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
        got, _ = _static_run(ctx, wl, host)
        report(f"static_{wl.name}", launches=exe.stats()[0], outputs=len(ref),
               intermediates=exe.describe().count("intermediate "))
        for o in ref:
            assert np.array_equal(got[o], ref[o]), f"{wl.name}: output {o}"


def test_static_launch_refuses_tables_beyond_the_scratch(ctx):
    """A 2048x2048 RGBA32F intermediate (64 MiB) does not fit the 32 MiB
    scratch: ZH_ERR_OUT_OF_MEMORY, nothing launched; a call after it works."""
    w = h = 2048
    cmd = z.CommandBuffer()
    d = z.Descriptor.rgba32f_linear(w, h)
    a = cmd.input(d)
    c = cmd.color_convert(a, z.Color(primaries=z.Primaries.DisplayP3, transfer=z.Transfer.Linear), d.texel)
    o1, od1 = cmd.output(cmd.color_convert(c, z.Color.SRGB, W.RGBA8))
    o2, od2 = cmd.output(cmd.color_convert(c, W.P3_SRGB, W.RGBA8))
    prog = z.Linker.from_included().compile(cmd)
    exe = prog.lower_to(None)
    assert "intermediate 0: 2048x2048 bits 19" in exe.describe()
    img = ctx.image(d)
    outs = {o1: ctx.image(od1), o2: ctx.image(od2)}
    with pytest.raises(z.ZosimosError) as e:
        prog.launch_static(ctx, {a: img}, outs)
    assert e.value.code == z.Status.OUT_OF_MEMORY and "static scratch" in e.value.message
    wl = W.c2(64, 32, seed=34)
    host = wl.make_inputs()
    ref, _ = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    got, _ = _static_run(ctx, wl, host)
    for o in ref:
        assert np.array_equal(got[o], ref[o])

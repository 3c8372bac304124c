"""Kernels compiled at run time (zh_jit.cpp): a pointwise chain with no
registered specialised kernel -- C2's operations with the 'over' moved
first -- gets the specialised body compiled for its signature with hiprtc.
Its results are checked against the FP64 oracle and against the general
interpreter (ZH_PLAN_NO_JIT), alone and as a batch of ragged images run by
one launch, and inside a captured graph."""
import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import workloads as W
from oracle import pipelines as PL

from gpu_util import compare_u8, report, run

pytestmark = pytest.mark.gpu


def test_unregistered_chain_runs_a_compiled_kernel(ctx):
    w, h = 320, 180
    wl = W.c2_reordered(w, h, seed=4)
    host = wl.make_inputs()
    (o, od), = wl.outputs.items()
    out, exe = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    text = exe.describe()
    assert "[kernel jit#" in text and "unavailable" not in text, text
    report("jit_compile", lines=[l for l in text.splitlines() if l.startswith("jit#")])
    src_r, layer_r = list(wl.inputs)
    lin, want = PL.c2_reordered(PL.u8_image(host[src_r], w, h), PL.f16_image(host[layer_r], w, h))
    mr, ma, exact = compare_u8("jit_c2_reordered", out[o], want, alpha_ok=want[..., 3] > 0)
    assert mr <= 1 and ma <= 1 and exact > 0.99
    interp, exe2 = run(ctx, wl.commands, wl.inputs, host, wl.outputs, flags=z.PLAN_NO_JIT)
    assert "interp-inline" in exe2.describe()
    diff = np.abs(out[o].astype(int) - interp[o].astype(int))
    report("jit_vs_interpreter", max_code_diff=int(diff.max()), equal=float((diff == 0).mean()))
    assert diff.max() <= 1


def test_compiled_kernel_batch_of_ragged_images(ctx):
    """Five images of the signature: one launch of the compiled batch kernel
    (chunks of 16,384 pixels), each image equal to its interpreter result."""
    sizes = [(1, 1), (7, 5), (129, 127), (300, 61), (16385, 2)]
    cmd = z.CommandBuffer()
    regs, inputs, host = [], {}, {}
    rng = np.random.default_rng(9)
    for w, h in sizes:
        d8 = z.Descriptor.rgba8_srgb(w, h)
        dl = z.Descriptor(w, h, W.RGBA16F, W.P3_LINEAR)
        a, b = cmd.input(d8), cmd.input(dl)
        p3 = cmd.color_convert(a, W.P3_LINEAR, W.RGBA16F)
        over = cmd.blend(p3, z.Rectangle(0, 0, w, h), b)
        lab = cmd.oklab_adjust(cmd.color_convert(over, z.Color.OKLAB, W.LABA16F), chroma=1.2, hue_rad=0.0)
        o, od = cmd.output(cmd.color_convert(cmd.color_convert(lab, W.P3_LINEAR, W.RGBA16F), W.P3_SRGB, W.RGBA8))
        inputs[a], inputs[b] = d8, dl
        host[a] = W.random_rgba8(rng, w, h)
        host[b] = W.random_rgba16f(rng, w, h)
        regs.append((a, b, o, od, w, h))
    outputs = {o: od for (_, _, o, od, _, _) in regs}
    out, exe = run(ctx, cmd, inputs, host, outputs)
    text = exe.describe()
    assert "[kernel jit#" in text and "over 5 stages" in text, text
    interp, _ = run(ctx, cmd, inputs, host, outputs, flags=z.PLAN_NO_JIT)
    for (a, b, o, od, w, h) in regs:
        lin, want = PL.c2_reordered(PL.u8_image(host[a], w, h), PL.f16_image(host[b], w, h))
        mr, ma, exact = compare_u8(f"jit_batch_{w}x{h}", out[o], want, alpha_ok=want[..., 3] > 0)
        assert mr <= 1 and ma <= 1
        assert np.abs(out[o].astype(int) - interp[o].astype(int)).max() <= 1


def test_compiled_kernel_inside_a_graph(ctx):
    """The compiled kernel between two other launches, captured once and
    replayed: every replay equals the eager program."""
    w, h = 256, 64
    wl = W.c2_reordered(w, h, seed=6)
    cmd = wl.commands
    (o, od), = wl.outputs.items()
    extra, extra_d = cmd.output(cmd.crop(list(wl.inputs)[0], z.Rectangle(0, 0, 64, 32)))
    host = wl.make_inputs()
    outputs = dict(wl.outputs)
    outputs[extra] = extra_d
    prog = z.Linker.from_included().compile(cmd)
    exe, eager = prog.lower_to(ctx), prog.lower_to(ctx, z.PLAN_NO_GRAPH)
    ins = {r: ctx.image(d) for r, d in wl.inputs.items()}
    for r, img in ins.items():
        img.upload(host[r])
    outs = {r: ctx.image(d) for r, d in outputs.items()}
    refs = {r: ctx.image(d) for r, d in outputs.items()}
    for _ in range(3):
        exe.launch(ins, outs)
        eager.launch(ins, refs)
        ctx.sync()
        for r in outs:
            assert np.array_equal(outs[r].download(), refs[r].download())
    g = exe.graph_info()
    report("jit_graph", **g)
    assert g["state"] == "captured" and g["replays"] >= 2


def test_unregistered_warp_form_runs_a_compiled_kernel(ctx):
    """A rotated RGBA16F layer composited onto a transparent canvas and
    converted from Display-P3 to sRGB: a post program (un-premultiply, matrix,
    store) no registered warp kernel has.  The compiled warp kernel against
    the FP64 oracle and the resample interpreter, alone and as a batch of two
    images (one launch)."""
    from oracle import color as C
    rng = np.random.default_rng(12)
    cmd = z.CommandBuffer()
    regs, host = [], {}
    for w, h, ow, oh in [(200, 120, 260, 200), (97, 64, 150, 101)]:
        d = z.Descriptor(w, h, W.RGBA16F, W.P3_LINEAR)
        a = cmd.input(d)
        canvas = cmd.solid_rgba(z.Descriptor(ow, oh, W.RGBA16F, W.P3_LINEAR), (0, 0, 0, 0))
        aff = z.Affine.new(z.AffineSample.BiLinear).rotate(0.3).shift(40.0, 10.0)
        r = cmd.affine(canvas, aff, a)
        o, od = cmd.output(cmd.color_convert(r, z.Color.SRGB, W.RGBA8))
        host[a] = W.random_rgba16f(rng, w, h)
        regs.append((a, d, o, od, aff, w, h, ow, oh))
    inputs = {a: d for (a, d, *_rest) in regs}
    outputs = {o: od for (_a, _d, o, od, *_rest) in regs}
    out, exe = run(ctx, cmd, inputs, host, outputs)
    text = exe.describe()
    assert "[kernel jitw#" in text and "over 2 stages" in text and "unavailable" not in text, text
    interp, exe2 = run(ctx, cmd, inputs, host, outputs, flags=z.PLAN_NO_JIT)
    assert "jitw#" not in exe2.describe()
    m709 = C.convert_matrix(7, 3, 5)
    for (a, d, o, od, aff, w, h, ow, oh) in regs:
        src = PL.f16_image(host[a], w, h)
        res = C.unpremultiply(C.bilinear_affine(C.premultiply(src), PL.mat_f32(aff.transformation), (oh, ow)))
        want = C.encode_u8(PL.apply3(m709, res), 5)
        mr, ma, exact = compare_u8(f"jitw_{w}x{h}", out[o], want, alpha_ok=res[..., 3] > 0.02)
        assert mr <= 1 and ma <= 1 and exact > 0.99
        diff = np.abs(out[o].astype(int) - interp[o].astype(int))
        report(f"jitw_vs_interpreter_{w}x{h}", max_code_diff=int(diff.max()), equal=float((diff == 0).mean()))
        assert diff.max() <= 1

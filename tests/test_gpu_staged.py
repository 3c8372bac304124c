"""LDS-staged warp tiles on the GPU: planned with ZH_PLAN_STAGED, one-layer
separable stages whose per-tile source box fits run the staged kernel (each
tile's box fetched by LDS-direct loads, decoded and premultiplied once); the
same programs planned by default run the direct-gather kernel.  Both are held to the FP64
oracle's bars (8-bit within 1 code, RGBA16F within 1 ulp where alpha >
1e-3) and to each other, on magnified, unit-scale-with-offset, flipped and
partly covered maps with partial tiles, alone and as a batch."""
import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import workloads as W
from oracle import color as C
from oracle import pipelines as PL

from gpu_util import compare_u8, report, run

pytestmark = pytest.mark.gpu

P3_SRGB = z.Color(primaries=z.Primaries.DisplayP3, transfer=z.Transfer.Srgb)
BIL = z.AffineSample.BiLinear

# source w, h -> output ow, oh and the map
CASES = {
    "magnify": (150, 77, 257, 131, lambda w, h, ow, oh: z.Affine.new(BIL).scale(ow / w, oh / h)),
    "unit_offset": (203, 45, 203, 45, lambda w, h, ow, oh: z.Affine.new(BIL).shift(0.37, -0.21)),
    "flip": (120, 90, 180, 160, lambda w, h, ow, oh: z.Affine.new(BIL).scale(-ow / w, oh / h).shift(ow, 0)),
    "partly_covered": (140, 100, 300, 220, lambda w, h, ow, oh: z.Affine.new(BIL).scale(1.3, 1.25).shift(60.5, -30.25)),
}


def program(fmt, w, h, ow, oh, aff):
    if fmt == 0:
        din, conv, tex = z.Descriptor.rgba8_srgb(w, h), P3_SRGB, W.RGBA8
    else:
        din, conv, tex = z.Descriptor(w, h, W.RGBA16F, W.LINEAR_709), W.P3_LINEAR, W.RGBA16F
    cmd = z.CommandBuffer()
    r = cmd.input(din)
    c = cmd.color_convert(r, conv, tex)
    canvas = cmd.solid_rgba(z.Descriptor(ow, oh, tex, conv), (0.0, 0.0, 0.0, 0.0))
    o, od = cmd.output(cmd.affine(canvas, aff, c))
    return cmd, r, din, o, od


def source(fmt, w, h, seed):
    rng = np.random.default_rng([seed, 77])
    return W.random_rgba8(rng, w, h) if fmt == 0 else W.random_rgba16f(rng, w, h)


def oracle(data, fmt, w, h, ow, oh, aff):
    """decode -> Rec.709 to Display-P3 -> premultiply -> bilinear through the
    map (transparent border) -> un-premultiply -> encode, in FP64."""
    m3 = C.convert_matrix(3, 7, 5)
    if fmt == 0:
        src = C.premultiply(PL.apply3(m3, C.decode_u8(PL.u8_image(data, w, h), 5)))
    else:
        src = C.premultiply(PL.apply3(m3, np.asarray(data).view(np.float16).reshape(h, w, 4).astype(np.float64)))
    res = C.unpremultiply(C.bilinear_affine(src, PL.mat_f32(aff.transformation), (oh, ow)))
    return res, (C.encode_u8(res, 5) if fmt == 0 else C.encode_f16(res))


def check(name, fmt, got, want, ow, oh):
    if fmt == 0:
        mr, ma, _ = compare_u8(name, got, want, alpha_ok=want[..., 3] > 0)
        assert mr <= 1 and ma <= 1
    else:
        g = got.view(np.float16).reshape(oh, ow, 4)
        ulp = np.abs(g.view(np.uint16).astype(np.int32) - want.view(np.uint16).astype(np.int32))
        ok = np.broadcast_to(want[..., 3:4] > 1e-3, ulp.shape) | (np.arange(4) == 3)
        report(name, max_ulp=int(ulp[ok].max(initial=0)))
        assert ulp[ok].max(initial=0) <= 1


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("case", list(CASES))
def test_staged_tiles_match_direct_and_oracle(ctx, fmt, case):
    w, h, ow, oh, mk = CASES[case]
    aff = mk(w, h, ow, oh)
    cmd, r, din, o, od = program(fmt, w, h, ow, oh, aff)
    data = source(fmt, w, h, 5 + fmt)
    staged, exe = run(ctx, cmd, {r: din}, {r: data}, {o: od}, flags=z.PLAN_STAGED)
    assert "_lds]" in exe.describe(), exe.describe()
    direct, exe2 = run(ctx, cmd, {r: din}, {r: data}, {o: od})
    assert "_lds" not in exe2.describe()
    _, want = oracle(data, fmt, w, h, ow, oh, aff)
    check(f"staged_{case}_{fmt}", fmt, staged[o], want, ow, oh)
    check(f"direct_{case}_{fmt}", fmt, direct[o], want, ow, oh)
    if fmt == 0:
        d = np.abs(staged[o].astype(int) - direct[o].astype(int))
        report(f"staged_vs_direct_{case}_u8", max_code_diff=int(d.max()), equal=f"{float((d == 0).mean()):.6f}")
        assert d.max() <= 1


def test_staged_batch_with_autotune(ctx):
    """A ragged batch (scales 0.5-2x, both formats): staged and direct units
    side by side, autotuned (staged units only try tile shapes whose boxes
    fit), against the oracle and the unstaged program."""
    wl = W.c3(n=12, seed=9, max_side=280)
    host = wl.make_inputs()
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx, z.PLAN_STAGED)
    text = exe.describe()
    assert "_lds]" in text and "_sep]" in text, text
    imgs, outs = {}, {}
    for reg, d in wl.inputs.items():
        imgs[reg] = ctx.image(d)
        imgs[reg].upload(host[reg])
    outs = {reg: ctx.image(d) for reg, d in wl.outputs.items()}
    exe.autotune(imgs, outs, iterations=2)
    for _ in range(2):
        exe.launch(imgs, outs)
    ctx.sync()
    got = {reg: img.download() for reg, img in outs.items()}
    plain, _ = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    for k, ((ri, din), (ro, dout)) in enumerate(zip(wl.inputs.items(), wl.outputs.items())):
        w, h, fmt, s = wl.meta["shapes"][k]
        _, want = PL.c3_image(host[ri], fmt, w, h, dout.width, dout.height)
        check(f"staged_batch_img{k}", fmt, got[ro], want, dout.width, dout.height)
        if fmt == 0:
            assert np.abs(got[ro].astype(int) - plain[ro].astype(int)).max() <= 1


def test_paired_forms_equal_two_queues(ctx):
    """ZH_PLAN_PAIR: the RGBA8 and RGBA16F images of a ragged batch run as
    one grid that deals the two forms' tile queues alternately (a kernel
    compiled at run time for the pair); autotuned and relaunched, every
    output equals the default plan's two queues byte for byte (the same
    kernel bodies), and the plan has one launch."""
    wl = W.c3(n=14, seed=21, max_side=300)
    host = wl.make_inputs()
    plain, _ = run(ctx, wl.commands, wl.inputs, host, wl.outputs)
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx, z.PLAN_PAIR)
    text = exe.describe()
    line = [l for l in text.splitlines() if l.startswith("launch 0:")][0]
    assert "[kernel pair " in line and "w_s8m_s8_sep" in line and "w_hm_h_sep" in line, text
    assert exe.stats()[0] == 1
    imgs = {}
    for reg, d in wl.inputs.items():
        imgs[reg] = ctx.image(d)
        imgs[reg].upload(host[reg])
    outs = {reg: ctx.image(d) for reg, d in wl.outputs.items()}
    exe.autotune(imgs, outs, iterations=2)
    for _ in range(3):
        exe.launch(imgs, outs)
    ctx.sync()
    for reg, img in outs.items():
        assert np.array_equal(img.download(), plain[reg]), reg
    report("paired_forms", launches=exe.stats()[0], images=len(outs))

"""Every specialised pointwise kernel against the FP64 oracle, next to the
general interpreter on the same program.

Each case is a one-stage program whose micro-op sequence matches one
registered signature.  It runs three ways: as planned (the specialised
kernel: flat, 16-byte aligned images), with padded rows (a different
descriptor, so the planner falls back to the 2-D interpreter), and -- for the
4- and 8-byte formats -- with the same executable bound to images at an
address that is not 16-byte aligned (the runtime then launches the
interpreter and captures its graph again).  Sizes leave a partial group of
pixels at the end of the image.

Bars: 8-bit outputs within 1 code of the FP64 oracle (colour where the 8-bit
alpha is non-zero after 'over'); f16 outputs within 1 ulp or 1e-6 absolute;
f32 outputs within 1e-5 relative plus 1e-6 absolute.
"""
import math

import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import workloads as W
from oracle import color as C
from oracle import pipelines as PL

from gpu_util import report

pytestmark = pytest.mark.gpu

P3 = z.Primaries.DisplayP3
LAB = z.Texel(z.SampleBits.Float16x4, 18)


def desc_of(kind, w, h):
    return {
        "s8": z.Descriptor.rgba8_srgb(w, h),
        "l8": z.Descriptor(w, h, z.Texel(z.SampleBits.UInt8x4, z.SampleParts.RgbA), z.Color(transfer=z.Transfer.Linear)),
        "h": z.Descriptor.rgba16f_linear(w, h),
        "f": z.Descriptor.rgba32f_linear(w, h),
        "lab": z.Descriptor(w, h, LAB, z.Color.OKLAB),
    }[kind]


def random_bytes(kind, w, h, rng):
    n = w * h
    if kind in ("s8", "l8"):
        return rng.integers(0, 256, size=n * 4, dtype=np.uint8)
    v = rng.random((n, 4))
    if kind == "lab":
        v[:, 1:3] = (v[:, 1:3] - 0.5) * 0.6
    dt = np.float32 if kind == "f" else np.float16
    return v.astype(dt).view(np.uint8).reshape(-1)


def decode(kind, raw, w, h):
    if kind == "s8":
        return C.decode_u8(PL.u8_image(raw, w, h), 5)
    if kind == "l8":
        return C.decode_u8(PL.u8_image(raw, w, h), 4)
    if kind == "f":
        return PL.f32_image(raw, w, h)
    return PL.f16_image(raw, w, h)


def oklab_enc(v):
    out = np.array(v)
    out[..., :3] = C.rgb_to_oklab(v[..., :3], 3)
    return out


def oklab_dec(v):
    out = np.array(v)
    out[..., :3] = C.oklab_to_rgb(v[..., :3], 3, clamp=True)
    return out


def over(below, above):
    return C.unpremultiply(C.over(C.premultiply(above), C.premultiply(below)))


# name: (input kinds, output kind, program builder, FP64 function of the decoded inputs)
def _conv(dst_kind, color=None):
    def build(cmd, regs, w, h):
        d = desc_of(dst_kind, w, h)
        return cmd.color_convert(regs[0], color or d.color, d.texel)
    return build


MAT = C.convert_matrix(3, 7, 5)
CASES = {
    "cv_s8l8": (["s8"], "l8", _conv("l8"), lambda a: a),
    "cv_l8s8": (["l8"], "s8", _conv("s8"), lambda a: a),
    "cv_s8h": (["s8"], "h", _conv("h"), lambda a: a),
    "cv_hs8": (["h"], "s8", _conv("s8"), lambda a: a),
    "cv_hf": (["h"], "f", _conv("f"), lambda a: a),
    "cv_fh": (["f"], "h", _conv("h"), lambda a: a),
    "cv_s8f": (["s8"], "f", _conv("f"), lambda a: a),
    "cv_fs8": (["f"], "s8", _conv("s8"), lambda a: a),
    "mat_s8s8": (["s8"], "s8", _conv("s8", z.Color(primaries=P3)), lambda a: PL.apply3(MAT, a)),
    "mat_hh": (["h"], "h", _conv("h", z.Color(primaries=P3, transfer=z.Transfer.Linear)), lambda a: PL.apply3(MAT, a)),
    "mat_s8h": (["s8"], "h", _conv("h", z.Color(primaries=P3, transfer=z.Transfer.Linear)),
                lambda a: PL.apply3(MAT, a)),
    "mat_hs8": (["h"], "s8", _conv("s8", z.Color(primaries=P3)), lambda a: PL.apply3(MAT, a)),
    "mat_ff": (["f"], "f", _conv("f", z.Color(primaries=P3, transfer=z.Transfer.Linear)), lambda a: PL.apply3(MAT, a)),
    "oklab_enc_s8h": (["s8"], "lab", lambda cmd, r, w, h: cmd.color_convert(r[0], z.Color.OKLAB, LAB), oklab_enc),
    "oklab_dec_hs8": (["lab"], "s8", lambda cmd, r, w, h: cmd.color_convert(r[0], z.Color.SRGB, desc_of("s8", w, h).texel),
                      oklab_dec),
    "oklab_rt_s8": (["s8"], "s8", lambda cmd, r, w, h: W.color_chain(cmd, r[0]), PL.chain),
    "over_s8": (["s8", "s8"], "s8", lambda cmd, r, w, h: cmd.blend(r[0], z.Rectangle(0, 0, w, h), r[1]), over),
    "over_h": (["h", "h"], "h", lambda cmd, r, w, h: cmd.blend(r[0], z.Rectangle(0, 0, w, h), r[1]), over),
}


def build_case(name, w, h, row_pad=0):
    kinds, out_kind, builder, _ = CASES[name]
    cmd = z.CommandBuffer()
    descs = []
    for k in kinds:
        d = desc_of(k, w, h)
        if row_pad:
            d = d.with_(row_stride=d.stride + row_pad)
        descs.append(d)
    regs = [cmd.input(d) for d in descs]
    o, od = cmd.output(builder(cmd, regs, w, h))
    return cmd, dict(zip(regs, descs)), o, od


def check(name, got_raw, want, out_kind, w, h, alpha_mask=None):
    if out_kind in ("s8", "l8"):
        code = C.encode_u8(want, 5 if out_kind == "s8" else 4).astype(np.int32)
        got = PL.u8_image(got_raw, w, h).astype(np.int32)
        d = np.abs(got - code)
        rgb = d[..., :3][alpha_mask] if alpha_mask is not None else d[..., :3]
        mr, ma = int(rgb.max(initial=0)), int(d[..., 3].max())
        report(name, max_rgb=mr, max_alpha=ma, exact=f"{float(np.mean(d == 0)):.6f}")
        assert mr <= 1 and ma <= 1
    elif out_kind in ("h", "lab"):
        exp = C.encode_f16(want)
        got = np.asarray(got_raw).view(np.float16).reshape(h, w, 4)
        ulp = np.abs(got.view(np.int16).astype(np.int32) - exp.view(np.int16).astype(np.int32))
        # Near zero (Oklab a, b of greys: a difference of nearly equal terms)
        # the f32 arithmetic is good to ~1e-7 absolute, several f16 ulps there.
        ulp = np.where(np.abs(got.astype(np.float64) - want) <= 1e-6, 0, ulp)
        if alpha_mask is not None:
            ulp = np.where(alpha_mask[..., None] | (np.arange(4) == 3), ulp, 0)
        report(name, max_ulp=int(ulp.max()))
        assert ulp.max() <= 1
    else:
        got = PL.f32_image(got_raw, w, h)
        err = np.abs(got - want) / (np.abs(want) + 1e-1)
        bad = np.abs(got - want) > 1e-5 * np.abs(want) + 1e-6
        report(name, max_rel=f"{float(err.max()):.2e}")
        assert not bad.any()


def launch(ctx, exe, inputs, o, od):
    out = ctx.image(od)
    exe.launch(inputs, {o: out})
    ctx.sync()
    return out.download().reshape(-1)


def expected(name, host, w, h):
    kinds, out_kind, _, fn = CASES[name]
    vals = [decode(k, host[i], w, h) for i, k in enumerate(kinds)]
    want = fn(*vals)
    mask = None
    if name.startswith("over_"):
        mask = want[..., 3] > (0.5 / 255.0 if out_kind == "s8" else 0.0)
    return want, mask


@pytest.mark.parametrize("w,h", [(37, 5), (301, 77)])
@pytest.mark.parametrize("name", list(CASES))
def test_specialised_kernel_matches_oracle(ctx, name, w, h):
    kinds, out_kind, _, _ = CASES[name]
    rng = np.random.default_rng([w, h, len(name)])
    host = [random_bytes(k, w, h, rng) for k in kinds]
    want, mask = expected(name, host, w, h)

    cmd, ins, o, od = build_case(name, w, h)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    # 8-bit sources reach the cube root through a non-negative matrix: the
    # planner picks the variant without sign handling (suffix p).
    text = exe.describe()
    assert f"[kernel {name}]" in text or f"[kernel {name}p]" in text or f"[kernel {name}_p]" in text
    imgs = {}
    for (r, d), raw in zip(ins.items(), host):
        imgs[r] = ctx.image(d)
        imgs[r].upload(raw)
    got = launch(ctx, exe, imgs, o, od)
    check(f"{name}_{w}x{h}_spec", got, want, out_kind, w, h, mask)

    # Same executable, images at an address that is not 16-byte aligned: the
    # interpreter runs (the graph is captured again), then the aligned images
    # once more.
    bpp = {"s8": 4, "l8": 4, "h": 8, "lab": 8, "f": 16}
    if all(bpp[k] < 16 for k in kinds + [out_kind]):
        keep, mis = [], {}
        for (r, d), raw, k in zip(ins.items(), host, kinds):
            big = ctx.image(z.Descriptor.rgba8_srgb(w * h * bpp[k] // 4 + 4, 1))
            keep.append(big)
            img = ctx.wrap(d, big.device_ptr + bpp[k], w * h * bpp[k])
            img.upload(raw)
            mis[r] = img
        big_o = ctx.image(z.Descriptor.rgba8_srgb(w * h * bpp[out_kind] // 4 + 4, 1))
        out = ctx.wrap(od, big_o.device_ptr + bpp[out_kind], w * h * bpp[out_kind])
        exe.launch(mis, {o: out})
        ctx.sync()
        check(f"{name}_{w}x{h}_unaligned", out.download().reshape(-1), want, out_kind, w, h, mask)
        got2 = launch(ctx, exe, imgs, o, od)
        assert np.array_equal(got2, got)
        for img in list(mis.values()) + [out]:
            img.free()
        del keep, big_o

    # Padded rows: the planner takes the 2-D interpreter.
    cmd, ins, o, od = build_case(name, w, h, row_pad=16)
    exe = z.Linker.from_included().compile(cmd).lower_to(ctx)
    assert "[kernel" not in exe.describe()
    imgs = {}
    for (r, d), raw in zip(ins.items(), host):
        imgs[r] = ctx.image(d)
        imgs[r].upload(raw)
    check(f"{name}_{w}x{h}_interp", launch(ctx, exe, imgs, o, od), want, out_kind, w, h, mask)


def test_c2_kernel_multi_step_grid(ctx):
    """C2 at 2049 x 1101 (more groups than one pass of the persistent grid,
    a partial group at the end) against the FP64 oracle."""
    w, h = 2049, 1101
    wl = W.c2(w, h, seed=5)
    host = wl.make_inputs()
    (o, od), = wl.outputs.items()
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx)
    assert "[kernel c2_over_p]" in exe.describe()
    imgs = {}
    for r, d in wl.inputs.items():
        imgs[r] = ctx.image(d)
        imgs[r].upload(host[r])
    got = launch(ctx, exe, imgs, o, od)
    src_r, layer_r = list(wl.inputs)
    _, want = PL.c2(PL.u8_image(host[src_r], w, h), PL.f16_image(host[layer_r], w, h))
    d = np.abs(PL.u8_image(got, w, h).astype(np.int32) - want.astype(np.int32))
    ok = want[..., 3] > 0
    report("c2_multi_step", max_rgb=int(d[..., :3][ok].max()), max_alpha=int(d[..., 3].max()),
           exact=f"{float(np.mean(d == 0)):.6f}")
    assert d[..., :3][ok].max() <= 1 and d[..., 3].max() <= 1


@pytest.mark.parametrize("scale,mode", [(0.75, "lds"), (0.02, "direct")])
def test_resample_interpreter_fallback(ctx, scale, mode):
    """A bilinear layer stack with no registered warp form (linear RGBA8
    canvas and output) gets a warp kernel compiled at run time; with run-time
    kernels off it runs on the interpreter resample kernels: LDS-staged tiles,
    or four gathered taps per pixel when a tile's source box exceeds the LDS
    budget.  Both match the oracle."""
    w, h = 192, 160
    d = desc_of("l8", w, h)
    cmd = z.CommandBuffer()
    a, b = cmd.input(d), cmd.input(d)
    ab = cmd.blend(a, z.Rectangle(0, 0, w, h), b)
    canvas = cmd.solid_rgba(d, (0.0, 0.0, 0.0, 0.0))
    aff = W.about_centre(w, h, scale, math.radians(15.0))
    o, od = cmd.output(cmd.affine(canvas, aff, ab))
    prog = z.Linker.from_included().compile(cmd)
    exe = prog.lower_to(ctx, z.PLAN_NO_JIT)
    text = exe.describe()
    assert "[kernel" not in text and f" {mode} " in text, text
    jit = prog.lower_to(ctx)
    assert "[kernel jitw#" in jit.describe(), jit.describe()
    rng = np.random.default_rng(11)
    ha, hb = (rng.integers(0, 256, size=w * h * 4, dtype=np.uint8) for _ in range(2))
    ia, ib = ctx.image(d), ctx.image(d)
    ia.upload(ha)
    ib.upload(hb)
    A = C.premultiply(C.decode_u8(PL.u8_image(ha, w, h), 4))
    Bv = C.premultiply(C.decode_u8(PL.u8_image(hb, w, h), 4))
    lin = C.unpremultiply(C.bilinear_affine(C.over(Bv, A), PL.mat_f32(aff.transformation), (h, w)))
    for tag, e in ((f"fallback_{mode}", exe), (f"fallback_{mode}_jit", jit)):
        got = launch(ctx, e, {a: ia, b: ib}, o, od)
        check(tag, got, lin, "l8", w, h, lin[..., 3] > 0.5 / 255.0)


def _bind_and_run(ctx, wl, host, exe):
    imgs = {}
    for r, d in wl.inputs.items():
        imgs[r] = ctx.image(d)
        imgs[r].upload(host[r])
    outs = {r: ctx.image(d) for r, d in wl.outputs.items()}
    return imgs, outs


def test_autotune_keeps_parity(ctx):
    """The autotuner re-tiles warp kernels (single stages and a batch); the
    tuned programs still match the FP64 oracle and replay their graph."""
    import re
    wl = W.c3(n=6, seed=12, max_side=420)
    host = wl.make_inputs()
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx)
    imgs, outs = _bind_and_run(ctx, wl, host, exe)
    exe.autotune(imgs, outs, iterations=2)
    tiles = set(re.findall(r"tile (\d+x\d+)", exe.describe()))
    assert tiles and tiles <= {"64x16", "32x32", "128x8", "16x64"}, tiles
    # each warp launch reports the shape its stages were tuned to
    launch_tiles = set(re.findall(r"launch \d+: tiles (\d+x\d+)", exe.describe()))
    assert launch_tiles and launch_tiles <= tiles, (launch_tiles, tiles)
    for _ in range(2):
        exe.launch(imgs, outs)
    ctx.sync()
    for k, ((ri, din), (ro, dout)) in enumerate(zip(wl.inputs.items(), wl.outputs.items())):
        w, h, fmt, s = wl.meta["shapes"][k]
        lin, want = PL.c3_image(host[ri], fmt, w, h, dout.width, dout.height)
        got = outs[ro].download().reshape(-1)
        mask = want[..., 3] > (0.5 / 255.0 if fmt == 0 else 1e-3)
        check(f"tuned_c3_img{k}", got, lin, "s8" if fmt == 0 else "h", dout.width, dout.height,
              mask)

    wl = W.c5(layers=4, width=200, height=120, seed=13)
    host = wl.make_inputs()
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx)
    imgs, outs = _bind_and_run(ctx, wl, host, exe)
    exe.autotune(imgs, outs, iterations=2)
    exe.launch(imgs, outs)
    ctx.sync()
    (ro, dout), = wl.outputs.items()
    layers = [PL.f16_image(host[r], 200, 120) for r in wl.inputs]
    lin, _ = PL.c5(layers, wl.meta["matrices"], (120, 200))
    check("tuned_c5", outs[ro].download().reshape(-1), lin, "s8", 200, 120, lin[..., 3] > 0.5 / 255.0)

"""Parity at the benchmark's full sizes, on the configuration the benchmark
runs (tile shapes chosen by the autotuner on the bound images).

The FP64 oracle is evaluated at 65,536 seeded output pixels per workload
(image corners and edges included), which it finishes in seconds at any
image size; the bars are those of the small-size tests:
- 8-bit outputs within 1 code of the oracle, colour compared where the
  oracle's 8-bit alpha is non-zero; alpha within 1 code;
- f16 outputs within 1 ulp where alpha > 1e-3 (un-premultiplying divides by
  the interpolated alpha); alpha within 1 ulp everywhere.
Every comparison also reports its max Oklab Delta-E against the oracle's
linear value.
"""
import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import workloads as W
from oracle import color as C
from oracle import pipelines as PL

from gpu_util import report

pytestmark = pytest.mark.gpu

N_SAMPLES = 65536


def sample_xy(w, h, n, seed):
    """n seeded pixels of a w x h image, its four corners and the middles of
    its edges first."""
    rng = np.random.default_rng(seed)
    fixed = [(0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1), (w // 2, 0), (w // 2, h - 1), (0, h // 2),
             (w - 1, h // 2)]
    k = max(0, n - len(fixed))
    xs = np.concatenate([[p[0] for p in fixed], rng.integers(0, w, k)])[:max(n, 1)]
    ys = np.concatenate([[p[1] for p in fixed], rng.integers(0, h, k)])[:max(n, 1)]
    return xs.astype(np.int64), ys.astype(np.int64)


def launch_tuned(ctx, wl, host):
    """Lower, upload, autotune on the bound images (as the benchmark does),
    launch once; returns {output register: image}."""
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx)
    ins = {}
    for r, d in wl.inputs.items():
        ins[r] = ctx.image(d)
        ins[r].upload(host[r], sync=False)
    outs = {o: ctx.image(d) for o, d in wl.outputs.items()}
    ctx.sync()
    exe.autotune(ins, outs, iterations=2)
    exe.launch(ins, outs)
    ctx.sync()
    return exe, ins, outs


def check_u8(name, got, want, lin):
    """got, want: (N, 4) uint8; lin: the oracle's linear straight RGBA."""
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    ok = want[:, 3] > 0
    mr = int(d[ok, :3].max(initial=0))
    ma = int(d[:, 3].max(initial=0))
    dec = C.decode_u8(got[None], 5)[0][:, :3]
    de = float(np.max(C.delta_e_ok(C.rgb_to_oklab(dec[ok], 3), C.rgb_to_oklab(np.clip(lin[ok, :3], 0, 1), 3)),
                      initial=0.0))
    report(name, samples=got.shape[0], max_rgb=mr, max_alpha=ma, exact=f"{np.mean(d == 0):.5f}",
           max_delta_e_ok=f"{de:.5f}")
    return mr, ma


def test_c4_full_size_sampled(ctx):
    """16 x 3840x2160 RGBA8 through the 7-step colour chain."""
    wl = W.c4()
    host = wl.make_inputs()
    exe, ins, outs = launch_tuned(ctx, wl, host)
    per = N_SAMPLES // len(wl.inputs)
    worst = (0, 0)
    for k, (ri, ro) in enumerate(zip(wl.inputs, wl.outputs)):
        xs, ys = sample_xy(3840, 2160, per, 900 + k)
        src = host[ri].reshape(2160, 3840, 4)[ys, xs]
        lin, want = PL.c4(src[None])
        got = outs[ro].download().reshape(2160, 3840, 4)[ys, xs]
        mr, ma = check_u8(f"c4_full_frame{k}", got, want[0], lin[0])
        worst = (max(worst[0], mr), max(worst[1], ma))
        assert np.array_equal(got[:, 3], src[:, 3])  # alpha passes through the colour chain
    assert worst[0] <= 1 and worst[1] == 0


def test_c5_full_size_sampled(ctx):
    """8 layers of 3840x2160 RGBA16F, rotations within +-45 deg and scales
    0.5-1.5, composited and run through the colour chain."""
    wl = W.c5()
    host = wl.make_inputs()
    exe, ins, outs = launch_tuned(ctx, wl, host)
    (o, od), = wl.outputs.items()
    xs, ys = sample_xy(3840, 2160, N_SAMPLES, 950)
    layers = [np.asarray(host[r]).view(np.float16).reshape(2160, 3840, 4) for r in wl.inputs]
    lin, want = PL.c5_at(layers, wl.meta["matrices"], xs, ys)
    got = outs[o].download().reshape(2160, 3840, 4)[ys, xs]
    mr, ma = check_u8("c5_full", got, want, lin)
    report("c5_full_plan", describe=exe.describe().splitlines()[0][:120])
    assert mr <= 1 and ma <= 1


def test_c3_full_size_sampled(ctx):
    """All 256 images of the seeded batch at full size (sides 64-4096, half
    RGBA8 sRGB, half RGBA16F), 256 samples per image."""
    wl = W.c3()
    host = wl.make_inputs()
    exe, ins, outs = launch_tuned(ctx, wl, host)
    assert "tile queue" in exe.describe()
    per = N_SAMPLES // len(wl.inputs)
    worst_u8, worst_ulp = (0, 0), 0
    n_u8 = n_f16 = 0
    for k, ((ri, din), (ro, dout)) in enumerate(zip(wl.inputs.items(), wl.outputs.items())):
        w, h, fmt, s = wl.meta["shapes"][k]
        ow, oh = dout.width, dout.height
        xs, ys = sample_xy(ow, oh, per, 1000 + k)
        lin, want = PL.c3_image_at(host[ri], fmt, w, h, ow, oh, xs, ys)
        raw = outs[ro].download()
        if fmt == 0:
            got = raw.reshape(oh, ow, 4)[ys, xs]
            d = np.abs(got.astype(np.int32) - want.astype(np.int32))
            ok = want[:, 3] > 0
            worst_u8 = (max(worst_u8[0], int(d[ok, :3].max(initial=0))), max(worst_u8[1], int(d[:, 3].max())))
            n_u8 += 1
        else:
            got = raw.view(np.float16).reshape(oh, ow, 4)[ys, xs]
            a_ok = want[:, 3:4].astype(np.float64) > 1e-3
            ulp = np.abs(got.view(np.uint16).astype(np.int32) - want.view(np.uint16).astype(np.int32))
            ok = np.broadcast_to(a_ok, ulp.shape) | (np.arange(4) == 3)
            worst_ulp = max(worst_ulp, int(ulp[ok].max(initial=0)))
            n_f16 += 1
    report("c3_full", images_u8=n_u8, images_f16=n_f16, max_rgb_u8=worst_u8[0], max_alpha_u8=worst_u8[1],
           max_ulp_f16=worst_ulp)
    assert worst_u8[0] <= 1 and worst_u8[1] <= 1 and worst_ulp <= 1

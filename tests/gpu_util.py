"""Helpers of the GPU parity tests: run a program through the C ABI and
compare with the FP64 oracle."""
import sys

import numpy as np

import zosimos_amd as z
from oracle import color as C


def run(ctx, commands, inputs, host, outputs, flags=0, launches=1):
    """Compile, lower and launch; returns {output register: bytes (H, W*bpp)}."""
    exe = z.Linker.from_included().compile(commands).lower_to(ctx, flags)
    imgs_in = {}
    for reg, desc in inputs.items():
        img = ctx.image(desc)
        img.upload(host[reg])
        imgs_in[reg] = img
    imgs_out = {reg: ctx.image(desc) for reg, desc in outputs.items()}
    for _ in range(launches):
        exe.launch(imgs_in, imgs_out)
    ctx.sync()
    return {reg: img.download() for reg, img in imgs_out.items()}, exe


def report(name, **kv):
    print(f"[parity] {name}: " + ", ".join(f"{k}={v}" for k, v in kv.items()), flush=True)
    sys.stdout.flush()


def compare_u8(name, got, want, alpha_ok=None):
    """Max |difference| in 8-bit codes; optional mask of pixels whose colour
    is well defined (8-bit alpha > 0)."""
    got = got.reshape(want.shape).astype(np.int32)
    want = want.astype(np.int32)
    d = np.abs(got - want)
    if alpha_ok is not None:
        d_rgb = d[..., :3][alpha_ok]
    else:
        d_rgb = d[..., :3]
    d_a = d[..., 3]
    exact = float(np.mean(d == 0))
    report(name, max_rgb=int(d_rgb.max(initial=0)), max_alpha=int(d_a.max(initial=0)), exact_fraction=f"{exact:.6f}")
    return int(d_rgb.max(initial=0)), int(d_a.max(initial=0)), exact


def delta_e_u8(got_u8, want_lin, primaries=3):
    """Max Oklab Delta-E of the decoded GPU output against the FP64 linear
    result (before quantisation)."""
    got = C.decode_u8(got_u8, 5)[..., :3]
    want = np.clip(want_lin[..., :3], 0, 1)
    return float(np.max(C.delta_e_ok(C.rgb_to_oklab(got, primaries), C.rgb_to_oklab(want, primaries))))

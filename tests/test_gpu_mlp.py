"""The learned colour surrogate on the GPU.

* The training kernels' loss and gradient against the FP64 oracle
  (``oracle/mlp.py``), which applies the same bfloat16 roundings and whose own
  gradient is pinned by central differences in ``test_oracle.py``.
* The matrix-core inference kernel against the oracle's forward pass.
* Learning on a program (the learned-chain workload's pointwise chain): what
  training gains over the untrained network, the lattice check's decision,
  the plain kernel as the fallback, and the surrogate kernel inside the
  program against the oracle's forward pass on the trained weights.

Tolerances: the GPU rounds FP32 activations to bfloat16 where the oracle
rounds FP64 ones, so a value within an FP32 error of a rounding boundary can
land one bfloat16 step away (0.4 %); such flips are rare and bounded.
"""
import numpy as np
import pytest

import zosimos_amd as z
from zosimos_amd import workloads as W
from oracle import color as C
from oracle import mlp as M
from oracle import pipelines as PL

from gpu_util import report

pytestmark = pytest.mark.gpu


def params(seed, noise=0.02):
    rng = np.random.default_rng(seed)
    return (M.init_params(seed) + rng.normal(0, noise, M.N_PARAMS)).astype(np.float32)


def test_training_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(7)
    p = params(5)
    x = rng.random((2000, 3)).astype(np.float32)
    t = rng.random((2000, 3)).astype(np.float32)
    loss, g = z.mlp_gradient(ctx, p, x, t)
    ref_loss, ref_g = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), quant=True)
    rel = float(np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g))
    groups = {"W1": (0, 192), "b1": (192, 256), "W2": (256, 4352), "b2": (4352, 4416), "W3": (4416, 4608),
              "b3": (4608, 4611)}
    per = {k: float(np.linalg.norm(g[a:b] - ref_g[a:b]) / (np.linalg.norm(ref_g[a:b]) + 1e-30))
           for k, (a, b) in groups.items()}
    report("mlp_gradient", loss=f"{loss:.6e}", ref_loss=f"{ref_loss:.6e}", rel_norm=f"{rel:.2e}",
           **{k: f"{v:.1e}" for k, v in per.items()})
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    assert rel < 2e-3
    assert max(per.values()) < 1e-2


def test_inference_matches_oracle_forward(ctx):
    rng = np.random.default_rng(8)
    p = params(6, noise=0.05)
    x = rng.random((8192, 3)).astype(np.float32)
    y = z.mlp_evaluate(ctx, p, x)
    ref, _ = M.forward(p.astype(np.float64), x.astype(np.float64), quant=True)
    d = np.abs(y - ref)
    spread = float(np.std(ref))
    report("mlp_inference", max_abs=f"{d.max():.2e}", q999=f"{np.quantile(d, 0.999):.2e}",
           output_std=f"{spread:.3f}")
    assert spread > 0.05  # the check is not vacuous
    assert np.quantile(d, 0.999) < 2e-3 and d.max() < 2e-2


def test_learned_chain_gain_check_and_fallback(ctx):
    w, h = 256, 128
    wl = W.c4(images=1, width=w, height=h, seed=9)
    (ri, din), = wl.inputs.items()
    (ro, dout), = wl.outputs.items()
    host = wl.make_inputs()
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx)
    src = ctx.image(din)
    src.upload(host[ri])

    def run():
        out = ctx.image(dout)
        exe.launch({ri: src}, {ro: out})
        ctx.sync()
        return out.download().reshape(h, w, 4)

    plain = run()
    reps = exe.learn(seed=1)
    assert len(reps) == 1
    r = reps[0]
    report("learned_chain", accepted=r["accepted"], max_de=f"{r['max_delta_e']:.5f}",
           mean_de=f"{r['mean_delta_e']:.5f}", max_de_untrained=f"{r['max_delta_e_initial']:.4f}",
           loss=f"{r['initial_loss']:.3e}->{r['final_loss']:.3e}", train_ms=f"{r['train_ms']:.0f}")
    # What training gains over the untrained network.
    assert r["final_loss"] < r["initial_loss"] / 100
    assert r["max_delta_e"] < r["max_delta_e_initial"] / 5
    assert r["mean_delta_e"] < 0.01
    # The check's decision follows its bound.
    assert r["accepted"] == (1 if r["max_delta_e"] < 0.005 else 0)

    # Plain kernels: bit-identical to the program before learning.
    exe.use_surrogates(0)
    assert np.array_equal(run(), plain)
    # Default mode: the surrogate runs only if accepted.
    exe.use_surrogates(1)
    if not r["accepted"]:
        assert np.array_equal(run(), plain)

    # Forced: the surrogate kernel inside the program against the oracle's
    # forward pass on the trained weights.  8-bit outputs agree to 1 code
    # except where a bfloat16 flip of a hidden activation (one step of 2^-8
    # relative, times a trained output weight) moves the value by a few codes.
    exe.use_surrogates(2)
    got = run()
    pw = exe.surrogate_params(0).astype(np.float64)
    x = host[ri].reshape(h, w, 4)[..., :3].reshape(-1, 3).astype(np.float64) / 255.0
    y, _ = M.forward(pw, x, quant=True)
    want = np.rint(np.clip(y, 0, 1) * 255.0).reshape(h, w, 3)
    dc = np.abs(got[..., :3].astype(np.int32) - want.astype(np.int32))
    report("surrogate_kernel", max_code=int(dc.max()), exact=f"{float(np.mean(dc == 0)):.5f}")
    assert np.array_equal(got[..., 3], host[ri].reshape(h, w, 4)[..., 3])  # alpha passes through
    assert dc.max() <= 3 and np.mean(dc <= 1) > 0.995
    # ... and against the plain chain's FP64 result, in Oklab.
    lin, _ = PL.c4(PL.u8_image(host[ri], w, h))
    de = C.delta_e_ok(C.rgb_to_oklab(C.decode_u8(got, 5)[..., :3], 3), C.rgb_to_oklab(np.clip(lin[..., :3], 0, 1), 3))
    report("surrogate_vs_fp64", max_de=f"{de.max():.5f}", mean_de=f"{de.mean():.5f}")
    assert de.mean() < 0.01
    exe.use_surrogates(1)


def test_surrogate_against_plain_chain(ctx):
    """The learned chain beside the method it would replace, on a 1920x1080
    frame: device time of the fused plain chain and of the forced surrogate
    (per-stage events over 20 launches each), and each output's max Delta-E_OK
    against the FP64 chain.  The plain chain is both faster and more accurate,
    so the surrogate stays off unless its lattice error is under the bound."""
    w, h = 1920, 1080
    wl = W.c4(images=1, width=w, height=h, seed=3)
    (ri, din), = wl.inputs.items()
    (ro, dout), = wl.outputs.items()
    host = wl.make_inputs()
    exe = z.Linker.from_included().compile(wl.commands).lower_to(ctx)
    src, out = ctx.image(din), ctx.image(dout)
    src.upload(host[ri])
    r = exe.learn(seed=1)[0]
    lin, _ = PL.c4(PL.u8_image(host[ri], w, h))
    ref = C.rgb_to_oklab(np.clip(lin[..., :3], 0, 1), 3)

    def measure(mode):
        exe.use_surrogates(mode)
        ms = sum(exe.time_stages({ri: src}, {ro: out}, 20))
        got = out.download().reshape(h, w, 4)
        de = C.delta_e_ok(C.rgb_to_oklab(C.decode_u8(got, 5)[..., :3], 3), ref)
        return ms, float(de.max()), float(de.mean())

    plain = measure(0)
    surrogate = measure(2)
    exe.use_surrogates(1)
    report("surrogate_vs_plain", plain_ms=f"{plain[0]:.4f}", surrogate_ms=f"{surrogate[0]:.4f}",
           speed_ratio=f"{plain[0] / surrogate[0]:.3f}", plain_max_de=f"{plain[1]:.5f}",
           surrogate_max_de=f"{surrogate[1]:.5f}", lattice_max_de=f"{r['max_delta_e']:.5f}", bound=0.005,
           accepted=r["accepted"])
    # 8-bit quantisation alone keeps the plain chain within ~0.005 of FP64
    assert plain[1] < surrogate[1]
    assert r["accepted"] == (1 if r["max_delta_e"] < 0.005 else 0)

"""The all-core FP32 CPU restatement of the workloads (``oracle/cpu_chain.c``,
the benchmark's CPU baseline) against the FP64 oracle, at small sizes, with
the GPU tests' bars: 8-bit outputs within 1 code (colour where alpha > 0),
f16 outputs within 1 ulp where alpha > 1e-3.  Its binary16 rounding is
checked against numpy's."""
import numpy as np
import pytest

from zosimos_amd import workloads as W
from oracle import color as C
from oracle import cpu as CPU
from oracle import pipelines as PL


def _u8_close(got, want):
    got = got.reshape(want.shape).astype(np.int32)
    d = np.abs(got - want.astype(np.int32))
    ok = want[..., 3] > 0
    return int(d[..., :3][ok].max(initial=0)), int(d[..., 3].max(initial=0))


def test_half_rounding_matches_numpy():
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 100,
                        rng.standard_normal(20000).astype(np.float32) * 1e-5,
                        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, np.inf, -np.inf, 6e-8, 3e-8,
                                  2.98e-8, 5.96e-8, 6.1e-5], np.float32)])
    got = np.array([CPU.lib().zsc_f2h(float(x)) for x in v], np.uint16)
    assert np.array_equal(got, v.astype(np.float16).view(np.uint16))


def test_c2_chain_matches_oracle():
    w, h = 256, 64
    rng = np.random.default_rng(2)
    src = rng.integers(0, 256, (h, w * 4), dtype=np.uint8)
    layer = W.random_rgba16f(rng, w, h)
    got = CPU.c2(src, layer.view(np.uint16))
    _, want = PL.c2(PL.u8_image(src, w, h), PL.f16_image(layer, w, h))
    mr, ma = _u8_close(got, want)
    assert mr <= 1 and ma <= 1


def test_c4_chain_matches_oracle():
    w, h = 256, 64
    src = np.random.default_rng(3).integers(0, 256, (h, w * 4), dtype=np.uint8)
    got = CPU.c4(src)
    _, want = PL.c4(PL.u8_image(src, w, h))
    mr, ma = _u8_close(got, want)
    assert mr <= 1 and ma == 0


@pytest.mark.parametrize("fmt,w,h,ow,oh", [(0, 61, 37, 97, 20), (1, 61, 37, 97, 20), (0, 300, 200, 150, 399),
                                           (1, 2, 2, 5, 7)])
def test_c3_image_matches_oracle(fmt, w, h, ow, oh):
    rng = np.random.default_rng(4 + fmt)
    data = W.random_rgba8(rng, w, h) if fmt == 0 else W.random_rgba16f(rng, w, h)
    got = CPU.c3_image(data, fmt, w, h, ow, oh)
    _, want = PL.c3_image(data, fmt, w, h, ow, oh)
    if fmt == 0:
        mr, ma = _u8_close(got, want)
        assert mr <= 1 and ma <= 1
    else:
        g = got.view(np.float16).reshape(oh, ow, 4)
        ulp = np.abs(g.view(np.uint16).astype(np.int32) - want.view(np.uint16).astype(np.int32))
        ok = np.broadcast_to(want[..., 3:4].astype(np.float64) > 1e-3, ulp.shape) | (np.arange(4) == 3)
        assert ulp[ok].max(initial=0) <= 1


def test_c5_matches_oracle():
    w, h = 96, 64
    wl = W.c5(8, w, h, seed=5)
    host = wl.make_inputs()
    got = CPU.c5([host[r] for r in wl.inputs], wl.meta["matrices"], w, h)
    layers = [PL.f16_image(host[r], w, h) for r in wl.inputs]
    _, want = PL.c5(layers, wl.meta["matrices"], (h, w))
    mr, ma = _u8_close(got, want)
    assert mr <= 1 and ma <= 1


def test_c1_matches_oracle():
    w, h = 128, 96
    wl = W.c1(w, h, seed=6)
    host = wl.make_inputs()
    a, b = list(wl.inputs)
    got = CPU.c1(host[a], host[b], wl.meta["affine"], w, h)
    _, want = PL.c1(PL.u8_image(host[a], w, h), PL.u8_image(host[b], w, h), wl.meta["affine"], (h, w))
    mr, ma = _u8_close(got, want)
    assert mr <= 1 and ma <= 1


def test_uses_every_core():
    assert CPU.threads() >= 1

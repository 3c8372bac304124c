"""Planning of LDS-staged warp tiles (no GPU): under ZH_PLAN_STAGED a
one-layer stage with a separable map whose per-tile source box fits the LDS
box runs the staged kernel, as its own launch unit; minifying maps whose
boxes do not fit, maps with rotation and the default plan keep the direct
gathers."""
import re

import zosimos_amd as z
from zosimos_amd import workloads as W


def plan(cmd, flags=z.PLAN_STAGED):
    return z.Linker.from_included().compile(cmd).lower_to(None, flags).describe()


def one(scale_x, scale_y, rotate=0.0, w=300, h=200):
    d = z.Descriptor.rgba8_srgb(w, h)
    cmd = z.CommandBuffer()
    r = cmd.input(d)
    c = cmd.color_convert(r, z.Color(primaries=z.Primaries.DisplayP3, transfer=z.Transfer.Srgb), W.RGBA8)
    ow, oh = int(w * abs(scale_x)), int(h * abs(scale_y))
    canvas = cmd.solid_rgba(z.Descriptor(ow, oh, W.RGBA8, z.Color(primaries=z.Primaries.DisplayP3,
                                                                    transfer=z.Transfer.Srgb)), (0, 0, 0, 0))
    aff = z.Affine.new(z.AffineSample.BiLinear).scale(scale_x, scale_y)
    if scale_x < 0:
        aff = aff.shift(ow, 0)
    if rotate:
        aff = aff.rotate(rotate)
    cmd.output(cmd.affine(canvas, aff, c))
    return cmd


def test_magnifying_separable_stage_is_staged():
    text = plan(one(1.7, 1.3))
    assert "[kernel w_s8m_s8_sep_lds]" in text, text
    assert "_lds" not in plan(one(1.7, 1.3), 0)
    # a flip keeps the box: its bounds come from both ends of the tile
    assert "[kernel w_s8m_s8_sep_lds]" in plan(one(-1.5, 1.0))
    # unit scale: 69 x 21 texels on 64 x 16 tiles
    assert "[kernel w_s8m_s8_sep_lds]" in plan(one(1.0, 1.0))


def test_boxes_that_do_not_fit_stay_direct():
    # 0.5x: a 64 x 16 tile's box is 132 x 36 texels
    assert "[kernel w_s8m_s8_sep]" in plan(one(0.5, 0.5))
    # rotation: not separable
    text = plan(one(1.5, 1.5, rotate=0.2))
    assert "[kernel w_s8m_s8]" in text and "_lds" not in text


def test_c3_batch_splits_staged_and_direct_units():
    text = plan(W.c3(n=24, seed=3, max_side=400).commands)
    staged = [l for l in text.splitlines() if l.startswith("stage ") and "_lds]" in l]
    direct = [l for l in text.splitlines() if l.startswith("stage ") and "_sep]" in l]
    assert staged and direct
    for line in staged:
        # the box bound: floor(|m| (tile - 1)) + 6 per axis (one more column for the odd pitch) within 1600 texels
        m = re.search(r"resample (\d+)x(\d+) .*layer0\(bilinear-over (\d+)x(\d+)\)", line)
        ow, oh, sw, sh = map(int, m.groups())
        assert (int(sw / ow * 63) + 7) * (int(sh / oh * 15) + 6) <= 1600, line
    launches = [l for l in text.splitlines() if l.startswith("launch ")]
    assert any("_lds]" in l for l in launches) and any("_sep]" in l for l in launches), launches


def test_warp_forms_need_fixed_point_steps():
    """The warp kernels map pixels with the layer's steps in 8.56 fixed point
    (|m| < 127): a 1/200 reduction stays on the interpreter's resampler, a
    1/100 one gets the warp kernel."""
    assert "[kernel w_s8m_s8" not in plan(one(1 / 200, 1 / 200, w=1000, h=800), 0)
    assert "[kernel w_s8m_s8_sep]" in plan(one(1 / 100, 1 / 100, w=1000, h=800), 0)

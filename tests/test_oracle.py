"""Pins of the CPU oracle against published values, library routines and
invariants (no GPU).  The oracle is the checker of every GPU parity test, so
each of its functions is checked here by something other than its own
formula."""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import color as C
from oracle import pipelines as PL
from oracle import stage as S


# ---- colour matrices: published values ---------------------------------------
def test_srgb_to_xyz_matches_iec_61966_2_1():
    # IEC 61966-2-1:1999 Table (4-decimal matrix).
    published = np.array([[0.4124, 0.3576, 0.1805], [0.2126, 0.7152, 0.0722], [0.0193, 0.1192, 0.9505]])
    assert np.allclose(C.rgb_to_xyz_matrix(3, 5), published, atol=6e-5)


def test_display_p3_to_xyz_matches_published():
    # Display P3 (SMPTE EG 432-1 primaries, D65), 4-decimal published matrix.
    published = np.array([[0.4866, 0.2657, 0.1982], [0.2290, 0.6917, 0.0793], [0.0000, 0.0451, 1.0439]])
    assert np.allclose(C.rgb_to_xyz_matrix(7, 5), published, atol=6e-5)


def test_bt2020_luma_coefficients():
    # ITU-R BT.2020-2: Y = 0.2627 R + 0.6780 G + 0.0593 B.
    assert np.allclose(C.rgb_to_xyz_matrix(5, 5)[1], [0.2627, 0.6780, 0.0593], atol=6e-5)


def test_white_maps_to_white_and_inverse():
    for p in (1, 2, 3, 4, 5, 7):
        m = C.rgb_to_xyz_matrix(p, 5)
        assert np.allclose(m @ np.ones(3), C.whitepoint_xyz(5), atol=1e-12)
    m = C.convert_matrix(3, 7, 5)
    assert np.allclose(C.convert_matrix(7, 3, 5) @ m, np.eye(3), atol=1e-12)
    assert np.allclose(C.convert_matrix(3, 3, 5), np.eye(3), atol=1e-12)


# ---- Oklab: Ottosson's table and CSS Color 4 ---------------------------------
@pytest.mark.parametrize("xyz,lab", [
    ((0.950, 1.000, 1.089), (1.000, 0.000, 0.000)),
    ((1.000, 0.000, 0.000), (0.450, 1.236, -0.019)),
    ((0.000, 1.000, 0.000), (0.922, -0.671, 0.263)),
    ((0.000, 0.000, 1.000), (0.153, -1.415, -0.449)),
])
def test_oklab_known_answers(xyz, lab):
    # Table of XYZ -> Oklab values published with the Oklab definition.
    assert np.allclose(C.xyz_to_oklab(np.array(xyz)), lab, atol=1.5e-3)


def test_oklab_srgb_red_css_color_4():
    # CSS Color 4: red = oklab(0.6279554 0.2248631 0.1258463).
    assert np.allclose(C.rgb_to_oklab(np.array([1.0, 0, 0]), 3), [0.6279554, 0.2248631, 0.1258463], atol=5e-4)


def test_oklab_round_trip_and_adjust():
    rng = np.random.default_rng(1)
    rgb = rng.random((1000, 3))
    lab = C.rgb_to_oklab(rgb, 3)
    assert np.allclose(C.oklab_to_rgb(lab, 3, clamp=False), rgb, atol=1e-9)
    # chroma scaling multiplies the polar radius; hue rotation preserves it
    adj = C.lch_adjust(lab, 1.2, 30.0)
    assert np.allclose(np.hypot(adj[:, 1], adj[:, 2]), 1.2 * np.hypot(lab[:, 1], lab[:, 2]), atol=1e-12)
    ang = np.arctan2(adj[:, 2], adj[:, 1]) - np.arctan2(lab[:, 2], lab[:, 1])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    ok = np.hypot(lab[:, 1], lab[:, 2]) > 1e-6
    assert np.allclose(ang[ok], math.radians(30.0), atol=1e-9)
    assert np.allclose(adj[:, 0], lab[:, 0])


# ---- transfers ---------------------------------------------------------------------
def test_srgb_known_values():
    assert C.eo_srgb(np.array(0.5)) == pytest.approx(0.214041140, abs=1e-9)
    assert C.eo_srgb(np.array(128 / 255)) == pytest.approx(0.215860500, abs=1e-8)
    assert C.oe_srgb(np.array(0.5)) == pytest.approx(0.735356983, abs=1e-9)
    assert C.eo_srgb(np.array(1.0)) == pytest.approx(1.0)
    # sign-symmetric extension (stage.frag:335-352)
    assert C.eo_srgb(np.array(-0.5)) == pytest.approx(-0.214041140, abs=1e-9)


def test_transfer_round_trips():
    v = np.linspace(0.0, 1.0, 10001)
    assert np.allclose(C.eo_srgb(C.oe_srgb(v)), v, atol=1e-12)
    assert np.allclose(C.eo_bt709(C.oe_bt709(v)), v, atol=1e-12)
    assert np.allclose(C.eo_smpte2084(C.eo_inv_smpte2084(v)), v, atol=1e-9)
    # PQ end points (BT.2100 Table 4)
    assert C.eo_smpte2084(np.array(1.0)) == pytest.approx(1.0)
    assert C.eo_smpte2084(np.array(0.0)) == pytest.approx(0.0)


def test_srgb_u8_encode_decode_identity():
    codes = np.arange(256, dtype=np.uint8)
    px = np.stack([codes, codes, codes, codes], axis=-1)
    assert np.array_equal(C.encode_u8(C.decode_u8(px, 5), 5), px)
    assert np.array_equal(C.encode_u8(C.decode_u8(px, 4), 4), px)


# ---- stage.frag restatement (FP32 C) -----------------------------------------------
def test_stage_u8_u16_exhaustive_round_trip():
    # Correctly rounded division round-trips every code.
    v8 = np.arange(256, dtype=np.uint32)
    assert np.array_equal(S.stage_encode(S.stage_decode(v8, 0, 4, 4), 0, 4, 4), v8)
    v16 = np.arange(65536, dtype=np.uint32)
    assert np.array_equal(S.stage_encode(S.stage_decode(v16, 3, 4, 4), 3, 4, 4), v16)


@pytest.mark.parametrize("bits,nbits", [(1, 8), (2, 8), (4, 16), (7, 16), (8, 16), (14, 32), (15, 32)])
def test_stage_packed_round_trip(bits, nbits):
    rng = np.random.default_rng(bits)
    v = np.arange(1 << nbits, dtype=np.uint32) if nbits <= 16 else rng.integers(0, 1 << 32, 200000, dtype=np.uint64).astype(np.uint32)
    parts = 8 if bits in (4, 14, 15) else (5 if bits == 8 else 6)
    dec = S.stage_decode(v, bits, parts, 4)
    assert np.array_equal(S.stage_encode(dec, bits, parts, 4), v)


def test_stage_demux_known_fields():
    # Int565 (stage.frag:549-550): low 5 bits are the first channel.
    out = S.stage_decode(np.array([0x001F, 0x07E0, 0xF800], dtype=np.uint32), 7, 6, 4)
    assert np.allclose(out[:, :3], [[1, 0, 0], [0, 1, 0], [0, 0, 1]])
    # Int8x4 RGBA: bytes in memory order
    out = S.stage_decode(np.array([0x80402010], dtype=np.uint32), 10, 8, 4)
    assert np.allclose(out[0], np.array([0x10, 0x20, 0x40, 0x80]) / 255.0)
    # truncating encode: 0.999 * 255 -> 254
    assert S.stage_encode(np.array([[0.999, 0.999, 0.999, 0.999]], np.float32), 0, 4, 4)[0] == 254


def test_stage_srgb_matches_fp64():
    v = np.arange(256, dtype=np.uint32)
    got = S.stage_decode(v, 0, 4, 5)[:, 0]
    assert np.allclose(got, C.srgb_u8_decode_table(), atol=2e-7)


# ---- bilinear resampling ------------------------------------------------------------
def test_bilinear_matches_scipy_grid_constant():
    rng = np.random.default_rng(7)
    src = rng.random((23, 31, 4))
    m = PL.mat_f32(C.affine_matrix((-15.5, -11.5), (0.8, 1.3), 0.4, (20.0, 14.0)))
    out = C.bilinear_affine(src, m, (29, 41))
    inv = np.linalg.inv(m)
    ys, xs = np.mgrid[0:29, 0:41].astype(np.float64)
    q = inv @ np.stack([xs.ravel() + 0.5, ys.ravel() + 0.5, np.ones(xs.size)])
    u, v = q[0] / q[2] - 0.5, q[1] / q[2] - 0.5
    for ch in range(4):
        ref = ndimage.map_coordinates(src[..., ch], [v, u], order=1, mode="grid-constant", cval=0.0)
        assert np.allclose(out[..., ch].ravel(), ref, atol=1e-12)


def test_bilinear_invariants():
    rng = np.random.default_rng(8)
    src = rng.random((16, 12, 4))
    assert np.allclose(C.bilinear_affine(src, np.eye(3), (16, 12)), src)
    shifted = C.bilinear_affine(src, C.affine_matrix(shift_post=(3, 2)), (16, 12))
    assert np.allclose(shifted[2:, 3:], src[:14, :9])
    assert np.allclose(shifted[:2], 0) and np.allclose(shifted[:, :3], 0)
    # rotation by 90 degrees about the centre of a square image is a permutation
    sq = rng.random((10, 10, 4))
    m = C.affine_matrix((-5, -5), (1, 1), math.pi / 2, (5, 5))
    rot = C.bilinear_affine(sq, m, (10, 10))
    assert np.allclose(np.sort(rot.ravel()), np.sort(sq.ravel()))


def test_porter_duff_invariants():
    rng = np.random.default_rng(9)
    a = C.premultiply(rng.random((50, 4)))
    b = C.premultiply(rng.random((50, 4)))
    opaque = a.copy(); opaque[:, 3] = 1.0
    assert np.allclose(C.over(opaque, b), opaque)
    clear = np.zeros_like(a)
    assert np.allclose(C.over(clear, b), b)
    straight = rng.random((50, 4)) + 0.01
    assert np.allclose(C.unpremultiply(C.premultiply(straight)), straight)


# ---- paint draws ---------------------------------------------------------------------
def test_paint_identity_and_offset_copy():
    rng = np.random.default_rng(10)
    base = rng.random((20, 30, 4))
    src = rng.random((8, 8, 4))
    out = PL.paint_nearest(base, src, (0, 0, 8, 8), PL.inscribe_quad((5, 7, 13, 15)), (0, 0, 30, 20))
    ref = base.copy()
    ref[7:15, 5:13] = src
    assert np.array_equal(out, ref)


def test_inscribe_normalize_quirk_stretches_height():
    # A 6x4 foreground becomes a 6x6 quad (max_y = y + width, command.rs:3541).
    assert PL.inscribe_quad((0, 0, 6, 4)) == [0, 0, 6, 0, 6, 6, 0, 6]


def test_crop_stretches_selection():
    src = np.arange(8 * 8 * 4, dtype=np.float64).reshape(8, 8, 4)
    out = PL.crop(src, (2, 2, 6, 6))
    # every 2x2 block of the output shows one texel of the 4x4 selection
    assert np.array_equal(out[0, 0], src[2, 2]) and np.array_equal(out[7, 7], src[5, 5])
    assert np.array_equal(out[1, 1], src[2, 2]) and np.array_equal(out[2, 2], src[3, 3])


# ---- the learned colour surrogate's oracle (oracle/mlp.py) ------------------------
def test_mlp_gradient_matches_central_differences():
    """Reverse-mode gradient of the FP64 network against central
    differences on every parameter of a small random problem."""
    from oracle import mlp as M
    rng = np.random.default_rng(1)
    p = M.init_params(3) + rng.normal(0, 0.05, M.N_PARAMS)
    x, t = rng.random((7, 3)), rng.random((7, 3))
    _, g = M.loss_and_grad(p, x, t, quant=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(M.N_PARAMS):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (M.loss_and_grad(p + e, x, t, quant=False)[0] - M.loss_and_grad(p - e, x, t, quant=False)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))


def test_mlp_bf16_rounding_and_silu():
    from oracle import mlp as M
    # ties to even at 1 + 2^-8 (halfway between 1 and 1 + 2^-7)
    assert M.bf16(1.0 + 2 ** -8) == 1.0
    assert M.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert M.bf16(-2.0 ** -3 * (1 + 2 ** -9)) == -2.0 ** -3
    assert M.bf16(3.140625) == 3.140625  # representable
    a = np.linspace(-8, 8, 101)
    assert np.allclose(M.silu(a), a / (1 + np.exp(-a)), rtol=1e-14, atol=1e-15)
    assert np.allclose(M.silu_grad(a), np.gradient(M.silu(a), a), atol=2e-2)
    assert M.silu(0.0) == 0.0 and M.silu_grad(0.0) == 0.5


def test_mlp_forward_layout_and_adam_first_step():
    """A network whose only path is x -> 2x: W1 picks x into feature 0, W2
    copies it, W3 reads it back; the first Adam step moves every parameter
    with a non-zero gradient by -lr * sign(g)."""
    from oracle import mlp as M
    w1 = np.zeros((3, 64)); w1[0, 0] = 1.0
    w2 = np.zeros((64, 64)); w2[0, 0] = 1.0
    w3 = np.zeros((64, 3)); w3[0, 2] = 2.0
    p = M.pack(w1, np.zeros(64), w2, np.zeros(64), w3, np.zeros(3))
    x = np.array([[0.5, 0.1, 0.2]])
    y, _ = M.forward(p, x, quant=False)
    assert np.isclose(y[0, 2], 2 * M.silu(M.silu(0.5))) and y[0, 0] == 0 and y[0, 1] == 0
    _, g = M.loss_and_grad(p, x, np.zeros((1, 3)), quant=False)
    q, m, v = M.adam_step(p, g, np.zeros_like(p), np.zeros_like(p), 1, 0.01)
    nz = np.abs(g) > 1e-12
    assert np.allclose((q - p)[nz], -0.01 * np.sign(g[nz]), rtol=1e-6)


# ---- perceptual hash of the reference's tests (oracle/blockhash.py) --------------
def _golden_rgba(name):
    import os
    from PIL import Image
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".png")
    return np.asarray(Image.open(here).convert("RGBA"))


def test_blockhash_known_answers():
    """The restated blockhash256 gives the reference's blessed hash of its
    red/green-swapped background for the unmodified background (a channel
    swap keeps every r + g + b sum), and all-ones for a uniform bright image
    (every block equals its band's median and exceeds half the block
    maximum)."""
    import json, os
    from oracle import blockhash as B
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    bg = _golden_rgba("background")
    assert bg.shape == (512, 512, 4)
    assert B.blockhash256(bg) in hashes["swapped"]
    swapped = bg[..., [1, 0, 2, 3]]
    assert B.blockhash256(swapped) == B.blockhash256(bg)
    assert B.blockhash256(np.full((64, 48, 4), 200, np.uint8)) == "f" * 64
    # transparent pixels count as white; the precise method spreads a
    # 17-pixel-wide image over 16 block columns without losing weight
    odd = np.random.default_rng(0).integers(0, 256, (33, 17, 4), dtype=np.uint8)
    ay, ax = B._axis_weights(33, 16), B._axis_weights(17, 16)
    assert np.allclose(ay.sum(1), 1) and np.allclose(ax.sum(1), 1) and np.isclose(ay.sum(), 33)
    assert len(B.blockhash256(odd)) == 64


def test_blockhash_of_lumaa16_matches_transmute_hash():
    """The reference transmutes its RGBA8 background to LumaA16 (a byte copy,
    ``blend.rs:340-375``) and hashes the LumaA16 image: the background's
    bytes read as LumaA16, rounded to 8 bits per channel, give the blessed
    ``transmute`` hash; truncating instead of rounding does not."""
    import json, os
    from oracle import blockhash as B
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    bg = _golden_rgba("background")
    assert B.blockhash256(B.lumaa16_to_rgba8(bg, 512, 512)) in hashes["transmute"]
    v = np.frombuffer(bg.tobytes(), dtype="<u2").reshape(512, 512, 2)
    t = (v // 257).astype(np.uint8)
    assert B.blockhash256(np.stack([t[..., 0]] * 3 + [t[..., 1]], -1)) not in hashes["transmute"]


def test_bilinear_grid_known_answers():
    """bilinear.frag: the reference's mgrid parameters (shaders/bilinear.rs:
    22-31) give each pixel centre's coordinates; constant corners give the
    constant; the uv term alone is (u v) scaled at the pixel centres."""
    from oracle import pipelines as PL
    g = PL.bilinear_grid(5, 3, [0, 0, 0, 0, 5, 0, 0, 0, 0, 0, 0, 0, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    xs, ys = np.meshgrid(np.arange(5) + 0.5, np.arange(3) + 0.5)
    assert np.array_equal(g[..., 0], xs) and np.array_equal(g[..., 1], ys) and not g[..., 2:].any()
    c = PL.bilinear_grid(4, 4, [0.25] * 4 * 4 + [0.0] * 8)
    assert np.allclose(c, 0.5)  # mix(0.25, 0.25, u) + mix(0.25, 0.25, v)
    uv = PL.bilinear_grid(2, 2, [0.0] * 20 + [1.0, 2.0, 3.0, 4.0])
    w = np.array([[0.0625, 0.1875], [0.1875, 0.5625]])
    assert np.allclose(uv, w[..., None] * np.array([1.0, 2.0, 3.0, 4.0]))


def test_oklab_lch_decode_known_answers():
    """8-bit LchA Oklab: zero chroma gives the same grey whatever the hue
    (neutral to 1e-3: Ottosson's matrices against the D65 sRGB matrix derived
    from chromaticities); L = 1, C = 0 decodes to white; hue 128/255 puts the
    chroma near +a."""
    from oracle import pipelines as PL
    grey = np.array([[[204, 0, h, 255] for h in (0, 64, 128, 255)]], np.uint8)
    lin, enc = PL.oklab_lch8_to_srgb(grey)
    assert np.all(lin == lin[:, :1]) and np.all(enc == enc[:, :1])
    assert np.ptp(lin[..., :3], axis=-1).max() < 1e-3
    white, _ = PL.oklab_lch8_to_srgb(np.array([[[255, 0, 0, 255]]], np.uint8))
    assert np.allclose(white[..., :3], 1.0, atol=1e-3)
    from oracle import color as C
    lab = C.decode_u8(np.array([[[128, 51, 128, 255]]], np.uint8), 0x100)
    assert np.isclose(lab[0, 0, 1], 0.2 * np.cos(np.radians(360.0 * (128 / 255 - 0.5))))


def test_oklab_hue_plane_reproduces_blessed_hash():
    """blend.rs:449-508 (run_oklab): a bilinear grid of constant lightness 0.8,
    chroma 0 to 1 down and hue 0 to 1 across, stored to RGBA8 with a Scalars
    colour (a staged format: truncating store), reinterpreted as LchA Oklab
    and converted to sRGB.  The restated pipeline hashes to a blessed line;
    rounding the staged store instead does not."""
    import json, os
    from oracle import blockhash as B
    from oracle import pipelines as PL
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    g = PL.bilinear_grid(400, 400, OKLAB_PLANE)
    _, enc = PL.oklab_lch8_to_srgb(PL.staged_u8_store(g))
    assert B.blockhash256(enc) in hashes["oklab"]
    _, rounded = PL.oklab_lch8_to_srgb(np.rint(np.clip(g, 0, 1) * 255).astype(np.uint8))
    assert B.blockhash256(rounded) not in hashes["oklab"]


# blend.rs:466-483: u_min, u_max, v_min, v_max, uv_min, uv_max
OKLAB_PLANE = [0.4, 0.0, 0.0, 1.0, 0.4, 0.0, 1.0, 1.0, 0.4, 0.0, 0.0, 1.0, 0.4, 1.0, 0.0, 1.0] + [0.0] * 8


def test_chromatic_adaptation_matrix_and_blessed_hash():
    """Adapting white to its own white point is the identity for every
    method; the source white maps to the target white in XYZ; and the
    reference's run_adaptation (blend.rs:169-197: the sRGB background adapted
    from D65 to D50 by von Kries) hashes to its blessed line, which the
    Bradford and XYZ-scaling matrices miss."""
    import json, os
    from oracle import blockhash as B
    from oracle import color as C
    for m in (0, 1, 2):
        assert np.allclose(C.adaptation_matrix(m, 5, 5), np.eye(3))
        assert np.allclose(C.adaptation_matrix(m, 5, 3) @ C.whitepoint_xyz(5), C.whitepoint_xyz(3))
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    lin = C.decode_u8(_golden_rgba("background"), 5)
    got = {}
    for m in (0, 1, 2):
        out = lin.copy()
        out[..., :3] = np.einsum("ij,...j->...i", C.adapt_rgb_matrix(3, 5, 3, m), lin[..., :3])
        got[m] = B.blockhash256(C.encode_u8(out, 5))
    assert got[1] in hashes["adapted"]
    assert got[0] not in hashes["adapted"] and got[2] not in hashes["adapted"]


def test_blockhash_of_an_image_narrower_than_the_grid():
    """buffer.rs:39-65 hashes an 8x8 image on the 16x16 grid: each pixel
    covers two block rows and columns, and the restated precise method
    reproduces the blessed from_buffer hash (row 1 black, the rest white)."""
    import json, os
    from oracle import blockhash as B
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    lum = np.full((8, 8), 255, np.uint8)
    lum[1] = 0
    assert B.blockhash256(np.stack([lum, lum, lum, np.full_like(lum, 255)], -1)) in hashes["from_buffer"]
    w = B._axis_weights(8, 16)
    assert np.allclose(w.sum(0), 0.5) and np.allclose(w.sum(1), 1.0)


def test_derivatives_reproduce_blessed_hashes():
    """blend.rs:577-618 (run_derivative, width direction): box3 of the linear
    background with each method's weights, encoded to sRGB, hashes to the
    blessed line of that method; weights applied transposed (x and y offsets
    swapped) do not.  A constant image has zero derivative."""
    import json, os
    from oracle import blockhash as B
    from oracle import color as C
    from oracle import pipelines as PL
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    lin = C.decode_u8(_golden_rgba("background"), 5)
    for m, name in ((1, "Prewitt"), (2, "Sobel"), (3, "Scharr3"), (6, "Scharr3To4Bit"), (8, "Scharr3To8Bit")):
        w = PL.derivative_weights(m)
        assert B.blockhash256(C.encode_u8(PL.box3(lin, w), 5)) in hashes["derived_" + name]
        assert B.blockhash256(C.encode_u8(PL.box3(lin, w.T), 5)) not in hashes["derived_" + name]
        assert np.allclose(PL.box3(np.full((5, 6, 4), 0.3), w)[..., :3], 0.0)


# ---- distribution_normal2d.frag -------------------------------------------------------
def _diag_params(v0, v1):
    """Inverse covariance diag(1/v0, 1/v1) and determinant (2 pi)^2 v0 v1."""
    return [0.0, 0.0, 1.0 / v0, 0.0, 0.0, 1.0 / v1, (2 * math.pi) ** 2 * v0 * v1, 0.0]


def test_normal2d_is_a_density():
    """Summed over the pixels of the screen square [-1, 1]^2 (pixel area
    (2/w)(2/h)) the density integrates to 1 when its mass lies inside; its
    peak at an exact pixel centre is 1 / sqrt(det); the marginal along x has
    the variance v0."""
    w, h = 301, 201
    v0, v1 = 0.02, 0.035
    g = PL.normal2d(w, h, _diag_params(v0, v1))[..., 0]
    assert abs(g.sum() * (2.0 / w) * (2.0 / h) - 1.0) < 1e-6
    assert np.isclose(g[h // 2, w // 2], 1.0 / math.sqrt(_diag_params(v0, v1)[6]), rtol=1e-12)
    x = 2.0 * (np.arange(w) + 0.5) / w - 1.0
    mx = g.sum(axis=0)
    assert np.isclose((mx * x * x).sum() / mx.sum(), v0, rtol=1e-5)
    # a shifted expectation moves the peak: mu = (0.5, 0) puts it 3/4 across
    mu = _diag_params(v0, v1)
    mu[0] = 0.5
    g2 = PL.normal2d(w, h, mu)[..., 0]
    assert np.unravel_index(np.argmax(g2), g2.shape) == (h // 2, int(0.75 * w))


def test_normal2d_one_dimensional_pseudo_inverse():
    """A 1-D distribution along d: with C = pinv(d d^T) the value is constant
    along the perpendicular of d through the centre."""
    d = np.array([0.12, 0.12])
    c = np.linalg.pinv(np.outer(d, d))
    p = [0.0, 0.0, c[0, 0], c[0, 1], c[1, 0], c[1, 1], 2 * math.pi * d @ d, 0.0]
    g = PL.normal2d(101, 101, p)[..., 0]
    anti = g[np.arange(101), 100 - np.arange(101)]  # the perpendicular diagonal
    assert np.allclose(anti, anti[50], rtol=1e-12)
    assert np.isclose(g.max(), 1.0 / math.sqrt(p[6]))


def _lumaa16_hash_image(values):
    from oracle import blockhash as B
    n = S.stage_encode(values.reshape(-1, 4).astype(np.float32), 11, 5, 5)
    return B.lumaa16_to_rgba8(n.astype("<u4").tobytes(), values.shape[1], values.shape[0])


def test_normal2d_reproduces_blessed_hashes():
    """blend.rs:226-311: the normal distributions stored to 400x400 sRGB
    LumaA16 (a staged format: sRGB transfer, truncating 16-bit store) hash to
    the blessed ``distribution_normal2d`` and ``distribution_normal1d`` lines;
    without the transfer the 1-D one does not.  The Luma8 image of
    ``run_distribution_u8`` hashes to its blessed line only when the 8-bit
    store rounds, not with the truncating store of stage.frag's ``mux_uint``
    (a documented divergence of the blessed data from the shader source)."""
    import json, os
    import zosimos_amd as z
    from oracle import blockhash as B
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    p2 = z.DistributionNormal2d.with_diagonal(0.2, 0.2).params()
    p1 = z.DistributionNormal2d.with_direction([0.04998, 0.0501]).params()
    v2 = PL.normal2d(400, 400, p2)
    v1 = PL.normal2d(400, 400, p1)
    assert B.blockhash256(_lumaa16_hash_image(v2)) in hashes["distribution_normal2d"]
    assert B.blockhash256(_lumaa16_hash_image(v1)) in hashes["distribution_normal1d"]
    lin = np.floor(np.clip(v1[..., 0], 0, 1) * 65535).astype(np.uint32)
    raw = np.stack([lin, np.full_like(lin, 65535)], -1).astype("<u2").tobytes()
    assert B.blockhash256(B.lumaa16_to_rgba8(raw, 400, 400)) not in hashes["distribution_normal1d"]
    enc = C.oe_srgb(np.clip(v1[..., 0], 0, 1))
    for store, blessed in ((np.floor, False), (np.rint, True)):
        c = store(enc * 255).astype(np.uint8)
        got = B.blockhash256(np.stack([c, c, c, np.full_like(c, 255)], -1))
        assert (got in hashes["distribution_u8"]) == blessed
    trunc = S.stage_encode(v1.reshape(-1, 4).astype(np.float32), 0, 4, 5).astype(np.uint8).reshape(400, 400)
    assert np.array_equal(trunc, np.floor(enc * 255).astype(np.uint8)) or \
        np.abs(trunc.astype(int) - np.floor(enc * 255)).max() <= 1


# ---- fractal_noise.frag ---------------------------------------------------------------
def test_pcg4d_is_a_bijection_of_its_inputs():
    """pcg4d is a composition of invertible steps (LCG, cross multiply-adds,
    xorshift): undoing them in reverse order returns the inputs."""
    rng = np.random.default_rng(7)
    v = [rng.integers(0, 2 ** 32, 1000, dtype=np.uint64).astype(np.uint32) for _ in range(4)]
    h = PL.pcg4d(*v)
    m = np.uint32
    with np.errstate(over="ignore"):
        def uncross(u):
            x, y, z, w = u
            w = w - y * z
            z = z - x * y
            y = y - z * x
            x = x - y * w
            return [x, y, z, w]

        def unshift(a):  # inverse of a ^ (a >> 16)
            return a ^ (a >> m(16))

        u = uncross(h)
        u = [unshift(a) for a in u]
        u = uncross(u)
        inv = m(pow(1664525, -1, 2 ** 32))
        u = [(a - m(1013904223)) * inv for a in u]
    for a, b in zip(u, v):
        assert np.array_equal(a, b)


def test_fractal_noise_range_and_blessed_hash():
    """blend.rs:314-338 (run_fractal_noise): four octaves into 400x400 RGBA8
    sRGB hash to the blessed ``distribution_fractal2d`` line (the hash pins
    the pcg4d constants and the cell blend); with amplitudes summing to 1 the
    noise is a convex combination of corner values in [0, 1]."""
    import json, os
    from oracle import blockhash as B
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    v = PL.fractal_noise(400, 400, 4, 0.25, 1.0)
    assert v.min() >= 0.0 and v.max() <= 1.0
    assert B.blockhash256(C.encode_u8(v, 5)) in hashes["distribution_fractal2d"]
    # One octave with one cell per pixel samples each cell at its centre,
    # where the smoothstep weights are 1/2: the value is the mean of the
    # four corner hashes.
    w = PL.fractal_noise(8, 6, 1, 1.0, 1.0, scale=(8.0, 6.0))
    for x, y in ((0, 0), (3, 4), (7, 5)):
        hs = [np.stack(PL.pcg4d([x + i], [y + j], [0], [0]))[:, 0].astype(np.float64) / 2.0 ** 32
              for i, j in ((0, 0), (1, 0), (0, 1), (1, 1))]
        assert np.allclose(w[y, x], np.mean(hs, axis=0), atol=1e-6)


# ---- srlab2.frag ----------------------------------------------------------------------
def test_srlab2_non_linearity_is_continuous_and_inverted():
    """The two branches meet at 216/24389 -> 0.08 (6^3/29^3 and 1.16 * 6/29 -
    0.16), and the inverse undoes the forward map on both sides."""
    k = 216.0 / 24389.0
    assert np.isclose(k * 24389.0 / 2700.0, 0.08) and np.isclose(1.16 * np.cbrt(k) - 0.16, 0.08)
    v = np.concatenate([np.linspace(-0.5, 1.5, 2001), [k * (1 - 1e-9), k * (1 + 1e-9)]])
    assert np.allclose(C.srlab_nl_inv(C.srlab_nl(v[v > -k])), v[v > -k], atol=1e-12)


def test_srlab2_decode_of_encode_is_the_white_point_scaling():
    """decode(encode(rgb)) is the linear map T^-1 CAT02^-1 diag(CAT02 w)
    CAT02 T (T: RGB -> XYZ, w: the SrLab2 white point): the decode scales the
    cone responses by the white's and the encode does not divide them out
    (srlab2.frag:37, 64)."""
    rng = np.random.default_rng(11)
    rgb = rng.uniform(0.05, 0.6, (200, 3))
    T = C.rgb_to_xyz_matrix(3, 5)
    M = np.linalg.inv(T) @ np.linalg.inv(C.CAT02) @ np.diag(C.CAT02 @ C.whitepoint_xyz(5)) @ C.CAT02 @ T
    got = C.srlab2_to_rgb(C.rgb_to_srlab2(rgb), clamp=False)
    assert np.allclose(got, rgb @ M.T, atol=1e-12)
    # L is the XYZ' luminance: a scaled grey stays on the grey axis of XYZ'
    lab = C.rgb_to_srlab2(np.array([[0.2, 0.2, 0.2], [0.7, 0.7, 0.7]]))
    assert lab[1, 0] > lab[0, 0]


def test_srlab2_lch_plane_reproduces_blessed_hash():
    """blend.rs:510-575 (run_srlab2): the bilinear grid of run_oklab stored to
    RGBA8 with a Scalars colour (truncating staged store), reinterpreted as
    SrLab2 LchA and converted to sRGB hashes to the blessed line; the decode
    without the white point's cone scaling does not."""
    import json, os
    from oracle import blockhash as B
    hashes = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "reference_hashes.json")))
    lab = C.decode_u8(PL.staged_u8_store(PL.bilinear_grid(400, 400, OKLAB_PLANE)), 0x100)
    out = lab.copy()
    out[..., :3] = C.srlab2_to_rgb(lab[..., :3])
    assert B.blockhash256(C.encode_u8(out, 5)) in hashes["srlab2"]
    xyz_e = lab[..., :3] @ np.linalg.inv(C.SRLAB_DIFF).T
    xyz = C.srlab_nl_inv(xyz_e @ C.HPE.T) @ np.linalg.inv(C.HPE).T
    out[..., :3] = np.clip(xyz @ np.linalg.inv(C.rgb_to_xyz_matrix(3, 5)).T, 0, 1)
    assert B.blockhash256(C.encode_u8(out, 5)) not in hashes["srlab2"]


def test_channel_swap_reproduces_blessed_hash():
    """blend.rs:423-447 (run_swap): extracting R and G of the background and
    injecting each into the other's channel swaps them; the swapped image
    hashes to the blessed ``swapped`` line.  The hash sums r + g + b, so it
    cannot tell the channel order: the swap itself is pinned byte for byte
    on the device (test_gpu_reference.py)."""
    import json, os
    from PIL import Image
    from oracle import blockhash as B
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    hashes = json.load(open(os.path.join(here, "reference_hashes.json")))
    bg = np.asarray(Image.open(os.path.join(here, "background.png")).convert("RGBA"))
    assert B.blockhash256(bg[..., [1, 0, 2, 3]]) in hashes["swapped"]


def test_generic_palette_reproduces_blessed_hash():
    """tests/generic.rs: the background sampled as a palette at the 2048x2048
    ramp u v stored in an RGBA8 Scalars texel (truncating staged store; the
    ramp in IEEE f32 as bilinear.frag computes it), height from R, width
    from G.  The oracle's image hashes to a blessed line exactly; rounding the
    store instead of truncating it does not."""
    import json
    import os
    from PIL import Image
    from oracle import blockhash as B
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    blessed = json.load(open(os.path.join(golden, "reference_hashes.json")))["generic"]
    bg = np.asarray(Image.open(os.path.join(golden, "background.png")).convert("RGBA"))
    f32 = np.float32
    u = (np.arange(2048, dtype=f32) + f32(0.5)) / f32(2048)
    uv = (u[:, None] * u[None, :]).astype(f32)

    def image(codes):
        ramp = np.zeros((2048, 2048, 4), f32)
        ramp[..., 0] = ramp[..., 1] = codes.astype(f32) / f32(255.0)
        return PL.palette_lookup(bg, ramp, [0, 1, 0, 0], [1, 0, 0, 0])

    assert B.blockhash256(image((uv * f32(255.0)).astype(np.uint32))) in blessed
    assert B.blockhash256(image(np.rint(uv * f32(255.0)).astype(np.uint32))) not in blessed


# ---- the reference's extension shaders (oracle/shaders.py) --------------------
def _blessed(name):
    import json
    import os
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return json.load(open(os.path.join(golden, "reference_hashes.json")))[name]


def _mandelbrot_srgb(lab):
    """The shader's Lab stored in the test's Oklab LchA 8-bit register, then
    converted to sRGB (custom.rs:76-90)."""
    u8 = S.stage_encode(lab.reshape(-1, 4), 10, 20, 0x100).view(np.uint8).reshape(lab.shape[0], lab.shape[1], 4)
    return PL.oklab_lch8_to_srgb(u8)[1]


def test_mandelbrot_reproduces_blessed_hash():
    """custom.rs:14-96: a 256x256 construct_dynamic at scale (3, 3), position
    (0.6, 0.5).  Swapping a and b or flipping rows leaves the blessed set."""
    from oracle import blockhash as B
    from oracle import shaders as SH
    lab = SH.mandelbrot(256, 256)
    assert B.blockhash256(_mandelbrot_srgb(lab)) in _blessed("mandelbrot")
    assert B.blockhash256(_mandelbrot_srgb(lab[..., [0, 2, 1, 3]].copy())) not in _blessed("mandelbrot")
    assert B.blockhash256(_mandelbrot_srgb(lab[::-1].copy())) not in _blessed("mandelbrot")


def test_mandelbrot_known_orbits():
    """Orbits with closed forms, at pixel centres chosen so that c is exact
    (u = 0.125 + 0.25 k, c = (u - position) * 8): c = -1 cycles 0, -1, so
    |z_2048| = 0, light 0.7, mean -1/2 and a = (-1/2 + 1) / 2 = 1/4;
    c = -2 lands on the fixed point 2 after two steps, light 0 and
    a = ((-2 + 2 * 2046) / 2048 + 2) / 2 exactly; c = 1 escapes (light 0)."""
    from oracle import shaders as SH
    lab = SH.mandelbrot(4, 4, scale=(8.0, 8.0), position=(0.25, 0.375))
    # row 1: v = 0.375, c.y = 0; columns: c.x = -1, 1, 3, 5
    assert tuple(lab[1, 0]) == (np.float32(0.7), 0.25, 0.0, 1.0)
    assert lab[1, 1, 0] == 0.0
    lab2 = SH.mandelbrot(4, 4, scale=(8.0, 8.0), position=(0.375, 0.375))
    # column 0: c.x = (0.125 - 0.375) * 8 = -2
    assert lab2[1, 0, 0] == 0.0 and lab2[1, 0, 1] == np.float32((4090.0 / 2048.0 + 2.0) * 0.5) and lab2[1, 0, 2] == 0
    # c = 0 (pixel 1 of position 0.375): z stays 0, colour (0.7, 0, 0, 1)
    assert tuple(lab2[1, 1]) == (np.float32(0.7), 0.0, 0.0, 1.0)


def test_crt_reproduces_blessed_hash_and_mask_texels():
    """custom.rs:98-185: the background magnified 3x through the shadow mask.
    On a constant grey the mask's texels follow the 3 x 6 staggered grid:
    (0, 0) cell 0 -> 0; (0, 2) cell 2 -> red 0.96; (4, 0) cell 3 -> green
    0.96; (4, 2) cell 5 -> 0; (2, 1) cell 1 -> blue 0.64; alpha kept."""
    from oracle import blockhash as B
    from oracle import shaders as SH
    bg = _golden_rgba("background")
    lin = C.decode_u8(bg, 5).astype(np.float32)
    out = SH.crt(lin, 3 * bg.shape[1], 3 * bg.shape[0])
    assert B.blockhash256(C.encode_u8(out, 5)) in _blessed("crt")
    grey = np.full((4, 4, 4), 0.5, np.float32)
    m = SH.crt(grey, 12, 12)
    f = np.float32
    assert tuple(m[0, 0]) == (0.0, 0.0, 0.0, 0.5)
    assert tuple(m[2, 0]) == (f(0.5) * (f(0.16) * f(6.0)), 0.0, 0.0, 0.5)
    assert tuple(m[0, 4]) == (0.0, f(0.5) * (f(0.16) * f(6.0)), 0.0, 0.5)
    assert tuple(m[2, 4]) == (0.0, 0.0, 0.0, 0.5)
    assert tuple(m[1, 2]) == (0.0, 0.0, f(0.5) * (f(0.16) * f(4.0)), 0.5)
    # the operand is sampled nearest: result texel (x, y) reads (x / 3, y / 3)
    ramp = np.zeros((4, 4, 4), np.float32)
    ramp[..., 3] = np.arange(16, dtype=np.float32).reshape(4, 4)
    assert np.array_equal(SH.crt(ramp, 12, 12)[..., 3], np.repeat(np.repeat(ramp[..., 3], 3, 0), 3, 1))


def test_flat_field_reproduces_blessed_hash():
    """custom.rs:187-285: the background corrected by 8-octave fractal noise
    (amplitude 0.1, damping 0.2, grid 10) stored as Luma8 sRGB, mean 0.062.
    Multiplying by the field instead of dividing, or transposing it, leaves
    the blessed set."""
    from oracle import blockhash as B
    from oracle import shaders as SH
    bg = _golden_rgba("background")
    lin = C.decode_u8(bg, 5).astype(np.float32)
    noise = PL.fractal_noise(512, 512, 8, 0.1, 0.2, scale=(10.0, 10.0))
    field = S.stage_decode(S.stage_encode(noise.reshape(-1, 4), 0, 4, 5), 0, 4, 5).reshape(512, 512, 4)
    assert B.blockhash256(C.encode_u8(SH.flat_field(lin, field, 0.124 / 2.0), 5)) in _blessed("flat_field")
    wrong = lin.copy()
    wrong[..., :3] = lin[..., :3] * field[..., :1] / np.float32(0.062)
    assert B.blockhash256(C.encode_u8(wrong, 5)) not in _blessed("flat_field")
    assert B.blockhash256(C.encode_u8(SH.flat_field(lin, field.transpose(1, 0, 2).copy(), 0.062), 5)) \
        not in _blessed("flat_field")
    # a field of exactly the mean leaves the image unchanged
    flat = np.full_like(field, np.float32(0.062))
    assert np.array_equal(SH.flat_field(lin, flat, 0.062), lin)

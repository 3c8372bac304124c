"""GPU parity of the texel codecs that the workloads do not exercise: the
non-sRGB transfer functions, native BGRA8, the wide 16-bit layouts and the
rarer packed layouts of stage.frag.  Every call goes through the C ABI.

Bars:
- packed layouts with a linear transfer: decode bit-exact against the FP32
  restatement of stage.frag (``oracle/stage_frag.c``); encode of the decoded
  values bit-exact against the restatement's truncating encode;
- transfers: decoded values within the accuracy of two ``pow``
  implementations (device ``powf`` against glibc ``powf``; the bar is written
  per transfer), and 8/16-bit encodes equal to the restatement's code, or one
  code off only where the restatement's value sits within 1e-4 of a code
  edge;
- native BGRA8: decode equal to the FP64 sRGB EOTF rounded once to f32 (sRGB)
  or to k * fl(1/255) (linear channels and alpha); encode equal to the FP64 oracle's
  rounded code for every value; BGRA8 <-> RGBA8 a byte permutation;
- 16x3 / 16x4 (above 32 bits, where stage.frag has no codec): IEEE
  v / 65535 and truncating ``uint(c * 65535)``, every code round-trips.
"""
import numpy as np
import pytest

import zosimos_amd as z
from oracle import color as C
from oracle import stage as S

from gpu_util import report, run

pytestmark = pytest.mark.gpu

SB, SP, T = z.SampleBits, z.SampleParts, z.Transfer
F32 = z.Texel(SB.Float32x4, SP.RgbA)


def _f32_desc(w, h):
    return z.Descriptor(w, h, F32, z.Color(transfer=T.Linear))


def _codes(nbits, n, seed):
    if nbits <= 16:
        return np.arange(1 << nbits, dtype=np.uint32)
    v = np.random.default_rng(seed).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    return v & np.uint32((1 << nbits) - 1) if nbits < 32 else v


def _raw(codes, bpp, w):
    n = codes.size
    return codes.astype(np.uint32).view(np.uint8).reshape(n, 4)[:, :bpp].reshape(n // w, w * bpp)


def decode_on_gpu(ctx, codes, bits, parts, transfer, w=256):
    """Staged texels -> RGBA32F linear through one conversion."""
    bpp = z.BITS_BYTES[bits]
    h = codes.size // w
    src = z.Descriptor(w, h, z.Texel(bits, parts), z.Color(transfer=transfer))
    cmd = z.CommandBuffer()
    a = cmd.input(src)
    o, od = cmd.output(cmd.color_convert(a, z.Color(transfer=T.Linear), F32))
    out, _ = run(ctx, cmd, {a: src}, {a: _raw(codes, bpp, w)}, {o: od})
    return out[o].view(np.float32).reshape(-1, 4)


def encode_on_gpu(ctx, values, bits, parts, transfer, w=256):
    """RGBA32F linear -> staged texels; returns the u32 texel values."""
    n = values.shape[0]
    h = n // w
    bpp = z.BITS_BYTES[bits]
    dst = z.Descriptor(w, h, z.Texel(bits, parts), z.Color(transfer=transfer))
    f32 = _f32_desc(w, h)
    cmd = z.CommandBuffer()
    a = cmd.input(f32)
    o, _ = cmd.output(cmd.color_convert(a, dst.color, dst.texel))
    out, _ = run(ctx, cmd, {a: f32}, {a: np.ascontiguousarray(values, np.float32).view(np.uint8).reshape(h, w * 16)},
                 {o: dst})
    raw = out[o].reshape(n, bpp)
    pad = np.zeros((n, 4), np.uint8)
    pad[:, :bpp] = raw[:, :min(bpp, 4)]
    return pad.view(np.uint32).ravel()


# ---- packed layouts (stage.frag:533-641) ------------------------------------------------
@pytest.mark.parametrize("bits,parts,nbits", [
    (SB.UInt233, SP.Rgb, 8),
    (5, SP.Rgb, 16),    # Inti444
    (6, SP.Rgb, 16),    # Int444i (its demux reads the second field at bit 9, stage.frag:547)
    (SB.UInt8x3, SP.Rgb, 24),
    (SB.UInt2101010, SP.RgbA, 32),
    (16, SP.Rgb, 32),   # Int101010i
    (17, SP.Rgb, 32),   # Inti101010
    (SB.UInt16x2, SP.LumaA, 32),
])
def test_packed_layout_decode_and_encode(ctx, bits, parts, nbits):
    codes = _codes(nbits, 65536, bits)
    got = decode_on_gpu(ctx, codes, bits, parts, T.Linear)
    want = S.stage_decode(codes, bits, parts, T.Linear)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"decode of bits {bits}"
    enc = encode_on_gpu(ctx, got, bits, parts, T.Linear)
    want_enc = S.stage_encode(want, bits, parts, T.Linear)
    mask = np.uint32(0xffffffff) if z.BITS_BYTES[bits] >= 4 else np.uint32((1 << (8 * z.BITS_BYTES[bits])) - 1)
    assert np.array_equal(enc & mask, want_enc & mask), f"encode of bits {bits}"
    # Layouts whose fields do not overlap round-trip every code.
    if bits not in (6,):
        keep = want_enc & mask
        report(f"layout_{bits}", round_trip=bool(np.array_equal(keep, codes & mask)))


# ---- 16x3 / 16x4: the build's 16-bit unorm codec above 32 bits ---------------------------
@pytest.mark.parametrize("bits,parts", [(SB.UInt16x3, SP.Rgb), (SB.UInt16x4, SP.RgbA)])
def test_wide_16bit_layouts_round_trip(ctx, bits, parts):
    ch = 3 if bits == SB.UInt16x3 else 4
    bpp = 2 * ch
    w, h = 256, 256
    rng = np.random.default_rng(bits)
    vals = np.stack([np.arange(65536, dtype=np.uint16)] + [rng.permutation(65536).astype(np.uint16)
                                                           for _ in range(ch - 1)], axis=-1)
    raw = vals.view(np.uint8).reshape(h, w * bpp)
    src = z.Descriptor(w, h, z.Texel(bits, parts), z.Color(transfer=T.Linear))
    cmd = z.CommandBuffer()
    a = cmd.input(src)
    o, od = cmd.output(cmd.color_convert(a, z.Color(transfer=T.Linear), F32))
    out, _ = run(ctx, cmd, {a: src}, {a: raw}, {o: od})
    got = out[o].view(np.float32).reshape(-1, 4)
    want = vals.astype(np.float32) / np.float32(65535.0)
    assert np.array_equal(got[:, :ch], want)
    if ch == 3:
        assert np.all(got[:, 3] == 1.0)
    # encode back: every code returns (truncation of an exact quotient)
    cmd = z.CommandBuffer()
    f32 = _f32_desc(w, h)
    a = cmd.input(f32)
    o, _ = cmd.output(cmd.color_convert(a, src.color, src.texel))
    out, _ = run(ctx, cmd, {a: f32}, {a: got.view(np.uint8).reshape(h, w * 16)}, {o: src})
    assert np.array_equal(out[o], raw)
    # truncating encode of arbitrary values, clamped to [0, 1]
    v = rng.uniform(-0.2, 1.2, (w * h, 4)).astype(np.float32)
    cmd = z.CommandBuffer()
    a = cmd.input(f32)
    o, _ = cmd.output(cmd.color_convert(a, src.color, src.texel))
    out, _ = run(ctx, cmd, {a: f32}, {a: v.view(np.uint8).reshape(h, w * 16)}, {o: src})
    exp = (np.clip(v, 0, 1) * np.float32(65535.0)).astype(np.uint32).astype(np.uint16)[:, :ch]
    assert np.array_equal(out[o].view(np.uint16).reshape(-1, ch), exp)


# ---- transfer functions (stage.frag:279-425) --------------------------------------------
# Relative and absolute bars of the decoded values against the FP32
# restatement: a single pow differs by a few ulps between the two libraries;
# the PQ EOTF raises a difference of two nearly equal numbers (n - c1) to the
# power 1/m1 = 6.28, so its relative bar is wider, and where its result is
# tiny the absolute bar applies.
TRANSFERS = {
    T.Bt709: (4e-6, 1e-7),
    T.Bt470M: (4e-6, 1e-7),
    T.Bt601: (4e-6, 1e-7),
    T.Smpte240: (4e-6, 1e-7),
    T.Bt2020_10bit: (4e-6, 1e-7),
    T.Bt2020_12bit: (4e-6, 1e-7),
    T.Smpte2084: (2e-4, 1e-6),
}


@pytest.mark.parametrize("transfer", list(TRANSFERS))
def test_transfer_decode_u16_exhaustive(ctx, transfer):
    """Every 16-bit Luma code through the transfer's EOTF (parts_untransfer,
    stage.frag:781-809)."""
    codes = np.arange(65536, dtype=np.uint32)
    got = decode_on_gpu(ctx, codes, SB.UInt16, SP.Luma, transfer)
    want = S.stage_decode(codes, SB.UInt16, SP.Luma, transfer)
    rtol, atol = TRANSFERS[transfer]
    err = _assert_close(got, want, rtol, atol)
    fp64 = C.transfer_decode(np.repeat((codes / 65535.0)[:, None], 4, axis=1), transfer)[:, 0]
    report(f"transfer_decode_{transfer}", max_abs=f"{err:.2e}",
           max_abs_vs_fp64=f"{np.nanmax(np.abs(got[:, 0] - fp64)):.2e}",
           nan_codes=int(np.isnan(got[:, 0]).sum()))
    assert np.all(got[:, 3] == 1.0)


@pytest.mark.parametrize("transfer", list(TRANSFERS))
def test_transfer_decode_rgba8_staged(ctx, transfer):
    """RGBA8 with a transfer other than sRGB or linear is a staged format
    (program.rs:794-838): random texels, alpha untouched by the transfer."""
    codes = _codes(32, 65536, 40 + transfer)
    got = decode_on_gpu(ctx, codes, SB.UInt8x4, SP.RgbA, transfer)
    want = S.stage_decode(codes, SB.UInt8x4, SP.RgbA, transfer)
    rtol, atol = TRANSFERS[transfer]
    _assert_close(got, want, rtol, atol)
    assert np.array_equal(got[:, 3], ((codes >> 24).astype(np.float32) / np.float32(255.0)))


def _encode_near_edge(values, transfer, nbits, channels):
    """Where the restatement's electrical value e times 2^n - 1 lies within
    the pow libraries' disagreement (4e-6 relative, plus 1e-4) of an integer,
    i.e. of a truncation edge; per texel and channel."""
    lib = S.lib()
    scale = float((1 << nbits) - 1)
    near = np.zeros((values.shape[0], channels), bool)
    for c in range(channels):
        e = np.array([lib.zso_transfer_encode1(float(x), transfer) for x in values[:, c]], np.float64)
        q = np.clip(e, 0, 1) * scale
        near[:, c] = np.abs(q - np.round(q)) <= 4e-6 * q + 1e-4
    return near


def _assert_close(got, want, rtol, atol):
    """Equal NaN positions (stage.frag's pow of a negative base, e.g. the
    SMPTE-240 EOTF as written between 0.0913 and 0.1115), finite values
    within the bar."""
    got = got.astype(np.float64)
    nan_g, nan_w = np.isnan(got), np.isnan(want)
    assert np.array_equal(nan_g, nan_w), f"{int((nan_g != nan_w).sum())} NaN positions differ"
    fin = ~nan_w
    err = np.abs(got[fin] - want[fin])
    assert np.all(err <= rtol * np.abs(want[fin]) + atol), f"max error {err.max():.3e}"
    return float(err.max(initial=0.0))


@pytest.mark.parametrize("transfer", list(TRANSFERS))
def test_transfer_encode_u8_u16(ctx, transfer):
    """Linear values in [0, 1] through the transfer's OETF and the truncating
    store (parts_transfer + mux_uint, stage.frag:750-778, 590-641)."""
    rng = np.random.default_rng(60 + transfer)
    n = 256 * 32
    v = rng.random((n, 4)).astype(np.float32)
    # 16-bit Luma (one channel)
    got = encode_on_gpu(ctx, v, SB.UInt16, SP.Luma, transfer) & np.uint32(0xffff)
    want = S.stage_encode(v, SB.UInt16, SP.Luma, transfer) & np.uint32(0xffff)
    diff = np.abs(got.astype(np.int64) - want.astype(np.int64))
    bad = diff != 0
    if bad.any():
        near = _encode_near_edge(v[bad], transfer, 16, 1)[:, 0]
        assert np.all(diff[bad] == 1) and near.all(), f"{int(bad.sum())} codes differ away from an edge"
    # RGBA8 staged (three transfer channels, alpha linear)
    got8 = encode_on_gpu(ctx, v, SB.UInt8x4, SP.RgbA, transfer)
    want8 = S.stage_encode(v, SB.UInt8x4, SP.RgbA, transfer)
    g = got8.view(np.uint8).reshape(n, 4).astype(np.int64)
    w_ = want8.view(np.uint8).reshape(n, 4).astype(np.int64)
    d8 = np.abs(g - w_)
    assert np.array_equal(g[:, 3], w_[:, 3])
    rows = np.nonzero(d8[:, :3].any(axis=1))[0]
    if rows.size:
        near = _encode_near_edge(v[rows], transfer, 8, 3)
        sub = d8[rows, :3]
        assert np.all(sub <= 1) and np.all(near[sub == 1])
    report(f"transfer_encode_{transfer}", u16_mismatch=int(bad.sum()), u8_rows_mismatch=int(rows.size))


def test_bt601_pair_is_swapped(ctx):
    """stage.frag:307-315 defines BT.601's EOTF as BT.709's OETF and the
    reverse; the device decode of BT.601 equals the FP64 BT.709 OETF."""
    codes = np.arange(65536, dtype=np.uint32)
    got = decode_on_gpu(ctx, codes, SB.UInt16, SP.Luma, T.Bt601)[:, 0]
    _assert_close(got, C.oe_bt709(codes / 65535.0), 4e-6, 1e-7)


# ---- native BGRA8 (program.rs:806-838) -----------------------------------------------
def _bgra(transfer):
    return z.Texel(SB.UInt8x4, SP.BgrA), z.Color(transfer=transfer)


@pytest.mark.parametrize("transfer", [T.Srgb, T.Linear])
def test_native_bgra8_decode_encode(ctx, transfer):
    w, h = 256, 64
    tex, col = _bgra(transfer)
    src = z.Descriptor(w, h, tex, col)
    rng = np.random.default_rng(70 + transfer)
    raw = rng.integers(0, 256, (h, w * 4), dtype=np.uint8)
    raw[0, :1024] = np.repeat(np.arange(256, dtype=np.uint8), 4)  # every code in every channel
    cmd = z.CommandBuffer()
    a = cmd.input(src)
    o, od = cmd.output(cmd.color_convert(a, z.Color(transfer=T.Linear), F32))
    out, _ = run(ctx, cmd, {a: src}, {a: raw}, {o: od})
    got = out[o].view(np.float32).reshape(h, w, 4)
    bgra = raw.reshape(h, w, 4)
    rgba = bgra[..., [2, 1, 0, 3]]
    lin = C.decode_u8(rgba, transfer)
    # Linear channels and alpha decode as k * fl(1/255) in f32 (two
    # roundings: within 2^-23 relative of the FP64 k / 255); sRGB colour
    # through the exact 8-bit EOTF table (the FP64 value rounded once).
    unorm = rgba.astype(np.float32) * np.float32(1.0 / 255.0)
    assert np.all(np.abs(unorm.astype(np.float64) - rgba / 255.0) <= 2.0 ** -23 * (rgba / 255.0))
    if transfer == T.Srgb:
        assert np.array_equal(got[..., :3], lin[..., :3].astype(np.float32))
    else:
        assert np.array_equal(got[..., :3], unorm[..., :3])
    assert np.array_equal(got[..., 3], unorm[..., 3])

    # encode: random linear values (some outside [0, 1]) to BGRA8
    v = rng.uniform(-0.1, 1.1, (h, w, 4)).astype(np.float32)
    f32 = _f32_desc(w, h)
    cmd = z.CommandBuffer()
    a = cmd.input(f32)
    o, _ = cmd.output(cmd.color_convert(a, col, tex))
    out, _ = run(ctx, cmd, {a: f32}, {a: v.view(np.uint8).reshape(h, w * 16)}, {o: src})
    gotb = out[o].reshape(h, w, 4)
    if transfer == T.Srgb:
        want_rgb = C.encode_u8(v.astype(np.float64), T.Srgb)[..., :3]
    else:
        want_rgb = np.rint(np.clip(v[..., :3], 0, 1) * np.float32(255.0)).astype(np.uint8)
    want_a = np.rint(np.clip(v[..., 3], 0, 1) * np.float32(255.0)).astype(np.uint8)
    assert np.array_equal(gotb[..., [2, 1, 0]], want_rgb)
    assert np.array_equal(gotb[..., 3], want_a)

    # round trip BGRA8 -> f32 -> BGRA8 returns every byte
    cmd = z.CommandBuffer()
    a = cmd.input(f32)
    o, _ = cmd.output(cmd.color_convert(a, col, tex))
    out, _ = run(ctx, cmd, {a: f32}, {a: got.view(np.uint8).reshape(h, w * 16)}, {o: src})
    assert np.array_equal(out[o], raw)


@pytest.mark.parametrize("transfer", [T.Srgb, T.Linear])
def test_bgra8_to_rgba8_is_a_byte_permutation(ctx, transfer):
    w, h = 130, 33
    tex, col = _bgra(transfer)
    src = z.Descriptor(w, h, tex, col)
    dst = z.Descriptor(w, h, z.Texel(SB.UInt8x4, SP.RgbA), col)
    raw = np.random.default_rng(80).integers(0, 256, (h, w * 4), dtype=np.uint8)
    cmd = z.CommandBuffer()
    a = cmd.input(src)
    o, _ = cmd.output(cmd.color_convert(a, col, dst.texel))
    out, _ = run(ctx, cmd, {a: src}, {a: raw}, {o: dst})
    assert np.array_equal(out[o].reshape(h, w, 4), raw.reshape(h, w, 4)[..., [2, 1, 0, 3]])

"""FP64 numpy oracle of every per-pixel operation on the compositing path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Conventions
-----------
* Images are numpy arrays of shape ``(H, W, 4)``; channel order after decode is
  the canonical RGBA of ``parts_normalize`` (``lib/std/src/stage.frag:654-699``).
* Identifiers of transfers, sample parts and sample bits are the numeric ids of
  ``lib/std/src/stage.frag:107-170`` (the std140 ``XyzParameter`` blob of
  ``lib/zosimos/src/shaders/stage.rs:53-61``).
* Every quantity is float64.  Functions that restate an FP32 GLSL expression
  evaluate the same formula in float64; bit-level FP32 behaviour lives in
  ``stage_frag.c``.
"""
from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------------
# Identifier tables (lib/std/src/stage.frag:107-170)
# --------------------------------------------------------------------------
TRANSFER = dict(
    Bt709=0, Bt470M=1, Bt601=2, Smpte240=3, Linear=4, Srgb=5, Bt2020_10bit=6,
    Bt2020_12bit=7, Smpte2084=8, Bt2100Pq=9, Bt2100Hlg=10, LinearScene=11,
    LabLch=0x100,
)
PARTS = dict(
    A=0, R=1, G=2, B=3, Luma=4, LumaA=5, Rgb=6, Bgr=7, Rgba=8, Rgbx=9, Bgra=10,
    Bgrx=11, Argb=12, Xrgb=13, Abgr=14, Xbgr=15, Yuv=16, Lab=17, LabA=18,
    LCh=19, LChA=20,
)
BITS = dict(
    Int8=0, Int332=1, Int233=2, Int16=3, Int4x4=4, Inti444=5, Int444i=6,
    Int565=7, Int8x2=8, Int8x3=9, Int8x4=10, Int16x2=11, Int16x3=12,
    Int16x4=13, Int1010102=14, Int2101010=15, Int101010i=16, Inti101010=17,
    Float16x4=18, Float32x4=19,
)
# Bytes per texel for each sample-bits id (image-canvas SampleBits::bytes()).
BITS_BYTES = {0: 1, 1: 1, 2: 1, 3: 2, 4: 2, 5: 2, 6: 2, 7: 2, 8: 2, 9: 3,
              10: 4, 11: 4, 12: 6, 13: 8, 14: 4, 15: 4, 16: 4, 17: 4,
              18: 8, 19: 16}

# Primaries (order of image-canvas ``Primaries`` as mirrored in
# lib/lang/src/serde_bindings.rs:65-75) plus Display-P3, which the reference
# lacks.  Chromaticities (x, y) of R, G, B from the published standards:
# ITU-R BT.709-6 Table 1, BT.601-7 (525 = SMPTE C, 625 = EBU), SMPTE 240M,
# ITU-R BT.2020-2 / BT.2100-2 Table 2, SMPTE EG 432-1 (Display P3).
PRIMARIES = dict(Xyz=0, Bt601_525=1, Bt601_625=2, Bt709=3, Smpte240=4,
                 Bt2020=5, Bt2100=6, DisplayP3=7)
_CHROMA = {
    1: ((0.630, 0.340), (0.310, 0.595), (0.155, 0.070)),
    2: ((0.640, 0.330), (0.290, 0.600), (0.150, 0.060)),
    3: ((0.640, 0.330), (0.300, 0.600), (0.150, 0.060)),
    4: ((0.630, 0.340), (0.310, 0.595), (0.155, 0.070)),
    5: ((0.708, 0.292), (0.170, 0.797), (0.131, 0.046)),
    6: ((0.708, 0.292), (0.170, 0.797), (0.131, 0.046)),
    7: ((0.680, 0.320), (0.265, 0.690), (0.150, 0.060)),
}
WHITEPOINT = dict(A=0, B=1, C=2, D50=3, D55=4, D65=5, D75=6, E=7, F2=8, F7=9,
                  F11=10)
# CIE 1931 2-degree chromaticities; D65 and D50 use the 4-digit values that
# BT.709 / IEC 61966-2-1 and ICC specify.
_WHITE = {0: (0.44757, 0.40745), 1: (0.34840, 0.35160), 2: (0.31006, 0.31616),
          3: (0.3457, 0.3585), 4: (0.33242, 0.34743), 5: (0.3127, 0.3290),
          6: (0.29902, 0.31485), 7: (1.0 / 3.0, 1.0 / 3.0),
          8: (0.37208, 0.37529), 9: (0.31292, 0.32933), 10: (0.38052, 0.37713)}


def whitepoint_xyz(wp: int) -> np.ndarray:
    x, y = _WHITE[wp]
    return np.array([x / y, 1.0, (1.0 - x - y) / y])


def rgb_to_xyz_matrix(primaries: int, wp: int) -> np.ndarray:
    """RGB->XYZ 3x3 from chromaticities and white point (textbook derivation,
    e.g. SMPTE RP 177).  Replaces image-canvas ``Primaries::to_xyz_row_matrix``
    (called at lib/zosimos/src/command.rs:1023-1024, 1035, 1048), which is an
    un-vendored dependency.  ``Xyz`` primaries give the identity."""
    if primaries == 0:
        return np.eye(3)
    prim = np.array([[x / y, 1.0, (1.0 - x - y) / y] for (x, y) in _CHROMA[primaries]]).T
    s = np.linalg.solve(prim, whitepoint_xyz(wp))
    return prim * s[None, :]


def convert_matrix(p_src: int, p_dst: int, wp: int) -> np.ndarray:
    """Linear RGB(src) -> linear RGB(dst) for the same white point:
    ``P_dst^-1 . P_src``.  The reference writes ``P_dst . P_src^-1``
    (lib/zosimos/src/command.rs:1022-1025, 3223-3229), which is the documented
    intent only if ``to_xyz_row_matrix`` were XYZ->RGB; this oracle follows the
    intent (identical primaries give the identity either way)."""
    return np.linalg.solve(rgb_to_xyz_matrix(p_dst, wp), rgb_to_xyz_matrix(p_src, wp))


# --------------------------------------------------------------------------
# Transfer functions (lib/std/src/stage.frag:279-425, 750-809), FP64
# --------------------------------------------------------------------------
def _pow(x, e):
    with np.errstate(invalid="ignore"):
        return np.power(x, e)


def oe_bt709(v):
    """stage.frag:280-286."""
    return np.where(v >= 0.018, 1.099 * _pow(v, 0.45) - 0.099, 4.5 * v)


def eo_bt709(v):
    """stage.frag:289-295 (threshold is oe_bt709(0.018))."""
    t = 1.099 * 0.018 ** 0.45 - 0.099
    return np.where(v >= t, _pow((v + 0.099) / 1.099, 1.0 / 0.45), v / 4.5)


def oe_srgb(v):
    """stage.frag:335-342 (IEC 61966-2-1, sign-symmetric)."""
    return np.where(v < -0.0031308, -1.055 * _pow(-v, 1.0 / 2.4) + 0.055,
                    np.where(v <= 0.0031308, v * 12.92,
                             1.055 * _pow(v, 1.0 / 2.4) - 0.055))


def eo_srgb(v):
    """stage.frag:345-352."""
    return np.where(v < -0.04045, -_pow((-v + 0.055) / 1.055, 2.4),
                    np.where(v <= 0.04045, v / 12.92,
                             _pow((v + 0.055) / 1.055, 2.4)))


def oe_smpte240(v):
    """stage.frag:318-323."""
    return np.where(v < 0.0228, 4.0 * v, 1.1115 * _pow(v, 0.45) - 0.1115)


def eo_smpte240(v):
    """stage.frag:326-331, kept as written: ``pow((v - 0.1115) / 1.1115, 1/0.45)``
    (the published inverse adds 0.1115; the reference subtracts it)."""
    return np.where(v < 0.0913, v / 4.0, _pow((v - 0.1115) / 1.1115, 1.0 / 0.45))


_PQ_M1 = 2610.0 / 16384.0
_PQ_M2 = 2523.0 / 4096.0
_PQ_C1 = 3424.0 / 4096.0
_PQ_C2 = 2413.0 / 128.0
_PQ_C3 = 2392.0 / 128.0


def eo_smpte2084(v):
    """stage.frag:374-379 (BT.2100-2 Table 4 PQ EOTF)."""
    n = _pow(v, 1.0 / _PQ_M2)
    nom = np.maximum(n - _PQ_C1, 0.0)
    return _pow(nom / (_PQ_C2 - _PQ_C3 * n), 1.0 / _PQ_M1)


def eo_inv_smpte2084(v):
    """stage.frag:382-387."""
    y = _pow(v, _PQ_M1)
    return _pow((_PQ_C1 + _PQ_C2 * y) / (_PQ_C3 * y + 1.0), _PQ_M2)


def oe_smpte2084(v):
    """stage.frag:392-405: scene->display OOTF then inverse EOTF."""
    display = _pow(oe_bt709(59.5208 * v), 2.4) / 100.0
    return eo_inv_smpte2084(display)


def lab_to_lch(lab):
    """stage.frag:415-420: hue scaled to [0, 1] with an offset of 0.5."""
    c = np.hypot(lab[..., 1], lab[..., 2])
    h = np.degrees(np.arctan2(lab[..., 2], lab[..., 1])) / 360.0 + 0.5
    return np.stack([lab[..., 0], c, h], axis=-1)


def lch_to_lab(lch):
    """stage.frag:422-425."""
    ang = np.radians(360.0 * (lch[..., 2] - 0.5))
    return np.stack([lch[..., 0], lch[..., 1] * np.cos(ang), lch[..., 1] * np.sin(ang)], axis=-1)


def transfer_decode(rgba, transfer: int):
    """``parts_untransfer`` (stage.frag:781-809): electrical -> linear on RGB,
    alpha untouched.  Bt601 uses ``transfer_eo_bt601 = transfer_oe_bt709``
    exactly as stage.frag:313-315 writes it; Bt2100Hlg, Bt2100Pq and
    LinearScene fall through to the identity as in the reference."""
    rgba = np.asarray(rgba, dtype=np.float64)
    out = rgba.copy()
    rgb = rgba[..., :3]
    if transfer == 0 or transfer in (6, 7):
        out[..., :3] = eo_bt709(rgb)
    elif transfer == 1:
        out[..., :3] = _pow(rgb, 2.2)
    elif transfer == 2:
        out[..., :3] = oe_bt709(rgb)
    elif transfer == 3:
        out[..., :3] = eo_smpte240(rgb)
    elif transfer == 5:
        out[..., :3] = eo_srgb(rgb)
    elif transfer == 8:
        out[..., :3] = eo_smpte2084(rgb)
    elif transfer == 0x100:
        out[..., :3] = lch_to_lab(rgb)
    return out


def transfer_encode(rgba, transfer: int):
    """``parts_transfer`` (stage.frag:750-778): linear -> electrical."""
    rgba = np.asarray(rgba, dtype=np.float64)
    out = rgba.copy()
    rgb = rgba[..., :3]
    if transfer == 0 or transfer in (6, 7):
        out[..., :3] = oe_bt709(rgb)
    elif transfer == 1:
        out[..., :3] = _pow(rgb, 1.0 / 2.2)
    elif transfer == 2:
        out[..., :3] = eo_bt709(rgb)
    elif transfer == 3:
        out[..., :3] = oe_smpte240(rgb)
    elif transfer == 5:
        out[..., :3] = oe_srgb(rgb)
    elif transfer == 8:
        out[..., :3] = oe_smpte2084(rgb)
    elif transfer == 0x100:
        out[..., :3] = lab_to_lch(rgb)
    return out


# --------------------------------------------------------------------------
# Chromatic adaptation (command.rs:1112-1174, 2527-2548, 3279-3340)
# --------------------------------------------------------------------------
# Cone-space matrices of the methods, as the reference documents them at
# command.rs:422-460 (the ``palette`` crate 0.7.6 computes with them; it is
# not vendored): XYZ scaling, von Kries (Hunt-Pointer-Estevez), Bradford.
ADAPT_CONES = {
    0: np.eye(3),
    1: np.array([[0.40024, 0.70760, -0.08081], [-0.22630, 1.16532, 0.04570], [0.0, 0.0, 0.91822]]),
    2: np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]]),
}


def adaptation_matrix(method: int, src_wp: int, dst_wp: int) -> np.ndarray:
    """von Kries transform in the method's cone space: M^-1 diag(M w_dst /
    M w_src) M (XYZ -> XYZ)."""
    m = ADAPT_CONES[method]
    ratio = (m @ whitepoint_xyz(dst_wp)) / (m @ whitepoint_xyz(src_wp))
    return np.linalg.inv(m) @ np.diag(ratio) @ m


def adapt_rgb_matrix(primaries: int, src_wp: int, dst_wp: int, method: int) -> np.ndarray:
    """The linear RGB matrix of ``chromatic_adaptation`` (command.rs:2527-2531):
    RGB relative to dst_wp <- XYZ <- adaptation <- XYZ <- RGB relative to
    src_wp."""
    return np.linalg.inv(rgb_to_xyz_matrix(primaries, dst_wp)) @ adaptation_matrix(method, src_wp, dst_wp) \
        @ rgb_to_xyz_matrix(primaries, src_wp)


# --------------------------------------------------------------------------
# Oklab (lib/std/src/oklab.frag:14-64; Ottosson 2020)
# --------------------------------------------------------------------------
# Row-major forms of the column-major GLSL constants at oklab.frag:14-24.
OKLAB_M1 = np.array([[0.8189330101, 0.3618667424, -0.1288597137],
                     [0.0329845436, 0.9293118715, 0.0361456387],
                     [0.0482003018, 0.2643662691, 0.6338517070]])
OKLAB_M2 = np.array([[0.2104542553, 0.7936177850, -0.0040720468],
                     [1.9779984951, -2.4285922050, 0.4505937099],
                     [0.0259040371, 0.7827717662, -0.8086757660]])


def xyz_to_oklab(xyz):
    """oklab.frag:40-43: ``Lab = M2 . cbrt(M1 . xyz)`` with a signed cube root."""
    lms = np.einsum("ij,...j->...i", OKLAB_M1, xyz)
    return np.einsum("ij,...j->...i", OKLAB_M2, np.cbrt(lms))


def oklab_to_xyz(lab):
    """oklab.frag:56-59: ``xyz = M1^-1 . (M2^-1 . Lab)^3``."""
    lms_ = np.einsum("ij,...j->...i", np.linalg.inv(OKLAB_M2), lab)
    return np.einsum("ij,...j->...i", np.linalg.inv(OKLAB_M1), lms_ ** 3)


def rgb_to_oklab(rgb, primaries: int = 3):
    """``color_convert`` Rgb(D65) -> Oklab: oklab.frag:34-47 with
    ``xyz_transform = to_xyz_row_matrix`` (command.rs:1027-1037, 3235-3242)."""
    xyz = np.einsum("ij,...j->...i", rgb_to_xyz_matrix(primaries, 5), rgb)
    return xyz_to_oklab(xyz)


def oklab_to_rgb(lab, primaries: int = 3, clamp: bool = True):
    """Oklab -> Rgb(D65): oklab.frag:50-64 including its ``clamp(rgb, 0, 1)``
    (oklab.frag:63)."""
    xyz = oklab_to_xyz(lab)
    rgb = np.linalg.solve(rgb_to_xyz_matrix(primaries, 5), xyz[..., None])[..., 0] \
        if xyz.ndim > 1 else np.linalg.solve(rgb_to_xyz_matrix(primaries, 5), xyz)
    return np.clip(rgb, 0.0, 1.0) if clamp else rgb


# CIECAM02's CAT02 and Hunt-Pointer-Estevez matrices, row-major
# (srlab2.frag:13-23, written column by column there).
CAT02 = np.array([[0.7328, 0.4296, -0.1624], [-0.7036, 1.6975, 0.0061], [0.0030, 0.0136, 0.9834]])
HPE = np.array([[0.38971, 0.68898, -0.07868], [-0.22981, 1.18340, 0.04641], [0.0, 0.0, 1.0]])
# SrLab2's difference scheme: (L, a, b) = (y, (x - y) 5 / 1.16, (z - y) 2 / 1.16).
SRLAB_DIFF = np.array([[0.0, 1.0, 0.0], [5.0 / 1.16, -5.0 / 1.16, 0.0], [0.0, -2.0 / 1.16, 2.0 / 1.16]])


def srlab_nl(v):
    """srlab2.frag:87-94: linear below 216/24389 in magnitude, 1.16 cbrt - 0.16
    above (the cube root taken with its sign for negative values, where the
    shader's pow is undefined)."""
    v = np.asarray(v, dtype=np.float64)
    return np.where(np.abs(v) < 216.0 / 24389.0, v * 24389.0 / 2700.0, 1.16 * np.cbrt(v) - 0.16)


def srlab_nl_inv(v):
    """srlab2.frag:104-111."""
    v = np.asarray(v, dtype=np.float64)
    return np.where(np.abs(v) < 0.08, v * 2700.0 / 24389.0, ((v + 0.16) / 1.16) ** 3)


def rgb_to_srlab2(rgb, primaries: int = 3, rgb_wp: int = 5):
    """srlab2.frag:35-59 (encode): XYZ of the linear RGB, into HPE cone
    space (the CAT02 round trip of the shader cancels), the non-linearity,
    back to XYZ' and the difference scheme.  The SrLab2 white point plays no
    part (its division is commented out in the shader)."""
    xyz = np.asarray(rgb, dtype=np.float64) @ rgb_to_xyz_matrix(primaries, rgb_wp).T
    xyz_e = srlab_nl(xyz @ HPE.T) @ np.linalg.inv(HPE).T
    return xyz_e @ SRLAB_DIFF.T


def srlab2_to_rgb(lab, primaries: int = 3, rgb_wp: int = 5, lab_wp: int = 5, clamp: bool = True):
    """srlab2.frag:62-85 (decode): undo the difference scheme and the
    non-linearity, into CAT02 cone space, scale by the CAT02 response of the
    SrLab2 white point (which the encode does not divide out), back to XYZ
    and RGB, clamped to [0, 1]."""
    xyz_e = np.asarray(lab, dtype=np.float64) @ np.linalg.inv(SRLAB_DIFF).T
    rgb_w = srlab_nl_inv(xyz_e @ HPE.T) @ (CAT02 @ np.linalg.inv(HPE)).T
    xyz = (rgb_w * (CAT02 @ whitepoint_xyz(lab_wp))) @ np.linalg.inv(CAT02).T
    rgb = xyz @ np.linalg.inv(rgb_to_xyz_matrix(primaries, rgb_wp)).T
    return np.clip(rgb, 0.0, 1.0) if clamp else rgb


def lch_adjust(lab, chroma: float = 1.0, hue_deg: float = 0.0):
    """Scale Oklab chroma and rotate hue: ``(a, b) <- k R(theta) (a, b)``.
    No reference implementation exists (the reference needs a user shader for
    it, command/dynamic.rs:7-28); this is the polar form of
    stage.frag:415-425 applied as one 2x2 map on (a, b)."""
    t = np.radians(hue_deg)
    c, s = np.cos(t), np.sin(t)
    a, b = lab[..., 1], lab[..., 2]
    return np.stack([lab[..., 0], chroma * (c * a - s * b), chroma * (s * a + c * b)], axis=-1)


def delta_e_ok(lab1, lab2):
    """Euclidean distance in Oklab (Delta-E OK)."""
    return np.sqrt(np.sum((np.asarray(lab1) - np.asarray(lab2)) ** 2, axis=-1))


# --------------------------------------------------------------------------
# Alpha compositing (Porter & Duff 1984).  Absent in the reference
# (command.rs:1510-1519 returns UNIMPLEMENTED).
# --------------------------------------------------------------------------
def premultiply(rgba):
    out = np.array(rgba, dtype=np.float64)
    out[..., :3] *= out[..., 3:4]
    return out


def unpremultiply(rgba):
    """Inverse of ``premultiply`` for alpha > 0; alpha == 0 gives (0,0,0,0)."""
    out = np.array(rgba, dtype=np.float64)
    a = out[..., 3:4]
    with np.errstate(divide="ignore", invalid="ignore"):
        out[..., :3] = np.where(a > 0.0, out[..., :3] / np.where(a > 0, a, 1.0), 0.0)
    return out


def over(src_premul, dst_premul):
    """Porter-Duff 'over' on premultiplied values: ``S + D (1 - alpha_S)``."""
    return src_premul + dst_premul * (1.0 - src_premul[..., 3:4])


# --------------------------------------------------------------------------
# Affine bilinear resampling with a transparent border.  The reference only
# rasterises a nearest quad (command.rs:1659-1665 rejects BiLinear).
# --------------------------------------------------------------------------
def affine_inverse(m):
    return np.linalg.inv(np.asarray(m, dtype=np.float64).reshape(3, 3))


def bilinear_affine_at(texels, src_hw, matrix, xs, ys):
    """The bilinear affine sample of ``bilinear_affine`` at the output pixels
    (xs[k], ys[k]) only.  ``texels(yi, xi)`` returns the premultiplied
    values of the source texels at integer arrays (yi, xi) (all in range);
    the source is ``src_hw`` = (H, W).  Returns (N, C)."""
    sh, sw = src_hw
    inv = affine_inverse(matrix)
    px = np.asarray(xs, dtype=np.float64) + 0.5
    py = np.asarray(ys, dtype=np.float64) + 0.5
    u = inv[0, 0] * px + inv[0, 1] * py + inv[0, 2]
    v = inv[1, 0] * px + inv[1, 1] * py + inv[1, 2]
    q = inv[2, 0] * px + inv[2, 1] * py + inv[2, 2]
    u, v = u / q - 0.5, v / q - 0.5
    x0, y0 = np.floor(u), np.floor(v)
    fx, fy = (u - x0)[..., None], (v - y0)[..., None]
    x0, y0 = x0.astype(np.int64), y0.astype(np.int64)

    def tap(xi, yi):
        ok = (xi >= 0) & (xi < sw) & (yi >= 0) & (yi < sh)
        vals = texels(np.clip(yi, 0, sh - 1), np.clip(xi, 0, sw - 1))
        return np.where(ok[..., None], vals, 0.0)

    return ((1 - fx) * (1 - fy) * tap(x0, y0) + fx * (1 - fy) * tap(x0 + 1, y0)
            + (1 - fx) * fy * tap(x0, y0 + 1) + fx * fy * tap(x0 + 1, y0 + 1))


def bilinear_affine(src_premul, matrix, out_hw):
    """Resample ``src_premul`` (H, W, C) onto an output grid of ``out_hw``.

    ``matrix`` is the row-major homogeneous transform of ``Affine``
    (command.rs:287-296) mapping source pixel coordinates (top-left corner of
    the first texel at (0, 0)) to output pixel coordinates, as
    ``QuadTarget::affine`` applies it (program.rs:1897-1906).  Output pixel
    (x, y) samples the source at ``M^-1 (x + 0.5, y + 0.5)``; texel (i, j) has
    its centre at (i + 0.5, j + 0.5); samples outside the source read zero.
    """
    h, w = out_hw
    ys, xs = np.mgrid[0:h, 0:w]
    src = np.asarray(src_premul)
    out = bilinear_affine_at(lambda yi, xi: src[yi, xi], src.shape[:2], matrix, xs.ravel(), ys.ravel())
    return out.reshape(h, w, -1)


def affine_matrix(shift_pre=(0.0, 0.0), scale=(1.0, 1.0), rotate_rad=0.0, shift_post=(0.0, 0.0)):
    """Compose ``Affine::new().shift(pre).scale(s).rotate(r).shift(post)`` as
    command.rs:3424-3485 does: each builder left-multiplies."""
    m = np.eye(3)
    m = np.array([[1, 0, shift_pre[0]], [0, 1, shift_pre[1]], [0, 0, 1.0]]) @ m
    m = np.diag([scale[0], scale[1], 1.0]) @ m
    c, s = np.cos(rotate_rad), np.sin(rotate_rad)
    m = np.array([[c, s, 0], [-s, c, 0], [0, 0, 1.0]]) @ m
    m = np.array([[1, 0, shift_post[0]], [0, 1, shift_post[1]], [0, 0, 1.0]]) @ m
    return m


# --------------------------------------------------------------------------
# Texel codecs
# --------------------------------------------------------------------------
def decode_u8(data, transfer: int = 5):
    """Native RGBA8 Unorm / UnormSrgb fetch (program.rs:794-838): ``x / 255``
    then the EOTF on RGB (IEC 61966-2-1 for sRGB); alpha is linear."""
    v = np.asarray(data, dtype=np.float64) / 255.0
    return transfer_decode(v, transfer)


def encode_u8(rgba, transfer: int = 5):
    """Native RGBA8 store: OETF on RGB, clamp to [0, 1], round to nearest
    (the float->unorm rule of the Vulkan/WebGPU render-target store; ties to
    even)."""
    e = transfer_encode(np.asarray(rgba, dtype=np.float64), transfer)
    e = np.nan_to_num(e, nan=0.0)
    return np.rint(np.clip(e, 0.0, 1.0) * 255.0).astype(np.uint8)


def decode_f16(halfs):
    """RGBA16F: IEEE 754 binary16 -> value (exact)."""
    return np.asarray(halfs, dtype=np.float16).astype(np.float64)


def encode_f16(values):
    """value -> IEEE 754 binary16, round to nearest even."""
    return np.asarray(values, dtype=np.float64).astype(np.float16)


def srgb_u8_decode_table():
    """The 256 linear values of an 8-bit sRGB code (FP64)."""
    return eo_srgb(np.arange(256, dtype=np.float64) / 255.0)

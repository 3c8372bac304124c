"""zosimos_amd -- MI355X (gfx950) execution backend for zosimos programs.

Python mirror of the reference's front end (``CommandBuffer``, ``Linker``,
``Program::lower_to``, ``Executable::launch``; lib/zosimos/src/command.rs,
program.rs, run.rs) over the C ABI of ``include/zosimos_hip.h`` and
``include/zosimos_command.h``.  All pixel work runs in the HIP kernels of
``csrc/``; importing the compute classes without the built library raises.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from ._abi import BindingC, ColorC, DescriptorC, PoolBindingC, RectC, Texel as TexelC

# ---- identifiers (lib/std/src/stage.frag:107-170) ---------------------------
class Transfer:
    Bt709 = 0
    Bt470M = 1
    Bt601 = 2
    Smpte240 = 3
    Linear = 4
    Srgb = 5
    Bt2020_10bit = 6
    Bt2020_12bit = 7
    Smpte2084 = 8
    LabLch = 0x100


class SampleParts:
    A, R, G, B, Luma, LumaA, Rgb, Bgr, RgbA, Rgbx, BgrA = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10
    Bgrx, ARgb, Xrgb, ABgr, Xbgr, Yuv, Lab, LabA, Lch, LchA = 11, 12, 13, 14, 15, 16, 17, 18, 19, 20


class SampleBits:
    UInt8, UInt332, UInt233, UInt16, UInt4x4 = 0, 1, 2, 3, 4
    UInt565, UInt8x2, UInt8x3, UInt8x4, UInt16x2, UInt16x3, UInt16x4 = 7, 8, 9, 10, 11, 12, 13
    UInt1010102, UInt2101010, Float16x4, Float32x4 = 14, 15, 18, 19


BITS_BYTES = {0: 1, 1: 1, 2: 1, 3: 2, 4: 2, 5: 2, 6: 2, 7: 2, 8: 2, 9: 3, 10: 4, 11: 4, 12: 6, 13: 8,
              14: 4, 15: 4, 16: 4, 17: 4, 18: 8, 19: 16}


class Primaries:
    Xyz, Bt601_525, Bt601_625, Bt709, Smpte240, Bt2020, Bt2100, DisplayP3 = range(8)


class DerivativeMethod:
    """command.rs:556-566, in declaration order."""
    Roberts, Prewitt, Sobel, Scharr3, Scharr4, Scharr5, Scharr3To4Bit, Scharr5Tab, Scharr3To8Bit = range(9)


class Direction:
    """command.rs:547-552."""
    Height, Width = range(2)


class ChromaticAdaptationMethod:
    """command.rs:422-460, in declaration order."""
    Xyz, VonKries, BradfordVonKries, BradfordNonLinear = range(4)


class Whitepoint:
    A, B, C, D50, D55, D65, D75, E, F2, F7, F11 = range(11)


class ColorModel:
    Rgb, Oklab, Scalars, SrLab2 = 0, 1, 2, 3


class Status:
    OK, INVALID_ARGUMENT, UNSUPPORTED, HIP, MISMATCHED_DESCRIPTOR, MISSING_KEY = 0, 1, 2, 3, 4, 5
    OUT_OF_MEMORY, TYPE, UNIMPLEMENTED, KNOB_LENGTH, NO_DEVICE, OTHER = 6, 7, 8, 9, 10, 11


class ZosimosError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message

    def is_type_err(self) -> bool:
        """CommandError::is_type_err (command.rs:3632-3639)."""
        return self.code == Status.TYPE


def _check(code: int):
    if code != 0:
        raise ZosimosError(code, _abi.lib().zh_last_error().decode())


# ---- descriptors (lib/zosimos/src/buffer.rs:13-30) ---------------------------
@dataclass(frozen=True)
class Color:
    model: int = ColorModel.Rgb
    primaries: int = Primaries.Bt709
    transfer: int = Transfer.Srgb
    whitepoint: int = Whitepoint.D65
    luminance: int = 0

    def c(self) -> ColorC:
        return ColorC(self.model, self.primaries, self.transfer, self.whitepoint, self.luminance)


Color.SRGB = Color()
Color.BT709_RGB = Color(transfer=Transfer.Bt709)
Color.OKLAB = Color(model=ColorModel.Oklab, primaries=0, transfer=Transfer.Linear, whitepoint=Whitepoint.D65)
Color.SRLAB2_D65 = Color(model=ColorModel.SrLab2, primaries=0, transfer=Transfer.Linear, whitepoint=Whitepoint.D65)


@dataclass(frozen=True)
class Texel:
    bits: int
    parts: int

    @staticmethod
    def new_u8(parts: int) -> "Texel":
        return Texel(SampleBits.UInt8x4 if parts in (SampleParts.RgbA, SampleParts.BgrA) else SampleBits.UInt8, parts)

    def c(self) -> TexelC:
        return TexelC(self.bits, self.parts)

    @property
    def bytes(self) -> int:
        return BITS_BYTES[self.bits]


@dataclass(frozen=True)
class Descriptor:
    width: int
    height: int
    texel: Texel
    color: Color = Color.SRGB
    row_stride: int = 0

    @staticmethod
    def rgba8_srgb(w: int, h: int) -> "Descriptor":
        """``Descriptor::with_srgb_image`` of an RGBA8 image (buffer.rs:201-210)."""
        return Descriptor(w, h, Texel(SampleBits.UInt8x4, SampleParts.RgbA), Color.SRGB)

    @staticmethod
    def rgba16f_linear(w: int, h: int, primaries: int = Primaries.Bt709) -> "Descriptor":
        return Descriptor(w, h, Texel(SampleBits.Float16x4, SampleParts.RgbA),
                          Color(primaries=primaries, transfer=Transfer.Linear))

    @staticmethod
    def rgba32f_linear(w: int, h: int, primaries: int = Primaries.Bt709) -> "Descriptor":
        return Descriptor(w, h, Texel(SampleBits.Float32x4, SampleParts.RgbA),
                          Color(primaries=primaries, transfer=Transfer.Linear))

    def with_(self, **kw) -> "Descriptor":
        d = dict(width=self.width, height=self.height, texel=self.texel, color=self.color, row_stride=self.row_stride)
        d.update(kw)
        return Descriptor(**d)

    @property
    def stride(self) -> int:
        return self.row_stride or self.width * self.texel.bytes

    @property
    def nbytes(self) -> int:
        return self.stride * self.height

    def c(self) -> DescriptorC:
        return DescriptorC(self.width, self.height, self.row_stride, self.texel.bytes, self.texel.c(), self.color.c())

    @staticmethod
    def from_c(d: DescriptorC) -> "Descriptor":
        return Descriptor(d.width, d.height, Texel(d.texel.bits, d.texel.parts),
                          Color(d.color.model, d.color.primaries, d.color.transfer, d.color.whitepoint,
                                d.color.luminance), int(d.row_stride))


@dataclass(frozen=True)
class Rectangle:
    """``Rectangle`` (command.rs:279-285)."""
    x: int
    y: int
    max_x: int
    max_y: int

    @staticmethod
    def with_width_height(w: int, h: int) -> "Rectangle":
        return Rectangle(0, 0, w, h)

    def c(self) -> RectC:
        return RectC(self.x, self.y, self.max_x, self.max_y)


class AffineSample:
    Nearest = 0
    BiLinear = 1


class Blend:
    Alpha = 0


def _mul3(a, b):
    return [sum(a[i * 3 + k] * b[k * 3 + j] for k in range(3)) for i in range(3) for j in range(3)]


@dataclass(frozen=True)
class Affine:
    """``Affine`` (command.rs:287-296, builders 3420-3485): a row-major
    homogeneous matrix; each builder left-multiplies."""
    transformation: Tuple[float, ...] = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    sampling: int = AffineSample.Nearest

    @staticmethod
    def new(sampling: int) -> "Affine":
        return Affine(sampling=sampling)

    def _post(self, m) -> "Affine":
        return Affine(tuple(_mul3(m, list(self.transformation))), self.sampling)

    def scale(self, x: float, y: float) -> "Affine":
        return self._post([x, 0, 0, 0, y, 0, 0, 0, 1.0])

    def rotate(self, rad: float) -> "Affine":
        c, s = math.cos(rad), math.sin(rad)
        return self._post([c, s, 0, -s, c, 0, 0, 0, 1.0])

    def shift(self, x: float, y: float) -> "Affine":
        return self._post([1, 0, x, 0, 1, y, 0, 0, 1.0])


# ---- CommandBuffer (command.rs:60-68, 739-1863) ----------------------------------
class CommandBuffer:
    def __init__(self):
        self._lib = _abi.lib()
        h = ctypes.c_void_p()
        _check(self._lib.zs_command_create(ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.zs_command_destroy(self._h)
            self._h = None

    def _reg(self, fn, *args) -> int:
        r = ctypes.c_uint32()
        _check(fn(self._h, *args, ctypes.byref(r)))
        return r.value

    def input(self, desc: Descriptor) -> int:
        return self._reg(self._lib.zs_command_input, ctypes.byref(desc.c()))

    def output(self, src: int):
        r = ctypes.c_uint32()
        d = DescriptorC()
        _check(self._lib.zs_command_output(self._h, src, ctypes.byref(r), ctypes.byref(d)))
        return r.value, Descriptor.from_c(d)

    def render(self, src: int):
        r = ctypes.c_uint32()
        d = DescriptorC()
        _check(self._lib.zs_command_render(self._h, src, ctypes.byref(r), ctypes.byref(d)))
        return r.value, Descriptor.from_c(d)

    def crop(self, src: int, rect: Rectangle) -> int:
        return self._reg(self._lib.zs_command_crop, src, rect.c())

    def color_convert(self, src: int, color: Color, texel: Texel) -> int:
        return self._reg(self._lib.zs_command_color_convert, src, ctypes.byref(color.c()), ctypes.byref(texel.c()))

    def inscribe(self, below: int, rect: Rectangle, above: int) -> int:
        return self._reg(self._lib.zs_command_inscribe, below, rect.c(), above)

    def transmute(self, src: int, target: Descriptor) -> int:
        return self._reg(self._lib.zs_command_transmute, src, ctypes.byref(target.c()))

    def affine(self, below: int, affine: Affine, above: int) -> int:
        m = (ctypes.c_float * 9)(*affine.transformation)
        return self._reg(self._lib.zs_command_affine, below, m, affine.sampling, above)

    def blend(self, below: int, rect: Rectangle, above: int, blend: int = Blend.Alpha) -> int:
        return self._reg(self._lib.zs_command_blend, below, rect.c(), above, blend)

    def solid_rgba(self, desc: Descriptor, color: Sequence[float]) -> int:
        c = (ctypes.c_float * 4)(*color)
        return self._reg(self._lib.zs_command_solid_rgba, ctypes.byref(desc.c()), c)

    def bilinear(self, desc: Descriptor, distribution: "Bilinear") -> int:
        """CommandBuffer::bilinear (command.rs:1615-1633): a generated image
        of bilinear.frag over the pixel centres' normalised coordinates."""
        p = (ctypes.c_float * 24)(*distribution.params())
        return self._reg(self._lib.zs_command_bilinear, ctypes.byref(desc.c()), p)

    def distribution_normal2d(self, desc: Descriptor, distribution: "DistributionNormal2d") -> int:
        """CommandBuffer::distribution_normal2d (command.rs:1559-1583): the
        density at the pixel centres as a grey (Luma or LumaA) image."""
        p = (ctypes.c_float * 8)(*distribution.params())
        return self._reg(self._lib.zs_command_distribution_normal2d, ctypes.byref(desc.c()), p)

    def invoke(self, function, args=None, num_results: int = 1) -> List[int]:
        """Call a function.

        ``invoke(fn, InvocationArguments(generics, arguments))``: the
        reference's ``CommandBuffer::invoke`` (command.rs:863-923) of a
        declared ``FunctionVar``, defined at ``Linker.link``; returns the
        result registers.  ``invoke(body, args)``: ``body`` (a command
        buffer without generics or calls) folded into this one at once."""
        if isinstance(function, FunctionVar):
            inv = args if isinstance(args, InvocationArguments) else InvocationArguments(arguments=args or ())
            g = (ctypes.c_uint32 * max(1, len(inv.generics)))(*inv.generics)
            a = (ctypes.c_uint32 * max(1, len(inv.arguments)))(*inv.arguments)
            out = (ctypes.c_uint32 * max(1, function.outputs))()
            _check(self._lib.zs_command_invoke_function(self._h, function.index, g, len(inv.generics), a,
                                                        len(inv.arguments), out, function.outputs))
            return list(out)[:function.outputs]
        body = function
        a = (ctypes.c_uint32 * max(1, len(args)))(*args)
        out = (ctypes.c_uint32 * max(1, num_results))()
        _check(self._lib.zs_command_invoke(self._h, body._h, a, len(args), out, num_results))
        return list(out)[:num_results]

    def generic(self, declaration: Optional["GenericDeclaration"] = None) -> int:
        """CommandBuffer::generic (command.rs:772-791): a generic descriptor
        parameter; returns its descriptor variable."""
        v = ctypes.c_uint32()
        _check(self._lib.zs_command_generic(self._h, ctypes.byref(v)))
        return v.value

    def input_generic(self, var: int) -> int:
        """CommandBuffer::input_generic (command.rs:756-762)."""
        return self._reg(self._lib.zs_command_input_generic, var)

    def register_descriptor(self, reg: int) -> int:
        """CommandBuffer::register_descriptor (command.rs:811-821)."""
        v = ctypes.c_uint32()
        _check(self._lib.zs_command_register_descriptor(self._h, reg, ctypes.byref(v)))
        return v.value

    def derive_descriptor(self, base: int, size: Optional[Tuple[int, int]] = None) -> int:
        """CommandBuffer::derive_descriptor (command.rs:794-808)."""
        v = ctypes.c_uint32()
        sz = (ctypes.c_uint32 * 2)(*size) if size is not None else None
        _check(self._lib.zs_command_derive_descriptor(self._h, base, sz, ctypes.byref(v)))
        return v.value

    def computed_signature(self) -> "CommandSignature":
        """CommandBuffer::computed_signature (command.rs:824-852)."""
        h = ctypes.c_void_p()
        _check(self._lib.zs_command_signature(self._h, ctypes.byref(h)))
        return CommandSignature(h)

    def function(self, signature: "CommandSignature") -> "FunctionVar":
        """CommandBuffer::function (command.rs:855-860)."""
        v = ctypes.c_uint32()
        _check(self._lib.zs_command_function(self._h, signature._h, ctypes.byref(v)))
        return FunctionVar(v.value, signature.vars, signature.inputs, signature.outputs)

    def resize(self, below: int, width: int, height: int) -> int:
        """CommandBuffer::resize (command.rs:1675-1705): a bilinear coordinate
        ramp of the target size looked up in ``below`` as a palette."""
        return self._reg(self._lib.zs_command_resize, below, width, height)

    def extract(self, src: int, channel: int) -> int:
        """CommandBuffer::extract (command.rs:1221-1265): one colour channel
        (``ColorChannel.R``, ``G`` or ``B``) as a single-channel image."""
        return self._reg(self._lib.zs_command_extract, src, channel)

    def inject(self, below: int, channel: int, above: int) -> int:
        """CommandBuffer::inject (command.rs:1360-1440): ``below`` with its
        ``channel`` replaced by the single-channel image ``above``."""
        return self._reg(self._lib.zs_command_inject, below, channel, above)

    def palette(self, palette: int, config: "Palette", indices: int) -> int:
        """CommandBuffer::palette (command.rs:1442-1485): per pixel of
        ``indices`` the ``palette`` texel its channels select (nearest)."""
        return self._reg(self._lib.zs_command_palette, palette, ctypes.byref(config.c()), indices)

    def distribution_fractal_noise(self, desc: Descriptor, noise: "FractalNoise") -> int:
        """CommandBuffer::distribution_fractal_noise (command.rs:1589-1607):
        fractal value noise in all four channels."""
        return self._reg(self._lib.zs_command_distribution_fractal_noise, ctypes.byref(desc.c()),
                         ctypes.byref(noise.c()))

    def oklab_adjust(self, src: int, chroma: float = 1.0, hue_rad: float = 0.0) -> int:
        return self._reg(self._lib.zs_command_oklab_adjust, src, chroma, hue_rad)

    def chromatic_adaptation(self, src: int, method: int, target: int) -> int:
        """CommandBuffer::chromatic_adaptation (command.rs:1112-1174): the RGB
        register adapted to white point ``target`` (``Whitepoint``) by a
        ``ChromaticAdaptationMethod``."""
        return self._reg(self._lib.zs_command_chromatic_adaptation, src, method, target)

    def derivative(self, src: int, method: int, direction: int) -> int:
        """CommandBuffer::derivative (command.rs:1487-1507): a 3x3 derivative
        filter (``DerivativeMethod``) along a ``Direction`` (box3.frag)."""
        return self._reg(self._lib.zs_command_derivative, src, method, direction)

    def buffer_init(self, data: bytes) -> int:
        """CommandBuffer::buffer_init (command.rs:1777-1790): a buffer register
        holding a copy of ``data``."""
        data = bytes(data)
        buf = ctypes.create_string_buffer(data, len(data)) if data else None
        return self._reg(self._lib.zs_command_buffer_init, buf, len(data))

    def buffer_zero(self, length: int) -> int:
        """CommandBuffer::buffer_zero (command.rs:1793-1801)."""
        return self._reg(self._lib.zs_command_buffer_zero, length)

    def buffer_size(self, reg: int) -> int:
        n = ctypes.c_uint64()
        _check(self._lib.zs_command_buffer_size(self._h, reg, ctypes.byref(n)))
        return n.value

    def from_buffer(self, buffer: int, desc: Descriptor) -> int:
        """CommandBuffer::from_buffer (command.rs:937-968): the image ``desc``
        read from a buffer whose rows are padded to 256 bytes."""
        return self._reg(self._lib.zs_command_from_buffer, buffer, ctypes.byref(desc.c()))

    def knob(self, reg: int) -> int:
        k = ctypes.c_int32()
        _check(self._lib.zs_command_knob(self._h, reg, ctypes.byref(k)))
        return k.value

    def describe(self, reg: int) -> Descriptor:
        d = DescriptorC()
        _check(self._lib.zs_command_describe(self._h, reg, ctypes.byref(d)))
        return Descriptor.from_c(d)

    def _dynamic(self, fn, regs, dynamic: "ShaderCommand") -> int:
        src = dynamic.source()
        sd = ShaderData()
        desc = dynamic.data(sd)
        buf = ctypes.create_string_buffer(sd.bytes, max(len(sd.bytes), 1))
        c = _abi.ShaderCommandC(src._h, ctypes.cast(buf, ctypes.c_void_p) if sd.bytes else None, len(sd.bytes),
                                desc.c())
        return self._reg(fn, *regs, ctypes.byref(c))

    def construct_dynamic(self, dynamic: "ShaderCommand") -> int:
        """``CommandBuffer::construct_dynamic`` (command.rs:2933-2961)."""
        return self._dynamic(self._lib.zs_command_construct_dynamic, (), dynamic)

    def unary_dynamic(self, op: int, dynamic: "ShaderCommand") -> int:
        """``CommandBuffer::unary_dynamic`` (command.rs:2963-3004)."""
        return self._dynamic(self._lib.zs_command_unary_dynamic, (op,), dynamic)

    def binary_dynamic(self, lhs: int, rhs: int, dynamic: "ShaderCommand") -> int:
        """``CommandBuffer::binary_dynamic`` (command.rs:3006-3048)."""
        return self._dynamic(self._lib.zs_command_binary_dynamic, (lhs, rhs), dynamic)


SHADER_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shaders")


class ShaderSource:
    """``ShaderSource`` (command/dynamic.rs:31-34) in its HIP form: a gfx950
    code object and the name of its stage kernel (``zh_shader_create``)."""

    def __init__(self, code: bytes, kernel: str):
        self._lib = _abi.lib()
        h = ctypes.c_void_p()
        _check(self._lib.zh_shader_create(code, len(code), kernel.encode(), ctypes.byref(h)))
        self._h = h
        self.kernel = kernel

    @staticmethod
    def hip_code_object(path: str, kernel: str) -> "ShaderSource":
        with open(path, "rb") as f:
            return ShaderSource(f.read(), kernel)

    @staticmethod
    def bundled(name: str) -> "ShaderSource":
        """One of the package's own stage kernels (``shaders/<name>.co``,
        kernel ``name``)."""
        return ShaderSource.hip_code_object(os.path.join(SHADER_DIR, name + ".co"), name)

    def __del__(self):
        if getattr(self, "_h", None) and _abi is not None:
            self._lib.zh_shader_release(self._h)
            self._h = None


class ShaderData:
    """``ShaderData`` (command/dynamic.rs:40-51): the shader's argument bytes."""

    def __init__(self):
        self.bytes = b""

    def set_data(self, values, dtype=np.float32):
        """``ShaderData::set_data``: the values as plain bytes of ``dtype``."""
        self.bytes = np.ascontiguousarray(values, dtype=dtype).tobytes()


class ShaderCommand:
    """The ``ShaderCommand`` trait (command/dynamic.rs:7-28): ``source`` gives
    the kernel, ``data`` sets its argument bytes and returns the result's
    descriptor."""

    def source(self) -> ShaderSource:
        raise NotImplementedError

    def data(self, data: ShaderData) -> Descriptor:
        raise NotImplementedError


@dataclass(frozen=True)
class GenericDeclaration:
    """command.rs:762-770: bounds of a generic parameter (none exist yet)."""
    bounds: Tuple = ()


@dataclass(frozen=True)
class InvocationArguments:
    """command.rs:703-710: generic arguments (descriptor variables) and
    argument registers of a call."""
    generics: Tuple[int, ...] = ()
    arguments: Tuple[int, ...] = ()


@dataclass(frozen=True)
class FunctionVar:
    index: int
    vars: int
    inputs: int
    outputs: int


class CommandSignature:
    """CommandSignature (command.rs:680-700) of a command buffer."""

    def __init__(self, handle):
        self._lib = _abi.lib()
        self._h = handle
        v, i, o = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(self._lib.zs_signature_counts(handle, ctypes.byref(v), ctypes.byref(i), ctypes.byref(o)))
        self.vars, self.inputs, self.outputs = v.value, i.value, o.value

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.zs_signature_destroy(self._h)
            self._h = None


@dataclass
class Bilinear:
    """shaders/bilinear.rs ShaderData: at the pixel centre (u, v) in [0, 1]^2
    the value is mix(u_min, u_max, u) + mix(v_min, v_max, v)
    + mix(uv_min, uv_max, u v), per channel."""
    u_min: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    u_max: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    v_min: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    v_max: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    uv_min: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    uv_max: Sequence[float] = (0.0, 0.0, 0.0, 0.0)

    @staticmethod
    def mgrid(width: float, height: float) -> "Bilinear":
        """shaders/bilinear.rs:22-31: the pixel centre's coordinates in r, g."""
        return Bilinear(u_max=(width, 0.0, 0.0, 0.0), v_max=(0.0, height, 0.0, 0.0))

    def params(self) -> List[float]:
        out: List[float] = []
        for v in (self.u_min, self.u_max, self.v_min, self.v_max, self.uv_min, self.uv_max):
            if len(v) != 4:
                raise ValueError("each Bilinear corner is a vec4")
            out += [float(x) for x in v]
        return out


class DistributionNormal2d:
    """shaders/distribution_normal2d.rs ShaderData: expectation, inverse
    covariance (row-major) and pseudo-determinant of a normal distribution in
    screen coordinates [-1, 1]^2."""

    def __init__(self, params: Sequence[float]):
        if len(params) != 8:
            raise ValueError("a normal distribution has 8 parameter floats")
        self._p = [float(x) for x in params]

    @staticmethod
    def _make(fn, a: float, b: float) -> "DistributionNormal2d":
        out = (ctypes.c_float * 8)()
        _check(fn(a, b, out))
        return DistributionNormal2d(list(out))

    @staticmethod
    def with_diagonal(var0: float, var1: float) -> "DistributionNormal2d":
        """Centred, with variances ``var0``, ``var1`` along the axes."""
        return DistributionNormal2d._make(_abi.lib().zs_normal2d_with_diagonal, var0, var1)

    @staticmethod
    def with_direction(direction: Sequence[float]) -> "DistributionNormal2d":
        """Centred, 1-D along ``direction``, whose squared length is the variance."""
        return DistributionNormal2d._make(_abi.lib().zs_normal2d_with_direction, direction[0], direction[1])

    def params(self) -> List[float]:
        return list(self._p)


class ColorChannel:
    """Channels that can index a palette (shaders/palette.rs:30-38)."""
    NONE, R, G, B, A = -1, 0, 1, 2, 3


@dataclass
class Palette:
    """command.rs:527-537: which channels give the palette coordinates."""
    width: int = ColorChannel.NONE
    height: int = ColorChannel.NONE
    width_base: int = 0
    height_base: int = 0

    def c(self) -> "_abi.PaletteC":
        return _abi.PaletteC(self.width, self.height, self.width_base, self.height_base)


class FractalNoise:
    """shaders/fractal_noise.rs ShaderData: octaves, initial amplitude,
    amplitude damping and the grid scale (cells per unit of uv)."""

    def __init__(self, num_octaves: int, initial_amplitude: float, amplitude_damping: float,
                 grid_scale: Sequence[float] = (100.0, 100.0)):
        self.num_octaves = int(num_octaves)
        self.initial_amplitude = float(initial_amplitude)
        self.amplitude_damping = float(amplitude_damping)
        self.grid_scale = (float(grid_scale[0]), float(grid_scale[1]))

    def c(self) -> "_abi.FractalNoiseC":
        n = _abi.FractalNoiseC()
        n.num_octaves = self.num_octaves
        n.initial_amplitude = self.initial_amplitude
        n.amplitude_damping = self.amplitude_damping
        n.grid_scale[0], n.grid_scale[1] = self.grid_scale
        return n

    @staticmethod
    def _from_c(n) -> "FractalNoise":
        return FractalNoise(n.num_octaves, n.initial_amplitude, n.amplitude_damping, tuple(n.grid_scale))

    @staticmethod
    def with_octaves(num_octaves: int) -> "FractalNoise":
        n = _abi.FractalNoiseC()
        _check(_abi.lib().zs_fractal_noise_with_octaves(num_octaves, ctypes.byref(n)))
        return FractalNoise._from_c(n)

    def set_damping(self, damping: float) -> "FractalNoise":
        """The damping, and the initial amplitude that makes the octaves'
        amplitudes sum to 1."""
        n = self.c()
        _check(_abi.lib().zs_fractal_noise_set_damping(ctypes.byref(n), damping))
        self.initial_amplitude, self.amplitude_damping = n.initial_amplitude, n.amplitude_damping
        return self


class Linker:
    """``Linker::from_included().compile(&commands)`` (command.rs:2063-2185)."""

    @staticmethod
    def from_included() -> "Linker":
        return Linker()

    def compile(self, commands: CommandBuffer) -> "Program":
        h = ctypes.c_void_p()
        _check(_abi.lib().zs_command_compile(commands._h, ctypes.byref(h)))
        return Program(h)

    def link(self, main: CommandBuffer, tys: Sequence[Descriptor], functions: Sequence[CommandBuffer],
             links: Sequence[Sequence[int]]) -> "Program":
        """Linker::link (command.rs:2083-2185): ``main`` instantiated with the
        descriptors ``tys``; ``links[b][k]`` names the buffer (0 = main, 1 +
        i = functions[i]) defining function symbol k of buffer b."""
        n = 1 + len(functions)
        t = (DescriptorC * max(1, len(tys)))(*[d.c() for d in tys])
        f = (ctypes.c_void_p * max(1, len(functions)))(*[fb._h.value for fb in functions])
        counts = [len(links[b]) if b < len(links) else 0 for b in range(n)]
        targets = [int(x) for b in range(n) for x in (links[b] if b < len(links) else [])]
        c = (ctypes.c_uint32 * n)(*counts)
        g = (ctypes.c_uint32 * max(1, len(targets)))(*targets)
        h = ctypes.c_void_p()
        _check(_abi.lib().zs_link(main._h, t, len(tys), f, len(functions), c, g, ctypes.byref(h)))
        prog = Program(h)
        prog._keep = (main, list(functions))
        return prog


# High op kinds of the program's inputs and outputs (zosimos_hip.h)
HIGH_INPUT = 0
HIGH_OUTPUT = 1

# Plan flags (zosimos_hip.h)
PLAN_DEFAULT = 0
PLAN_STRICT_REGISTERS = 1
PLAN_NO_GRAPH = 2
PLAN_NO_JIT = 4
PLAN_STAGED = 8
PLAN_PAIR = 16


class Program:
    """A linked program: the High IR list (program.rs:24-39)."""

    def __init__(self, handle):
        self._lib = _abi.lib()
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.zs_linked_destroy(self._h)
            self._h = None

    def high(self):
        tex = ctypes.POINTER(DescriptorC)()
        ops = ctypes.POINTER(_abi.OpC)()
        nt, no = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self._lib.zs_linked_high(self._h, ctypes.byref(tex), ctypes.byref(nt), ctypes.byref(ops),
                                        ctypes.byref(no)))
        return [tex[i] for i in range(nt.value)], [ops[i] for i in range(no.value)]

    def register_descriptor(self, reg: int, output: bool) -> Descriptor:
        """The descriptor of an input (or output) register of this program,
        from its High::Input (High::Output) entry."""
        slot = self.output_slot(reg) if output else self.input_slot(reg)
        tex, ops = self.high()
        for op in ops:
            if not output and op.high == HIGH_INPUT and op.src == slot:
                return Descriptor.from_c(tex[op.dst])
            if output and op.high == HIGH_OUTPUT and op.dst == slot:
                return Descriptor.from_c(tex[op.src])
        raise ZosimosError(Status.INVALID_ARGUMENT, f"register {reg} is not an {'output' if output else 'input'}")

    def input_slot(self, reg: int) -> int:
        s = ctypes.c_uint32()
        _check(self._lib.zs_linked_input_slot(self._h, reg, ctypes.byref(s)))
        return s.value

    def output_slot(self, reg: int) -> int:
        s = ctypes.c_uint32()
        _check(self._lib.zs_linked_output_slot(self._h, reg, ctypes.byref(s)))
        return s.value

    def query_knob(self, link_idx: int, register: int) -> int:
        """Executable::query_knob(RegisterKnob {link_idx, register})
        (run.rs:1012-1014): the knob id of a knob register of buffer
        ``link_idx`` (0 = the main buffer) in this linked program."""
        k = ctypes.c_int32()
        _check(self._lib.zs_linked_knob(self._h, link_idx, register, ctypes.byref(k)))
        return k.value

    def launch_static(self, ctx: "Context", inputs: Dict[int, "Image"], outputs: Dict[int, "Image"],
                      flags: int = PLAN_DEFAULT):
        """Lower and launch once without allocating device memory: the plan's
        device tables go to the library's static scratch
        (``zh_program_launch_static``; asynchronous, no hipGraph)."""
        tex, ops = self.high()
        t = (DescriptorC * max(1, len(tex)))(*tex)
        o = (_abi.OpC * max(1, len(ops)))(*ops)
        binds = [BindingC(self.input_slot(r), 0, img._h) for r, img in inputs.items()]
        binds += [BindingC(self.output_slot(r), 1, img._h) for r, img in outputs.items()]
        b = (BindingC * max(1, len(binds)))(*binds)
        _check(self._lib.zh_program_launch_static(ctx._h, t, len(tex), o, len(ops), flags, b, len(binds)))

    def lower_to(self, ctx: Optional["Context"] = None, flags: int = PLAN_DEFAULT) -> "Executable":
        """``Program::lower_to`` (program.rs:1304).  ``ctx=None`` plans only."""
        h = ctypes.c_void_p()
        _check(self._lib.zs_linked_lower(self._h, ctx._h if ctx else None, flags, ctypes.byref(h)))
        return Executable(h, self, ctx)


class Executable:
    """``run::Executable`` (run.rs:38-62) of one lowered program."""

    def __init__(self, handle, program: Program, ctx: Optional["Context"]):
        self._lib = _abi.lib()
        self._h = handle
        self.program = program
        self.ctx = ctx

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.zh_program_destroy(self._h)
            self._h = None

    def describe(self) -> str:
        need = ctypes.c_size_t()
        _check(self._lib.zh_program_describe(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(self._lib.zh_program_describe(self._h, buf, need.value, None))
        return buf.value.decode()

    def stats(self) -> Tuple[int, int]:
        n, b = ctypes.c_uint32(), ctypes.c_uint64()
        _check(self._lib.zh_program_stats(self._h, ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    GRAPH_STATES = {0: "not yet", 1: "captured", 2: "failed", 3: "disabled", 4: "single launch"}

    def graph_info(self) -> dict:
        """hipGraph state (``zh_program_graph_info``): state name, nodes,
        captures so far and launches that replayed the graph."""
        g = _abi.GraphInfoC()
        _check(self._lib.zh_program_graph_info(self._h, ctypes.byref(g)))
        return {"state": self.GRAPH_STATES.get(g.state, str(g.state)), "nodes": g.nodes, "captures": g.captures,
                "replays": g.replays}

    def knob(self, knob: int, data: bytes):
        """``Environment::knob`` (run.rs:1292-1306)."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _check(self._lib.zh_program_set_knob(self._h, knob, buf, len(data)))

    def _bindings(self, inputs: Dict[int, "Image"], outputs: Dict[int, "Image"]):
        binds = []
        for reg, img in inputs.items():
            binds.append(BindingC(self.program.input_slot(reg), 0, img._h))
        for reg, img in outputs.items():
            binds.append(BindingC(self.program.output_slot(reg), 1, img._h))
        arr = (BindingC * len(binds))(*binds)
        return arr, len(binds)

    def knob_data(self, knob: int) -> bytes:
        """The knob's current parameter bytes."""
        n = ctypes.c_uint32()
        _check(self._lib.zh_program_knob_data(self._h, knob, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(max(n.value, 1))
        _check(self._lib.zh_program_knob_data(self._h, knob, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def from_pool(self, pool: "Pool") -> "Environment":
        """``Executable::from_pool`` (run.rs:471-488): the bindings and knob
        data of one launch with images from ``pool``."""
        return Environment(self, pool)

    def launch(self, inputs, outputs: Optional[Dict[int, "Image"]] = None):
        """``Executable::launch`` (run.rs:1016-1167).  With an
        ``Environment``: bind and trade its pool images and return the
        ``Execution``, which queues nothing until it is stepped.  With
        dictionaries of registers to images: bind them and launch
        asynchronously, stepped to the end."""
        if isinstance(inputs, Environment):
            return inputs._launch()
        arr, n = self._bindings(inputs, outputs or {})
        _check(self._lib.zh_program_launch(self._h, arr, n))

    def _pool_launch(self, fn, pool: "Pool", inputs: Dict[int, int], outputs: Dict[int, int]) -> "Execution":
        binds = [PoolBindingC(self.program.input_slot(r), 0, k) for r, k in inputs.items()]
        binds += [PoolBindingC(self.program.output_slot(r), 1, k) for r, k in outputs.items()]
        arr = (PoolBindingC * max(len(binds), 1))(*binds)
        h = ctypes.c_void_p()
        _check(fn(self._h, pool._h, arr, len(binds), ctypes.byref(h)))
        return Execution(h, self, pool)

    def launch_pool(self, pool: "Pool", inputs: Dict[int, int], outputs: Dict[int, int]) -> "Execution":
        """Launch with inputs read from pool keys and outputs written to pool
        keys (0: a new key at retirement); ``Executable::launch`` with a pool
        (run.rs:1016-1167), stepped to the end.  Returns the execution."""
        return self._pool_launch(self._lib.zh_program_launch_pool, pool, inputs, outputs)

    def start_pool(self, pool: "Pool", inputs: Dict[int, int], outputs: Dict[int, int]) -> "Execution":
        """``Executable::launch`` before any step: bound and traded, nothing
        queued until ``Execution.step``."""
        return self._pool_launch(self._lib.zh_program_start_pool, pool, inputs, outputs)

    def launch_once(self, inputs: Dict[int, "Image"], outputs: Dict[int, "Image"]):
        """``Executable::launch_once`` (run.rs:1050-1079): one launch that
        consumes this executable."""
        arr, n = self._bindings(inputs, outputs)
        h, self._h = self._h, None
        _check(self._lib.zh_program_launch_once(h, arr, n))

    def dot(self) -> str:
        """``Executable::dot`` (run.rs:490-1010): a Graphviz digraph of the plan."""
        need = ctypes.c_size_t()
        _check(self._lib.zh_program_dot(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(self._lib.zh_program_dot(self._h, buf, need.value, None))
        return buf.value.decode()

    def time_stages(self, inputs, outputs, iterations: int = 10):
        arr, n = self._bindings(inputs, outputs)
        out = (ctypes.c_float * 64)()
        cnt = ctypes.c_uint32()
        _check(self._lib.zh_program_time(self._h, arr, n, iterations, out, 64, ctypes.byref(cnt)))
        return [out[i] for i in range(cnt.value)]

    def launch_bytes(self) -> List[int]:
        """Algorithmic HBM bytes of each kernel launch, in launch order."""
        out = (ctypes.c_uint64 * 256)()
        n = ctypes.c_uint32()
        _check(self._lib.zh_program_launch_bytes(self._h, out, 256, ctypes.byref(n)))
        return [out[i] for i in range(min(n.value, 256))]

    def autotune(self, inputs, outputs, iterations: int = 5):
        """Time every tile shape of the program's warp kernels on these images
        and keep the fastest (``zh_program_autotune``)."""
        arr, n = self._bindings(inputs, outputs)
        _check(self._lib.zh_program_autotune(self._h, arr, n, iterations))

    def learn(self, **options) -> List[dict]:
        """Fit the learned colour surrogate of every eligible stage
        (``zh_program_learn``); returns one report per stage."""
        o = _abi.LearnOptionsC()
        for k, v in options.items():
            setattr(o, k, v)
        reps = (_abi.LearnReportC * 16)()
        cnt = ctypes.c_uint32()
        _check(self._lib.zh_program_learn(self._h, ctypes.byref(o), reps, 16, ctypes.byref(cnt)))
        return [{f: getattr(reps[i], f) for f, _ in _abi.LearnReportC._fields_} for i in range(min(cnt.value, 16))]

    def use_surrogates(self, mode: int = 1):
        """0: plain kernels; 1: accepted surrogates; 2: every fitted surrogate."""
        _check(self._lib.zh_program_use_surrogates(self._h, int(mode)))

    def surrogate_params(self, stage: int):
        import numpy as np
        out = np.zeros(MLP_PARAMS, dtype=np.float32)
        _check(self._lib.zh_program_surrogate_params(self._h, stage, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                     MLP_PARAMS))
        return out


class Context:
    """One HIP device and stream (the analogue of a ``Pool`` device)."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        self._lib = _abi.lib()
        h = ctypes.c_void_p()
        _check(self._lib.zh_context_create(device, stream, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.zh_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def stream(self) -> int:
        return self._lib.zh_context_stream(self._h)

    def sync(self):
        _check(self._lib.zh_sync(self._h))

    def is_running(self) -> bool:
        r = ctypes.c_int32()
        _check(self._lib.zh_query(self._h, ctypes.byref(r)))
        return bool(r.value)

    def image(self, desc: Descriptor) -> "Image":
        h = ctypes.c_void_p()
        _check(self._lib.zh_image_alloc(self._h, ctypes.byref(desc.c()), ctypes.byref(h)))
        return Image(self, desc, h)

    def wrap(self, desc: Descriptor, device_ptr: int, nbytes: int) -> "Image":
        h = ctypes.c_void_p()
        _check(self._lib.zh_image_wrap(self._h, ctypes.byref(desc.c()), device_ptr, nbytes, ctypes.byref(h)))
        return Image(self, desc, h, keep=None)


class Image:
    def __init__(self, ctx: Context, desc: Descriptor, handle, keep=None):
        self.ctx = ctx
        self.desc = desc
        self._h = handle
        self._keep = keep

    def free(self):
        if getattr(self, "_h", None):
            if _abi is not None:  # module globals are cleared at interpreter exit
                _abi.lib().zh_image_free(self._h)
            self._h = None

    def __del__(self):
        self.free()

    @property
    def device_ptr(self) -> int:
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        _check(_abi.lib().zh_image_device_ptr(self._h, ctypes.byref(p), ctypes.byref(n)))
        return p.value

    def upload(self, data: np.ndarray, sync: bool = True):
        """Copy tightly packed host rows into the image."""
        arr = np.ascontiguousarray(data)
        row = self.desc.width * self.desc.texel.bytes
        if arr.nbytes != row * self.desc.height:
            raise ValueError(f"expected {row * self.desc.height} bytes, got {arr.nbytes}")
        _check(_abi.lib().zh_image_upload(self._h, arr.ctypes.data, row))
        if sync:
            self.ctx.sync()

    def download(self) -> np.ndarray:
        """Tightly packed bytes of the image (uint8, shape (H, W * bpp))."""
        row = self.desc.width * self.desc.texel.bytes
        out = np.empty((self.desc.height, row), dtype=np.uint8)
        _check(_abi.lib().zh_image_download(self._h, out.ctypes.data, row))
        self.ctx.sync()
        return out


class Pool:
    """``zosimos::pool::Pool`` (pool.rs:19-156): images under keys, as host
    bytes, a device image, or only a descriptor (late bound)."""

    KINDS = {0: "host", 1: "device", 2: "late bound"}

    def __init__(self):
        self._lib = _abi.lib()
        h = ctypes.c_void_p()
        _check(self._lib.zh_pool_create(ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _abi is not None:
            self._lib.zh_pool_destroy(self._h)
            self._h = None

    def insert(self, desc: Descriptor, data) -> int:
        """``Pool::insert`` (pool.rs:244-252): tightly packed host bytes."""
        arr = np.ascontiguousarray(data)
        if arr.nbytes != desc.width * desc.texel.bytes * desc.height:
            raise ValueError(f"expected {desc.width * desc.texel.bytes * desc.height} bytes, got {arr.nbytes}")
        k = ctypes.c_uint64()
        _check(self._lib.zh_pool_insert_host(self._h, ctypes.byref(desc.c()), arr.ctypes.data, 0, ctypes.byref(k)))
        return k.value

    def declare(self, desc: Descriptor) -> int:
        """``Pool::declare`` (pool.rs:281-291)."""
        k = ctypes.c_uint64()
        _check(self._lib.zh_pool_declare(self._h, ctypes.byref(desc.c()), ctypes.byref(k)))
        return k.value

    def insert_device(self, image: "Image") -> int:
        """Hand a device image to the pool (``ImageData::GpuBuffer``); the
        Image object gives up its handle."""
        k = ctypes.c_uint64()
        _check(self._lib.zh_pool_insert_device(self._h, image._h, ctypes.byref(k)))
        image._h = None
        return k.value

    def info(self, key: int) -> Tuple[Descriptor, str]:
        d, kind = DescriptorC(), ctypes.c_int32()
        _check(self._lib.zh_pool_image_info(self._h, key, ctypes.byref(d), ctypes.byref(kind)))
        return Descriptor.from_c(d), self.KINDS.get(kind.value, str(kind.value))

    def descriptor(self, key: int) -> Descriptor:
        return self.info(key)[0]

    def kind(self, key: int) -> str:
        return self.info(key)[1]

    def as_bytes(self, key: int) -> np.ndarray:
        """A copy of the host bytes (``PoolImage::as_bytes``, pool.rs:669)."""
        p, n = ctypes.c_void_p(), ctypes.c_uint64()
        _check(self._lib.zh_pool_host_bytes(self._h, key, ctypes.byref(p), ctypes.byref(n)))
        return np.frombuffer(ctypes.string_at(p.value, n.value), dtype=np.uint8).copy()

    def device_ptr(self, key: int) -> int:
        img = ctypes.c_void_p()
        _check(self._lib.zh_pool_device_image(self._h, key, ctypes.byref(img)))
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        _check(self._lib.zh_image_device_ptr(img, ctypes.byref(p), ctypes.byref(n)))
        return p.value

    def set_no_read(self, key: int, no_read: bool = True):
        _check(self._lib.zh_pool_set_no_read(self._h, key, int(no_read)))

    def upload(self, key: int, ctx: "Context"):
        _check(self._lib.zh_pool_upload(self._h, key, ctx._h))

    def download(self, key: int):
        _check(self._lib.zh_pool_download(self._h, key))

    def remove(self, key: int):
        _check(self._lib.zh_pool_remove(self._h, key))

    def clear_cache(self):
        _check(self._lib.zh_pool_clear_cache(self._h))

    def cached_buffers(self) -> int:
        n = ctypes.c_uint32()
        _check(self._lib.zh_pool_cached_buffers(self._h, ctypes.byref(n)))
        return n.value


class Execution:
    """A launch from a pool and its ``Retire`` (run.rs:2786-2998)."""

    def __init__(self, handle, exe: Executable, pool: Pool, outputs: Optional[Dict[int, int]] = None,
                 restore: Optional[Dict[int, bytes]] = None):
        self._lib = _abi.lib()
        self._h = handle
        self.exe = exe
        self.pool = pool  # the pool outlives the execution
        self._queued = 0
        self._outputs = dict(outputs or {})  # register -> the pool key it was bound to (0: none)
        self._restore = dict(restore or {})  # knob -> the bytes it held before this run

    def retire_output(self, reg: int, to_host: bool = True) -> int:
        """``Retire::output`` (run.rs:2809-2875): the output's pool key."""
        k = ctypes.c_uint64()
        _check(self._lib.zh_execution_retire_output(self._h, self.exe.program.output_slot(reg), int(to_host),
                                                    ctypes.byref(k)))
        return k.value

    def retire_input(self, reg: int) -> int:
        k = ctypes.c_uint64()
        _check(self._lib.zh_execution_retire_input(self._h, self.exe.program.input_slot(reg), ctypes.byref(k)))
        return k.value

    def step_to(self, instructions: int = 1) -> "SyncPoint":
        """``Execution::step_to(StepLimits { instructions })`` (run.rs:1394-
        1464): queue up to ``instructions`` more kernel launches; the
        ``SyncPoint`` waits for them and says how many were queued."""
        import time
        t0 = time.perf_counter()
        first = self._queued
        n = ctypes.c_uint32()
        _check(self._lib.zh_execution_step(self._h, instructions, ctypes.byref(n)))
        self._queued += n.value
        return SyncPoint(self, n.value, t0, first)

    def step(self) -> "SyncPoint":
        """``Execution::step`` (run.rs:1389-1391): one launch."""
        return self.step_to(1)

    def retire_gracefully(self, pool: "Pool") -> "Retire":
        """``Execution::retire_gracefully`` (run.rs:1479-1492): stop, with the
        images and buffers going back into ``pool``."""
        if pool is not self.pool:
            raise ZosimosError(Status.INVALID_ARGUMENT, "an execution retires into the pool it was launched from")
        return Retire(self)

    def is_running(self) -> bool:
        r = ctypes.c_int32()
        _check(self._lib.zh_execution_is_running(self._h, ctypes.byref(r)))
        return bool(r.value)

    def resources_used(self) -> dict:
        """``Execution::resources_used`` (run.rs:212-221)."""
        u = _abi.ResourcesUsedC()
        _check(self._lib.zh_execution_resources(self._h, ctypes.byref(u)))
        return {f: getattr(u, f) for f, _ in _abi.ResourcesUsedC._fields_}

    def prune(self):
        """``Retire::prune``: free this execution's buffers at finish."""
        _check(self._lib.zh_execution_prune(self._h))

    def _restore_knobs(self):
        # A run's knob data (Environment::knob) held for that run only.
        restore, self._restore = self._restore, {}
        for knob, data in restore.items():
            self.exe.knob(knob, data)

    def finish(self):
        """``Retire::finish`` (run.rs:2962-2998)."""
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _check(self._lib.zh_execution_finish(h))
            self._restore_knobs()

    def finish_by_discarding(self):
        """``Retire::finish_by_discarding``: drop everything not retired."""
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _check(self._lib.zh_execution_discard(h))
            self._restore_knobs()

    def __del__(self):
        if getattr(self, "_h", None) and _abi is not None:
            self._lib.zh_execution_finish(self._h)
            self._h = None


class Environment:
    """``Environment`` (run.rs:1170-1348): the pool images and knob data of
    one launch, from ``Executable.from_pool``; ``Executable.launch(env)``
    starts it."""

    def __init__(self, exe: Executable, pool: Pool):
        self.exe = exe
        self.pool = pool
        self._inputs: Dict[int, int] = {}
        self._outputs: Dict[int, int] = {}
        self._knobs: Dict[int, bytes] = {}

    def _check_image(self, reg: int, key: int, output: bool):
        want = self.exe.program.register_descriptor(reg, output)
        have = self.pool.descriptor(key)
        # As the reference: layout and texel must match, colour is not compared.
        if (have.width, have.height, have.texel) != (want.width, want.height, want.texel) or (
                have.row_stride and want.row_stride and have.row_stride != want.row_stride):
            raise ZosimosError(Status.MISMATCHED_DESCRIPTOR,
                               f"pool image {key} does not match register {reg}'s descriptor")

    def bind(self, reg: int, key: int):
        """``Environment::bind`` (run.rs:1171-1205): the input ``reg`` reads
        pool image ``key``."""
        self._check_image(reg, key, False)
        self._inputs[reg] = key

    def bind_output(self, reg: int, key: int):
        """``Environment::bind_output`` (run.rs:1207-1242): the output ``reg``
        is written into pool image ``key``."""
        self._check_image(reg, key, True)
        self._outputs[reg] = key

    def bind_render(self, reg: int, key: int):
        """``Environment::bind_render`` (run.rs:1244-1279): as
        ``bind_output``, into a pool image that lives on the device (the
        reference's render targets are GPU textures)."""
        if self.pool.kind(key) != "device":
            raise ZosimosError(Status.INVALID_ARGUMENT, f"render target {key} is not a device image")
        self.bind_output(reg, key)

    def knob(self, knob: int, data: bytes):
        """``Environment::knob`` (run.rs:1292-1306): the knob's data for this
        run only; the length must equal the knob's."""
        have = len(self.exe.knob_data(knob))
        if len(data) != have:
            raise ZosimosError(Status.KNOB_LENGTH, f"knob data length {len(data)} differs from {have}")
        self._knobs[knob] = bytes(data)

    def knob_by_register(self, knob: Tuple[int, int], data: bytes):
        """``Environment::knob_by_register`` (run.rs:1283-1290) with a
        ``RegisterKnob`` (link index, register)."""
        self.knob(self.exe.program.query_knob(knob[0], knob[1]), data)

    def recover_buffers(self) -> dict:
        """``Environment::recover_buffers`` (run.rs:1312-1347): the launch
        takes device buffers of matching size from the pool's cache; this
        reports how many the cache holds for it to take."""
        return {"cached_buffers": self.pool.cached_buffers()}

    def _launch(self) -> "Execution":
        restore = {}
        try:
            for knob, data in self._knobs.items():
                restore[knob] = self.exe.knob_data(knob)
                self.exe.knob(knob, data)
            # Outputs left unbound go to new pool images when retired.
            prog = self.exe.program
            bound = {prog.output_slot(reg): key for reg, key in self._outputs.items()}
            _, ops = prog.high()
            slots = sorted({op.dst for op in ops if op.high == HIGH_OUTPUT})
            binds = [PoolBindingC(prog.input_slot(r), 0, k) for r, k in self._inputs.items()]
            binds += [PoolBindingC(s, 1, bound.get(s, 0)) for s in slots]
            arr = (PoolBindingC * max(len(binds), 1))(*binds)
            h = ctypes.c_void_p()
            _check(self.exe._lib.zh_program_start_pool(self.exe._h, self.pool._h, arr, len(binds), ctypes.byref(h)))
            run = Execution(h, self.exe, self.pool)
        except BaseException:
            for knob, data in restore.items():
                self.exe.knob(knob, data)
            raise
        run._outputs = dict(self._outputs)
        run._restore = restore
        return run


class SyncPoint:
    """``SyncPoint`` (run.rs:3015-3184): the launches one step queued."""

    def __init__(self, run: "Execution", launches: int, t0: float, first: int):
        self.run = run
        self.launches = launches
        self._t0 = t0
        self._t1: Optional[float] = None
        self._first = first

    def block_on(self):
        """``SyncPoint::block_on``: wait until the queued launches finished."""
        import time
        self.run.exe.ctx.sync()
        self._t1 = time.perf_counter()

    def finish(self):
        """``SyncPoint::finish``."""
        self.block_on()

    def time_spent(self) -> float:
        """``SyncPoint::time_spent``: seconds from the step to the end of
        ``block_on`` (or to now, before it)."""
        import time
        return (self._t1 if self._t1 is not None else time.perf_counter()) - self._t0

    def debug_mark(self) -> Optional[str]:
        """``SyncPoint::debug_mark``: the first launch this step queued."""
        if self.launches == 0:
            return None
        head = f"launch {self._first}: "
        lines = [l for l in self.run.exe.describe().splitlines() if l.startswith(head)]
        return lines[0] if lines else None


class Retire:
    """``Retire`` (run.rs:2786-2998), from ``Execution.retire_gracefully``:
    images and buffers back into the pool."""

    def __init__(self, run: "Execution"):
        self.run = run
        self.pool = run.pool

    def input(self, reg: int) -> int:
        """``Retire::input``: a ``no_read`` input moved out for the launch
        goes back under its key."""
        return self.run.retire_input(reg)

    def output(self, reg: int) -> int:
        """``Retire::output``: the output's pool key (the bound key, else a
        new one), with its bytes on the host."""
        return self.run.retire_output(reg, to_host=True)

    def render(self, reg: int) -> int:
        """``Retire::render``: the render target's pool key, kept on the
        device."""
        return self.run.retire_output(reg, to_host=False)

    def output_key(self, reg: int) -> Optional[int]:
        """``Retire::output_key`` (run.rs:2860-2874): the key ``output``
        will use, when the output was bound to one."""
        self.run.exe.program.output_slot(reg)  # raises for a register that is no output
        key = self.run._outputs.get(reg, 0)
        return key or None

    def retire_buffers(self) -> dict:
        """``Retire::retire_buffers`` (run.rs:2876-2945): the execution's
        device buffers go to the pool's cache at ``finish``; returns what it
        used."""
        return self.run.resources_used()

    def prune(self):
        """``Retire::prune``: give no buffer back to the pool's cache."""
        self.run.prune()

    def finish(self):
        """``Retire::finish``."""
        self.run.finish()

    def finish_by_discarding(self):
        """``Retire::finish_by_discarding``."""
        self.run.finish_by_discarding()


MLP_PARAMS = 4611


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mlp_gradient(ctx: Context, params, x, t):
    """(loss, gradient) of the surrogate's training kernels on samples
    x[n][3] -> t[n][3] (``zh_mlp_gradient``)."""
    import numpy as np
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(MLP_PARAMS, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _check(_abi.lib().zh_mlp_gradient(ctx._h, _f32p(p), x.shape[0], _f32p(x), _f32p(t), _f32p(g), _f32p(loss)))
    return float(loss[0]), g


def mlp_evaluate(ctx: Context, params, x):
    """The surrogate's matrix-core inference on x[n][3] (``zh_mlp_evaluate``)."""
    import numpy as np
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x)
    _check(_abi.lib().zh_mlp_evaluate(ctx._h, _f32p(p), x.shape[0], _f32p(x), _f32p(y)))
    return y


def version() -> str:
    return _abi.lib().zh_version().decode()

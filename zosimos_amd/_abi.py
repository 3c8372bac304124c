"""ctypes declarations of ``include/zosimos_hip.h`` and ``include/zosimos_command.h``.

Loading fails loudly when the built library is missing: there is no CPU
fallback behind these entry points.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzosimos_hip.so")

u32, i32, u64, f32 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_float
vp, sz = ctypes.c_void_p, ctypes.c_size_t


class Texel(ctypes.Structure):
    _fields_ = [("bits", u32), ("parts", u32)]


class ColorC(ctypes.Structure):
    _fields_ = [("model", u32), ("primaries", u32), ("transfer", u32), ("whitepoint", u32), ("luminance", u32)]


class DescriptorC(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("row_stride", u64), ("texel_stride", u32), ("texel", Texel),
                ("color", ColorC)]


class FractalNoiseC(ctypes.Structure):
    _fields_ = [("num_octaves", u32), ("initial_amplitude", f32), ("amplitude_damping", f32), ("grid_scale", f32 * 2)]


class PaletteC(ctypes.Structure):
    _fields_ = [("width", i32), ("height", i32), ("width_base", i32), ("height_base", i32)]


class RectC(ctypes.Structure):
    _fields_ = [("x", u32), ("y", u32), ("max_x", u32), ("max_y", u32)]


class OpC(ctypes.Structure):
    _fields_ = [("high", u32), ("dst", u32), ("src", u32), ("target", u32), ("initializer", u32), ("shader", u32),
                ("num_operands", u32), ("operands", u32 * 2), ("knob", i32), ("params", vp), ("params_len", u32),
                ("selection", RectC), ("quad", f32 * 8), ("viewport", RectC)]


class BindingC(ctypes.Structure):
    _fields_ = [("slot", u32), ("is_output", u32), ("image", vp)]


class GraphInfoC(ctypes.Structure):
    _fields_ = [("state", u32), ("nodes", u32), ("captures", u32), ("pad", u32), ("replays", u64)]


class PoolBindingC(ctypes.Structure):
    _fields_ = [("slot", u32), ("is_output", u32), ("key", u64)]


class ShaderCommandC(ctypes.Structure):
    _fields_ = [("shader", vp), ("data", vp), ("data_len", u32), ("desc", DescriptorC)]


class ResourcesUsedC(ctypes.Structure):
    _fields_ = [("buffer_mem", u64), ("buffer_reused", u64), ("program_mem", u64), ("buffers_allocated", u32),
                ("buffers_reused", u32)]


class LearnOptionsC(ctypes.Structure):
    _fields_ = [("uniform_points", u32), ("face_points", u32), ("steps", u32), ("batch", u32), ("lr", f32),
                ("seed", u32), ("lattice", u32), ("max_delta_e", f32)]


class LearnReportC(ctypes.Structure):
    _fields_ = [("stage", u32), ("accepted", u32), ("steps", u32), ("samples", u32), ("initial_loss", f32),
                ("final_loss", f32), ("max_delta_e", ctypes.c_double), ("mean_delta_e", ctypes.c_double),
                ("max_delta_e_initial", ctypes.c_double), ("train_ms", f32)]


# name: (restype, argtypes)
_SIGNATURES = {
    # zosimos_hip.h
    "zh_context_create": (i32, [i32, vp, ctypes.POINTER(vp)]),
    "zh_context_destroy": (i32, [vp]),
    "zh_sync": (i32, [vp]),
    "zh_query": (i32, [vp, ctypes.POINTER(i32)]),
    "zh_context_stream": (vp, [vp]),
    "zh_image_alloc": (i32, [vp, ctypes.POINTER(DescriptorC), ctypes.POINTER(vp)]),
    "zh_image_wrap": (i32, [vp, ctypes.POINTER(DescriptorC), vp, sz, ctypes.POINTER(vp)]),
    "zh_image_free": (i32, [vp]),
    "zh_image_device_ptr": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(sz)]),
    "zh_image_upload": (i32, [vp, vp, u64]),
    "zh_image_download": (i32, [vp, vp, u64]),
    "zh_program_build": (i32, [vp, ctypes.POINTER(DescriptorC), u32, ctypes.POINTER(OpC), u32, u32,
                               ctypes.POINTER(vp)]),
    "zh_program_destroy": (i32, [vp]),
    "zh_program_describe": (i32, [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]),
    "zh_program_stats": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u64)]),
    "zh_program_set_knob": (i32, [vp, i32, vp, u32]),
    "zh_program_knob_data": (i32, [vp, i32, vp, u32, ctypes.POINTER(u32)]),
    "zh_program_launch_bytes": (i32, [vp, ctypes.POINTER(u64), u32, ctypes.POINTER(u32)]),
    "zh_program_launch": (i32, [vp, ctypes.POINTER(BindingC), u32]),
    "zh_program_graph_info": (i32, [vp, ctypes.POINTER(GraphInfoC)]),
    "zh_program_launch_static": (i32, [vp, ctypes.POINTER(DescriptorC), u32, ctypes.POINTER(OpC), u32, u32,
                                       ctypes.POINTER(BindingC), u32]),
    "zh_program_time": (i32, [vp, ctypes.POINTER(BindingC), u32, u32, ctypes.POINTER(f32), u32,
                              ctypes.POINTER(u32)]),
    "zh_program_autotune": (i32, [vp, ctypes.POINTER(BindingC), u32, u32]),
    "zh_program_learn": (i32, [vp, ctypes.POINTER(LearnOptionsC), ctypes.POINTER(LearnReportC), u32,
                               ctypes.POINTER(u32)]),
    "zh_program_use_surrogates": (i32, [vp, i32]),
    "zh_program_surrogate_params": (i32, [vp, u32, ctypes.POINTER(f32), u32]),
    "zh_mlp_gradient": (i32, [vp, ctypes.POINTER(f32), u32, ctypes.POINTER(f32), ctypes.POINTER(f32),
                              ctypes.POINTER(f32), ctypes.POINTER(f32)]),
    "zh_mlp_evaluate": (i32, [vp, ctypes.POINTER(f32), u32, ctypes.POINTER(f32), ctypes.POINTER(f32)]),
    "zh_srgb_tables": (i32, [vp, sz]),
    "zh_pool_create": (i32, [ctypes.POINTER(vp)]),
    "zh_pool_destroy": (i32, [vp]),
    "zh_pool_insert_host": (i32, [vp, ctypes.POINTER(DescriptorC), vp, u64, ctypes.POINTER(u64)]),
    "zh_pool_declare": (i32, [vp, ctypes.POINTER(DescriptorC), ctypes.POINTER(u64)]),
    "zh_pool_insert_device": (i32, [vp, vp, ctypes.POINTER(u64)]),
    "zh_pool_image_info": (i32, [vp, u64, ctypes.POINTER(DescriptorC), ctypes.POINTER(i32)]),
    "zh_pool_host_bytes": (i32, [vp, u64, ctypes.POINTER(vp), ctypes.POINTER(u64)]),
    "zh_pool_device_image": (i32, [vp, u64, ctypes.POINTER(vp)]),
    "zh_pool_set_no_read": (i32, [vp, u64, i32]),
    "zh_pool_upload": (i32, [vp, u64, vp]),
    "zh_pool_download": (i32, [vp, u64]),
    "zh_pool_remove": (i32, [vp, u64]),
    "zh_pool_clear_cache": (i32, [vp]),
    "zh_pool_cached_buffers": (i32, [vp, ctypes.POINTER(u32)]),
    "zh_program_launch_pool": (i32, [vp, vp, ctypes.POINTER(PoolBindingC), u32, ctypes.POINTER(vp)]),
    "zh_execution_retire_output": (i32, [vp, u32, i32, ctypes.POINTER(u64)]),
    "zh_execution_retire_input": (i32, [vp, u32, ctypes.POINTER(u64)]),
    "zh_execution_finish": (i32, [vp]),
    "zh_program_start_pool": (i32, [vp, vp, ctypes.POINTER(PoolBindingC), u32, ctypes.POINTER(vp)]),
    "zh_execution_step": (i32, [vp, u32, ctypes.POINTER(u32)]),
    "zh_execution_is_running": (i32, [vp, ctypes.POINTER(i32)]),
    "zh_execution_resources": (i32, [vp, ctypes.POINTER(ResourcesUsedC)]),
    "zh_execution_prune": (i32, [vp]),
    "zh_execution_discard": (i32, [vp]),
    "zh_program_launch_once": (i32, [vp, ctypes.POINTER(BindingC), u32]),
    "zh_program_dot": (i32, [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]),
    "zh_shader_create": (i32, [vp, sz, ctypes.c_char_p, ctypes.POINTER(vp)]),
    "zh_shader_retain": (i32, [vp]),
    "zh_shader_release": (i32, [vp]),
    "zh_last_error": (ctypes.c_char_p, []),
    "zh_version": (ctypes.c_char_p, []),
    # zosimos_command.h
    "zs_command_create": (i32, [ctypes.POINTER(vp)]),
    "zs_command_destroy": (i32, [vp]),
    "zs_command_input": (i32, [vp, ctypes.POINTER(DescriptorC), ctypes.POINTER(u32)]),
    "zs_command_output": (i32, [vp, u32, ctypes.POINTER(u32), ctypes.POINTER(DescriptorC)]),
    "zs_command_render": (i32, [vp, u32, ctypes.POINTER(u32), ctypes.POINTER(DescriptorC)]),
    "zs_command_crop": (i32, [vp, u32, RectC, ctypes.POINTER(u32)]),
    "zs_command_color_convert": (i32, [vp, u32, ctypes.POINTER(ColorC), ctypes.POINTER(Texel), ctypes.POINTER(u32)]),
    "zs_command_inscribe": (i32, [vp, u32, RectC, u32, ctypes.POINTER(u32)]),
    "zs_command_transmute": (i32, [vp, u32, ctypes.POINTER(DescriptorC), ctypes.POINTER(u32)]),
    "zs_command_affine": (i32, [vp, u32, ctypes.POINTER(f32), u32, u32, ctypes.POINTER(u32)]),
    "zs_command_blend": (i32, [vp, u32, RectC, u32, u32, ctypes.POINTER(u32)]),
    "zs_command_solid_rgba": (i32, [vp, ctypes.POINTER(DescriptorC), ctypes.POINTER(f32), ctypes.POINTER(u32)]),
    "zs_command_bilinear": (i32, [vp, ctypes.POINTER(DescriptorC), ctypes.POINTER(f32), ctypes.POINTER(u32)]),
    "zs_command_oklab_adjust": (i32, [vp, u32, f32, f32, ctypes.POINTER(u32)]),
    "zs_command_chromatic_adaptation": (i32, [vp, u32, u32, u32, ctypes.POINTER(u32)]),
    "zs_command_derivative": (i32, [vp, u32, u32, u32, ctypes.POINTER(u32)]),
    "zs_normal2d_with_diagonal": (i32, [ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "zs_normal2d_with_direction": (i32, [ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "zs_command_distribution_normal2d": (i32, [vp, ctypes.POINTER(DescriptorC), ctypes.POINTER(ctypes.c_float),
                                               ctypes.POINTER(u32)]),
    "zs_fractal_noise_with_octaves": (i32, [u32, ctypes.POINTER(FractalNoiseC)]),
    "zs_fractal_noise_set_damping": (i32, [ctypes.POINTER(FractalNoiseC), ctypes.c_float]),
    "zs_command_distribution_fractal_noise": (i32, [vp, ctypes.POINTER(DescriptorC), ctypes.POINTER(FractalNoiseC),
                                                    ctypes.POINTER(u32)]),
    "zs_command_palette": (i32, [vp, u32, ctypes.POINTER(PaletteC), u32, ctypes.POINTER(u32)]),
    "zs_command_extract": (i32, [vp, u32, i32, ctypes.POINTER(u32)]),
    "zs_command_resize": (i32, [vp, u32, u32, u32, ctypes.POINTER(u32)]),
    "zs_command_invoke": (i32, [vp, vp, ctypes.POINTER(u32), u32, ctypes.POINTER(u32), u32]),
    "zs_command_inject": (i32, [vp, u32, i32, u32, ctypes.POINTER(u32)]),
    "zs_command_buffer_init": (i32, [vp, vp, ctypes.c_uint64, ctypes.POINTER(u32)]),
    "zs_command_buffer_zero": (i32, [vp, ctypes.c_uint64, ctypes.POINTER(u32)]),
    "zs_command_buffer_size": (i32, [vp, u32, ctypes.POINTER(ctypes.c_uint64)]),
    "zs_command_from_buffer": (i32, [vp, u32, ctypes.POINTER(DescriptorC), ctypes.POINTER(u32)]),
    "zs_command_construct_dynamic": (i32, [vp, ctypes.POINTER(ShaderCommandC), ctypes.POINTER(u32)]),
    "zs_command_unary_dynamic": (i32, [vp, u32, ctypes.POINTER(ShaderCommandC), ctypes.POINTER(u32)]),
    "zs_command_binary_dynamic": (i32, [vp, u32, u32, ctypes.POINTER(ShaderCommandC), ctypes.POINTER(u32)]),
    "zs_command_knob": (i32, [vp, u32, ctypes.POINTER(i32)]),
    "zs_command_describe": (i32, [vp, u32, ctypes.POINTER(DescriptorC)]),
    "zs_command_generic": (i32, [vp, ctypes.POINTER(u32)]),
    "zs_command_input_generic": (i32, [vp, u32, ctypes.POINTER(u32)]),
    "zs_command_register_descriptor": (i32, [vp, u32, ctypes.POINTER(u32)]),
    "zs_command_derive_descriptor": (i32, [vp, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "zs_command_signature": (i32, [vp, ctypes.POINTER(vp)]),
    "zs_signature_destroy": (i32, [vp]),
    "zs_signature_counts": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "zs_command_function": (i32, [vp, vp, ctypes.POINTER(u32)]),
    "zs_command_invoke_function": (i32, [vp, u32, ctypes.POINTER(u32), u32, ctypes.POINTER(u32), u32,
                                         ctypes.POINTER(u32), u32]),
    "zs_command_compile": (i32, [vp, ctypes.POINTER(vp)]),
    "zs_link": (i32, [vp, ctypes.POINTER(DescriptorC), u32, ctypes.POINTER(vp), u32, ctypes.POINTER(u32),
                      ctypes.POINTER(u32), ctypes.POINTER(vp)]),
    "zs_linked_knob": (i32, [vp, u32, u32, ctypes.POINTER(i32)]),
    "zs_linked_destroy": (i32, [vp]),
    "zs_linked_high": (i32, [vp, ctypes.POINTER(ctypes.POINTER(DescriptorC)), ctypes.POINTER(u32),
                             ctypes.POINTER(ctypes.POINTER(OpC)), ctypes.POINTER(u32)]),
    "zs_linked_input_slot": (i32, [vp, u32, ctypes.POINTER(u32)]),
    "zs_linked_output_slot": (i32, [vp, u32, ctypes.POINTER(u32)]),
    "zs_linked_lower": (i32, [vp, vp, u32, ctypes.POINTER(vp)]),
}

_lib = None


def lib():
    """The loaded backend library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m zosimos_amd.build` "
                "(the HIP backend has no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def exported_symbols():
    return list(_SIGNATURES)

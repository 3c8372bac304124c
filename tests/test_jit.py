"""Planning of run-time kernels (no GPU): a flat pointwise chain without a
registered specialised kernel is marked for one compiled at run time, unless
ZH_PLAN_NO_JIT asks for the interpreter; registered chains keep their
kernels, and chains with operations the specialised bodies lack stay on the
interpreter."""
import zosimos_amd as z
from zosimos_amd import workloads as W


def describe(cmd, flags=0):
    return z.Linker.from_included().compile(cmd).lower_to(None, flags).describe()


def test_unregistered_chain_is_compiled_at_run_time():
    wl = W.c2_reordered(64, 32)
    assert "[kernel jit#" in describe(wl.commands)
    assert "[kernel interp-inline]" in describe(wl.commands, z.PLAN_NO_JIT)
    # the same signature twice: one kernel id
    a = describe(W.c2_reordered(64, 32).commands).split("jit#")[1].split("]")[0]
    b = describe(W.c2_reordered(128, 16).commands).split("jit#")[1].split("]")[0]
    assert a == b


def test_registered_and_unsupported_chains():
    assert "[kernel c2_over" in describe(W.c2(64, 32).commands)
    # SrLab2's non-linearity has no specialised body: the interpreter runs it
    d = z.Descriptor.rgba8_srgb(32, 32)
    cmd = z.CommandBuffer()
    i = cmd.input(d)
    s = cmd.color_convert(i, z.Color(model=z.ColorModel.SrLab2),
                          z.Texel(z.SampleBits.Float16x4, z.SampleParts.LabA))
    cmd.output(s)
    text = describe(cmd)
    assert "jit#" not in text


def test_unregistered_warp_form_is_compiled_at_run_time():
    d = z.Descriptor(200, 120, W.RGBA16F, W.P3_LINEAR)
    cmd = z.CommandBuffer()
    a = cmd.input(d)
    canvas = cmd.solid_rgba(z.Descriptor(300, 180, W.RGBA16F, W.P3_LINEAR), (0, 0, 0, 0))
    r = cmd.affine(canvas, z.Affine.new(z.AffineSample.BiLinear).rotate(0.3), a)
    cmd.output(cmd.color_convert(r, z.Color.SRGB, W.RGBA8))
    assert "post[unpremul mat3 store#1] [kernel jitw#" in describe(cmd)
    assert "jitw#" not in describe(cmd, z.PLAN_NO_JIT)
    # axis-aligned maps pick the separable form
    cmd2 = z.CommandBuffer()
    a2 = cmd2.input(d)
    c2 = cmd2.solid_rgba(z.Descriptor(300, 180, W.RGBA16F, W.P3_LINEAR), (0, 0, 0, 0))
    r2 = cmd2.affine(c2, z.Affine.new(z.AffineSample.BiLinear).scale(1.5, 1.5), a2)
    cmd2.output(cmd2.color_convert(r2, z.Color.SRGB, W.RGBA8))
    assert "_sep]" in describe(cmd2).split("jitw#")[1]
    assert "_sep_lds]" in describe(cmd2, z.PLAN_STAGED).split("jitw#")[1]

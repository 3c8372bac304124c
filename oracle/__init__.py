"""CPU oracle for the compositing hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import anything from this package, and there only as the checker (or
the timed CPU baseline), never as the thing measured or shipped.  The product
library (``zosimos_amd``) never imports, links or executes this code.

Contents
--------
``color``      FP64 numpy restatement of every per-pixel operation on the path
               (transfer functions, primaries matrices, Oklab, Lab/LCh
               adjustment, premultiply, Porter-Duff over, affine bilinear
               resampling, 8-bit / 16-bit / f16 / f32 texel codecs, Delta-E OK).
``stage``      ctypes binding of ``stage_frag.c``: an FP32, bit-level C
               restatement of the reference's ``lib/std/src/stage.frag``
               staging codec (demux/mux, parts permutation, transfers,
               truncating quantisation).
``blockhash``  a restatement of the blockhash-256 perceptual hash that the
               reference's tests compare against (``tests/util.rs:22``).
``mlp``        FP64 forward/backward of the 3-64-64-3 SiLU colour surrogate.

Each function cites the reference file:line it restates.  Where no reference
implementation exists (blend, bilinear affine, f16, chroma/hue) the docstring
says so and names the published method instead.
"""

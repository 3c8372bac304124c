"""Copy the reference's own test data for the perceptual-hash checks into
this folder: its two input images (``lib/zosimos/tests/input``) and the
blessed ``blockhash256`` lines of the tests on the hot path
(``lib/zosimos/tests/reference/<name>.crc.png``, text files despite the
suffix).  Data only; nothing of the reference is run.

    python tests/golden/make_reference_fixtures.py
"""
import json
import os
import shutil

REF = "/root/reference/lib/zosimos/tests"
HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ["composed", "affine", "convert_bt709", "swapped", "transmute", "oklab", "adapted", "solid",
         "bilinear-knob-0", "bilinear-knob-1", "bilinear-knob-2", "bilinear-knob-3", "bilinear-knob-4",
         "bilinear_from_buffer", "from_buffer", "from_buffer-with-knob", "distribution_normal2d",
         "distribution_normal1d", "distribution_u8", "distribution_fractal2d", "srlab2", "derived_Scharr3", "derived_Scharr3To4Bit",
         "derived_Scharr3To8Bit", "derived_Prewitt", "derived_Sobel", "palette", "generic", "mandelbrot", "crt",
         "flat_field"]


def main():
    for f in ("background.png", "foreground.png"):
        shutil.copyfile(os.path.join(REF, "input", f), os.path.join(HERE, f))
    hashes = {}
    for n in NAMES:
        with open(os.path.join(REF, "reference", n + ".crc.png")) as fh:
            hashes[n] = [l.strip() for l in fh.read().splitlines() if l.strip()]
    with open(os.path.join(HERE, "reference_hashes.json"), "w") as fh:
        json.dump(hashes, fh, indent=1)


if __name__ == "__main__":
    main()

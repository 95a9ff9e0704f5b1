"""Writes tests/golden/srgb8_to_float.txt: the reference's sRGB8 -> float
texel table (ryg-srgb.h:33-56) as 256 hex words, computed by the
reference's own header compiled here (oracle/Makefile target `ref`).

    python tests/golden/gen_srgb8_to_float.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

if __name__ == "__main__":
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_srgb_check"), "decode"],
                         capture_output=True, text=True, check=True).stdout
    words = out.split()
    assert len(words) == 256
    with open(os.path.join(HERE, "srgb8_to_float.txt"), "w") as f:
        f.write("\n".join(words) + "\n")

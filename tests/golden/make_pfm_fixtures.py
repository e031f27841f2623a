"""Writes the PFM fixtures byte by byte from the netpbm PFM definition
(three text header lines -- "Pf" gray or "PF" RGB, "<width> <height>",
scale with the sign giving the byte order: negative little-endian,
positive big-endian -- then float32 rows from the bottom row up), and the
arrays a reader must return (top row first, RGB extended to RGBA with
alpha 1, as cvt's PFMLoader does).

Run from the repository root: python tests/golden/make_pfm_fixtures.py
"""
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    cases = {}
    # 3x2 gray, big-endian, a comment line and a blank line in the header
    top, bottom = [1.0, 2.0, 3.0], [4.5, -5.25, 6e-3]
    body = b"".join(struct.pack(">f", v) for v in bottom + top)
    with open(os.path.join(HERE, "gray_be.pfm"), "wb") as f:
        f.write(b"Pf\n# written by hand\n\n3 2\n1.0\n" + body)
    cases["gray_be.pfm"] = {"shape": [2, 3], "data": [top, bottom]}
    # 2x2 RGB, little-endian, scale -2.5, tab-separated size
    rows_top = [[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]]
    rows_bot = [[-1.0, 0.0, 1.0], [7.0, 8.0, 9.0]]
    body = b"".join(struct.pack("<fff", *px) for px in rows_bot + rows_top)
    with open(os.path.join(HERE, "rgb_le.pfm"), "wb") as f:
        f.write(b"PF\n2\t2\n-2.5\n" + body)
    rgba = [[px + [1.0] for px in rows_top], [px + [1.0] for px in rows_bot]]
    cases["rgb_le.pfm"] = {"shape": [2, 2, 4], "data": rgba}
    with open(os.path.join(HERE, "pfm_expected.json"), "w") as f:
        json.dump(cases, f, indent=1)


if __name__ == "__main__":
    main()

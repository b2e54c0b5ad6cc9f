"""Writes tests/golden/kat.json from the reference's hand-calculated golden files.

Run in the container that holds the reference (reads /root/reference/tests/golden/*.hex
as data).  The pixel inputs are the values the reference tests encode to produce
those files (reference tests/cpu_codec_test.cpp:91-107, tests/fp32_codec_test.cpp:80-83).
The output is a data fixture: inputs, expected archive bytes and expected mode.
"""
import json
import os

GOLDEN = "/root/reference/tests/golden"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")

zero_width = [0x0102] * 8 + [0x0304] * 8
half_range = [0, 32768, 0, 32768, 0, 32768, 0, 32768, 0] + [5] * 9
CASES = [
    ("uint16_raw", 1, 1, 1, [0x1234], 0),
    ("uint16_constant", 1, 3, 2, [0xBEEF] * 6, 1),
    ("uint16_wrap", 1, 8, 2,
     [65535, 0, 1, 2, 3, 4, 5, 6, 0, 65535, 65534, 65533, 65532, 65531, 65530, 65529], 2),
    ("uint16_zero_width", 1, 8, 2, zero_width, 2),
    ("uint16_raw_tie", 1, 6, 1, [0, 1, 2, 3, 4, 5], 0),
    ("uint16_half_range", 1, 9, 2, half_range, 2),
    ("float32_raw", 2, 1, 1, [0x7FC01234], 0),
    ("float32_negative_zero_constant", 2, 3, 2, [0x80000000] * 6, 1),
]


def main():
    cases = []
    for name, ptype, w, h, pixels, mode in CASES:
        with open(os.path.join(GOLDEN, name + ".hex")) as f:
            data = bytes(int(tok, 16) for tok in f.read().split())
        cases.append({"name": name, "type": ptype, "width": w, "height": h,
                      "pixels": pixels, "mode": mode, "file": data.hex()})
    with open(OUT, "w") as f:
        json.dump({"source": "reference tests/golden/*.hex", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()

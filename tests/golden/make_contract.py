"""Writes tests/golden/calibration_contract.json: the reference's 42 calibration
contract rows (sample word, offset word, gain word, expected result word) as data.

Run in the container that holds the reference; it reads the numeric rows of
reference tests/numerics/support.h:413-456 and keeps only the four words and
the note of each row.  Each expected word is also recomputed here as the
binary32 result RN(RN(x - offset) * gain) (NumPy float32 arithmetic) and must
agree (any NaN matches an expected NaN).
"""
import json
import os
import re

import numpy as np

SOURCE = "/root/reference/tests/numerics/support.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "calibration_contract.json")
ROW = re.compile(r"\{(0x[0-9a-fA-F]+),\s*(0x[0-9a-fA-F]+),\s*(0x[0-9a-fA-F]+),\s*(0x[0-9a-fA-F]+),\s*\"([^\"]*)\"\}")


def binary32(x, o, g):
    f = lambda w: np.array([w], np.uint32).view(np.float32)  # noqa: E731
    with np.errstate(all="ignore"):
        return int(((f(x) - f(o)) * f(g)).view(np.uint32)[0])


def main():
    text = open(SOURCE).read()
    block = text[text.index("contract_rows[]"):]
    block = block[:block.index("};")]
    rows = [(int(a, 16), int(b, 16), int(c, 16), int(d, 16), note) for a, b, c, d, note in ROW.findall(block)]
    assert len(rows) == 42, len(rows)
    for x, o, g, want, note in rows:
        got = binary32(x, o, g)
        nan = lambda w: (w & 0x7F800000) == 0x7F800000 and (w & 0x7FFFFF) != 0  # noqa: E731
        assert got == want or (nan(got) and nan(want)), (note, hex(got), hex(want))
    json.dump({"source": "reference tests/numerics/support.h:413-456",
               "rows": [{"x": x, "offset": o, "gain": g, "expected": e, "note": n} for x, o, g, e, n in rows]},
              open(OUT, "w"), indent=1)
    print(f"wrote {len(rows)} rows to {OUT}")


if __name__ == "__main__":
    main()
